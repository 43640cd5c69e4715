"""Minimal ctypes access to the HIP runtime that libmppi_amd.so links (device buffers, streams, events).

This is plumbing for the Python host layer, the tests and the benchmark: device memory that the C-ABI kernel-level
entry points read and write. It binds /opt/rocm/lib/libamdhip64.so, the same runtime the kernel library uses.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HIP_PATHS = ["/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"]
_hip = None

H2D = 1
D2H = 2
D2D = 3


def _load():
    global _hip
    if _hip is not None:
        return _hip
    err = None
    for p in _HIP_PATHS:
        try:
            _hip = ct.CDLL(p, mode=os.RTLD_NOW | os.RTLD_GLOBAL)
            break
        except OSError as e:  # try the next name
            err = e
    if _hip is None:
        raise ImportError(f"HIP runtime not found: {err}")
    h = _hip
    h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    h.hipFree.argtypes = [ct.c_void_p]
    h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    h.hipMemcpyAsync.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int, ct.c_void_p]
    h.hipMemset.argtypes = [ct.c_void_p, ct.c_int, ct.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    h.hipGetDeviceCount.argtypes = [ct.POINTER(ct.c_int)]
    h.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
    h.hipStreamDestroy.argtypes = [ct.c_void_p]
    h.hipStreamSynchronize.argtypes = [ct.c_void_p]
    h.hipEventCreate.argtypes = [ct.POINTER(ct.c_void_p)]
    h.hipEventDestroy.argtypes = [ct.c_void_p]
    h.hipEventRecord.argtypes = [ct.c_void_p, ct.c_void_p]
    h.hipEventSynchronize.argtypes = [ct.c_void_p]
    h.hipEventElapsedTime.argtypes = [ct.POINTER(ct.c_float), ct.c_void_p, ct.c_void_p]
    h.hipGetErrorString.argtypes = [ct.c_int]
    h.hipGetErrorString.restype = ct.c_char_p
    h.hipDeviceGetAttribute.argtypes = [ct.POINTER(ct.c_int), ct.c_int, ct.c_int]
    return h


def check(st: int, what: str):
    if st != 0:
        raise RuntimeError(f"{what}: {_load().hipGetErrorString(st).decode()}")


def device_count() -> int:
    n = ct.c_int(0)
    st = _load().hipGetDeviceCount(ct.byref(n))
    return n.value if st == 0 else 0


def synchronize():
    check(_load().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """An owning device allocation with numpy copy-in/copy-out."""

    def __init__(self, nbytes: int):
        self.nbytes = int(max(nbytes, 4))
        p = ct.c_void_p()
        check(_load().hipMalloc(ct.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            check(_load().hipMemcpy(b.ptr, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")
        return b

    @classmethod
    def zeros(cls, shape, dtype=np.float32) -> "DeviceBuffer":
        return cls.from_numpy(np.zeros(shape, dtype=dtype))

    def copy_from(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(_load().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")

    def to_numpy(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            check(_load().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _load().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown
            pass


class Stream:
    def __init__(self):
        s = ct.c_void_p()
        check(_load().hipStreamCreate(ct.byref(s)), "hipStreamCreate")
        self.ptr = s.value

    def synchronize(self):
        check(_load().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.ptr:
                _load().hipStreamDestroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ct.c_void_p()
        check(_load().hipEventCreate(ct.byref(e)), "hipEventCreate")
        self.ptr = e.value

    def record(self, stream=None):
        check(_load().hipEventRecord(self.ptr, stream.ptr if isinstance(stream, Stream) else stream), "hipEventRecord")

    def synchronize(self):
        check(_load().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ct.c_float()
        check(_load().hipEventElapsedTime(ct.byref(ms), self.ptr, end.ptr), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                _load().hipEventDestroy(self.ptr)
        except Exception:
            pass
