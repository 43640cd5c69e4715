"""mppi_weighted_update_noalloc (the allocation-free variant of mppi_weighted_update) against mppi_weighted_update,
through the C-ABI: two calls in a row on the shared scratch array, the second with a smaller input, each equal bit
for bit to the allocating entry point (the same kernels over the same partial rows); sizes beyond the array are
rejected with MPPI_E_SIZE before any device work."""
import ctypes as ct

import numpy as np
import pytest

from mppi_generic_vendor_amd import capi

pytestmark = pytest.mark.gpu


def _update(name, w, stats, u):
    from mppi_generic_vendor_amd.hip import DeviceBuffer, synchronize

    K, T, C = u.shape
    wb = DeviceBuffer.from_numpy(np.ascontiguousarray(w, np.float32))
    stb = DeviceBuffer.from_numpy(np.frombuffer(bytes(stats), dtype=np.uint8).copy())
    sb = DeviceBuffer.from_numpy(np.ascontiguousarray(u, np.float32))
    ob = DeviceBuffer(T * C * 4)
    capi.call(name, wb.ptr, stb.ptr, sb.ptr, K, T, C, ob.ptr, None)
    synchronize()
    return ob.to_numpy((T, C))


def test_noalloc_update_equals_allocating_update(hip, engine):
    from gpu_util import banner, gpu_weights

    rng = np.random.default_rng(31)
    # (8192, 100, 2): 128 partial rows of 200 floats; then (700, 37, 3): 11 rows of 111 floats over the array that
    # still holds the first call's rows
    for K, T, C in ((8192, 100, 2), (700, 37, 3)):
        banner(f"weighted_update_noalloc K={K} T={T} C={C}")
        u = rng.normal(size=(K, T, C)).astype(np.float32)
        c = rng.uniform(0.0, 30.0, K).astype(np.float32)
        w, st = gpu_weights(c, 1.5)
        got = _update("mppi_weighted_update_noalloc", w, st, u)
        want = _update("mppi_weighted_update", w, st, u)
        print(f"[noalloc] K={K} T={T} C={C}: max |diff| {float(np.max(np.abs(got - want))):.3g}")
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, want)


def test_noalloc_update_rejects_sizes_beyond_the_scratch_array(hip, engine):
    lib = capi.load()
    # ceil(65536 / 64) * 256 * 4 * 4 bytes = 4 MiB > 2 MiB
    st = lib.mppi_weighted_update_noalloc(ct.c_void_p(16), ct.c_void_p(16), ct.c_void_p(16), 65536, 256, 4,
                                          ct.c_void_p(16), None)
    assert st == capi.MPPI_E_SIZE
    assert b"scratch" in lib.mppi_last_error()
