"""Helpers of the GPU tests: run the C-ABI kernel-level entry points on device buffers."""
from __future__ import annotations

import ctypes as ct
import sys

import numpy as np

from mppi_generic_vendor_amd import capi
from mppi_generic_vendor_amd.hip import DeviceBuffer, synchronize


def banner(msg: str):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def device_problem(prob, with_sampler=False):
    """Copies of the problem's parameter blocks with device pointers for the cost map, the MLP parameters and the
    goal trajectory (and, with_sampler, a sampler copy pointing at the per-step standard deviations)."""
    dyn = capi.DynamicsParams.from_buffer_copy(prob.dyn)
    cost = capi.CostParams.from_buffer_copy(prob.cost)
    smp = capi.SamplerParams.from_buffer_copy(prob.sampler)
    keep = []
    if prob.cost_map is not None:
        mb = DeviceBuffer.from_numpy(prob.cost_map.astype(np.float32))
        keep.append(mb)
        cost.map = mb.ptr
    if prob.nn_theta is not None:
        tb = DeviceBuffer.from_numpy(prob.nn_theta.astype(np.float32))
        keep.append(tb)
        dyn.nn_theta = tb.ptr
    if prob.goal_traj is not None:
        gb = DeviceBuffer.from_numpy(np.ascontiguousarray(prob.goal_traj, np.float32))
        keep.append(gb)
        cost.goal_traj = gb.ptr
        cost.goal_traj_len = prob.goal_traj.shape[0]
    if prob.std_dev_t is not None:
        sb = DeviceBuffer.from_numpy(np.ascontiguousarray(prob.std_dev_t, np.float32))
        keep.append(sb)
        smp.std_dev_t = sb.ptr
    if with_sampler:
        return dyn, cost, smp, keep
    return dyn, cost, keep


def gpu_normals(seed: int, offset: int, n: int) -> np.ndarray:
    b = DeviceBuffer(n * 4)
    capi.call("mppi_generate_normals", ct.c_uint64(seed), ct.c_uint64(offset), b.ptr, n, None)
    synchronize()
    return b.to_numpy(n)


def _with_std_dev_t(sp, std_dev_t):
    """A sampler copy whose std_dev_t points at a device copy of `std_dev_t` (None: the block as given)."""
    if std_dev_t is None:
        return sp, None
    buf = DeviceBuffer.from_numpy(np.ascontiguousarray(std_dev_t, np.float32))
    s = capi.SamplerParams.from_buffer_copy(sp)
    s.std_dev_t = buf.ptr
    return s, buf


def gpu_sample_gaussian(sp, K, T, C, mean, stride, iteration, rng_offset=0, eps=None, std_dev_t=None):
    sp, _sb = _with_std_dev_t(sp, std_dev_t)
    mb = DeviceBuffer.from_numpy(np.ascontiguousarray(mean, np.float32))
    out = DeviceBuffer(K * T * C * 4)
    eb = DeviceBuffer.from_numpy(np.ascontiguousarray(eps, np.float32)) if eps is not None else None
    capi.call("mppi_sample_gaussian", ct.byref(sp), K, T, mb.ptr, out.ptr, stride, iteration,
              ct.c_uint64(rng_offset), eb.ptr if eb else None, None)
    synchronize()
    return out.to_numpy((K, T, C))


def gpu_sample_colored(sp, K, T, C, mean, stride, iteration, rng_offset=0, spectra=None, std_dev_t=None):
    sp, _sb = _with_std_dev_t(sp, std_dev_t)
    mb = DeviceBuffer.from_numpy(np.ascontiguousarray(mean, np.float32))
    out = DeviceBuffer(K * T * C * 4)
    sb = DeviceBuffer.from_numpy(np.ascontiguousarray(spectra, np.float32)) if spectra is not None else None
    capi.call("mppi_sample_colored", ct.byref(sp), K, T, mb.ptr, out.ptr, stride, iteration,
              ct.c_uint64(rng_offset), sb.ptr if sb else None, None)
    synchronize()
    return out.to_numpy((K, T, C))


def gpu_rollout(prob, samples, mean, x0=None):
    K, T, C = samples.shape
    dyn, cost, smp, keep = device_problem(prob, with_sampler=True)
    x0 = prob.x0 if x0 is None else x0
    xb = DeviceBuffer.from_numpy(np.ascontiguousarray(x0, np.float32))
    mb = DeviceBuffer.from_numpy(np.ascontiguousarray(mean, np.float32))
    sb = DeviceBuffer.from_numpy(np.ascontiguousarray(samples, np.float32))
    cb = DeviceBuffer(K * 4)
    capi.call("mppi_rollout", ct.byref(dyn), ct.byref(cost), ct.byref(smp), K, T, prob.dt, prob.lambda_,
              prob.alpha, xb.ptr, mb.ptr, sb.ptr, cb.ptr, None)
    synchronize()
    return cb.to_numpy(K), sb.to_numpy((K, T, C))


def gpu_weights(costs, lam):
    K = costs.shape[0]
    cb = DeviceBuffer.from_numpy(np.ascontiguousarray(costs, np.float32))
    st = DeviceBuffer(ct.sizeof(capi.WeightStats))
    capi.call("mppi_compute_weights", cb.ptr, K, lam, st.ptr, None)
    synchronize()
    w = cb.to_numpy(K)
    raw = st.to_numpy(ct.sizeof(capi.WeightStats), np.uint8)
    return w, capi.WeightStats.from_buffer_copy(raw.tobytes())


def gpu_weighted_update(weights, stats, samples):
    K, T, C = samples.shape
    wb = DeviceBuffer.from_numpy(np.ascontiguousarray(weights, np.float32))
    stb = DeviceBuffer.from_numpy(np.frombuffer(bytes(stats), dtype=np.uint8).copy())
    sb = DeviceBuffer.from_numpy(np.ascontiguousarray(samples, np.float32))
    ob = DeviceBuffer(T * C * 4)
    capi.call("mppi_weighted_update", wb.ptr, stb.ptr, sb.ptr, K, T, C, ob.ptr, None)
    synchronize()
    return ob.to_numpy((T, C))


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
