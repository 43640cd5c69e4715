"""A user-written device plugin (tests/plugin: a unicycle model and cost the engine does not ship, compiled against
include/mppi_amd/device_plugin.hpp) through the engine, against the oracle's rollout loop (oracle/rollout_loop.hpp)
over the same model's host formulas.

- kernel-level plugin rollout: clamped samples bit-exact (deadband and limits), costs within 1e-4 relative of the
  FP32 loop (the reference's rollout tolerance, tests/mppi_core/rollout_kernel_tests.cu:258), argmin equal; errors
  against the FP64 loop printed;
- the C-ABI plugin controller (mppi_controller_create_plugin): eager and graph-replayed calls against the oracle chain
  (weights, update), smoothing, the plugin's constraint on the returned sequence and its state propagation; colored
  samples drawn by the library's sampler around the plugin; the device loop;
- the C++ VanillaMPPIController<Unicycle, UnicycleCost, ...> in a closed loop around an obstacle to the goal.
"""
import ctypes as ct
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import banner, gpu_normals, gpu_sample_colored, gpu_sample_gaussian, rel_err
from mppi_generic_vendor_amd import capi

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent
PLUGIN_LIB = HERE / "plugin" / "_build" / "libunicycle_plugin.so"
_f = ct.c_float


class UnicycleSetup(ct.Structure):
    _fields_ = [("speed_scale", _f), ("turn_scale", _f),
                ("goal_x", _f), ("goal_y", _f), ("goal_heading", _f), ("q_pos", _f), ("q_heading", _f),
                ("obs_x", _f), ("obs_y", _f), ("obs_radius", _f), ("obs_weight", _f), ("obs_time_gain", _f),
                ("control_weight", _f), ("terminal_weight", _f),
                ("control_min", _f * 2), ("control_max", _f * 2), ("control_deadband", _f * 2),
                ("zero_control", _f * 2)]


def setup():
    s = UnicycleSetup()
    s.speed_scale, s.turn_scale = 1.0, 1.0
    s.goal_x, s.goal_y, s.goal_heading = 2.0, 1.5, 0.6
    s.q_pos, s.q_heading = 1.0, 0.05
    s.obs_x, s.obs_y, s.obs_radius, s.obs_weight, s.obs_time_gain = 1.0, 0.75, 0.3, 20.0, 0.01
    s.control_weight, s.terminal_weight = 0.05, 10.0
    s.control_min[0], s.control_max[0] = -0.5, 2.0
    s.control_min[1], s.control_max[1] = -1.5, 1.5
    s.control_deadband[0], s.zero_control[0] = 0.05, 0.0
    return s


def sampler():
    sp = capi.sampler_defaults(capi.SAMPLER_GAUSSIAN, 2)
    sp.std_dev[0], sp.std_dev[1] = 0.8, 0.7
    sp.control_cost_coeff[0], sp.control_cost_coeff[1] = 0.2, 0.1
    sp.seed = 2024
    return sp


def plugin_lib():
    capi.load()
    if not PLUGIN_LIB.exists():
        raise RuntimeError(f"{PLUGIN_LIB} is not built (python tests/plugin/build_plugin.py)")
    lib = ct.CDLL(str(PLUGIN_LIB))
    P, i, f, d = ct.c_void_p, ct.c_int, ct.c_float, ct.c_double
    S = ct.POINTER(UnicycleSetup)
    sigs = {
        "unicycle_rollout_gpu": ([S, P, i, i, f, f, f, P, P, P, P], ct.c_int),
        "unicycle_rollout_oracle_f32": ([S, P, i, i, f, f, f, P, P, P, P, i], None),
        "unicycle_rollout_oracle_f64": ([S, P, i, i, d, d, d, P, P, P, P, i], None),
        "unicycle_state_trajectory_f32": ([S, P, P, i, f, P], None),
        "unicycle_controller_create": ([P, S, P], ct.c_int),
        "unicycle_cpp_closed_loop": ([S, P, i, f, P, P], ct.c_int),
    }
    for name, (a, r) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = a
        fn.restype = r
    return lib


def oracle_rollout(lib, s, sp, samples, mean, x0, dt, lam, alpha, f64=False):
    K, T, C = samples.shape
    u = np.ascontiguousarray(samples, np.float32).copy()
    x0 = np.ascontiguousarray(x0, np.float32)
    mean = np.ascontiguousarray(mean, np.float32)
    if f64:
        c = np.empty(K, np.float64)
        lib.unicycle_rollout_oracle_f64(ct.byref(s), ct.byref(sp), K, T, dt, lam, alpha, x0.ctypes.data,
                                        mean.ctypes.data, u.ctypes.data, c.ctypes.data, 8)
    else:
        c = np.empty(K, np.float32)
        lib.unicycle_rollout_oracle_f32(ct.byref(s), ct.byref(sp), K, T, dt, lam, alpha, x0.ctypes.data,
                                        mean.ctypes.data, u.ctypes.data, c.ctypes.data, 8)
    return c, u


def constrain(s, u):
    """The plugin's control constraint (deadband + limits) through the oracle's restatement."""
    dyn = capi.DynamicsParams()
    dyn.control_dim = 2
    for i in range(2):
        dyn.control_min[i], dyn.control_max[i] = s.control_min[i], s.control_max[i]
        dyn.control_deadband[i], dyn.zero_control[i] = s.control_deadband[i], s.zero_control[i]
    v = np.ascontiguousarray(u, np.float32).copy()
    for row in v.reshape(-1, 2):
        ol.load().oracle_enforce_constraints_f32(ct.byref(dyn), row.ctypes.data)
    return v


@pytest.mark.parametrize("K,T,alpha", [(1000, 80, 0.0), (64, 17, 0.3), (4096, 120, 0.1)])
def test_plugin_rollout_matches_oracle(hip, engine, K, T, alpha):
    from mppi_generic_vendor_amd.hip import DeviceBuffer, synchronize

    banner(f"plugin rollout K={K} T={T} alpha={alpha}")
    lib = plugin_lib()
    s, sp = setup(), sampler()
    rng = np.random.default_rng(K + T)
    mean = (rng.normal(size=(T, 2)) * [0.4, 0.3] + [0.6, 0.0]).astype(np.float32)
    x0 = np.array([0.2, -0.3, 0.5], np.float32)
    samples = gpu_sample_gaussian(sp, K, T, 2, mean, 1, 0, rng_offset=3)
    assert np.mean(np.abs(samples[..., 0]) < 0.05) > 0.01  # the deadband branch is exercised
    xb, mb = DeviceBuffer.from_numpy(x0), DeviceBuffer.from_numpy(mean)
    sb, cb = DeviceBuffer.from_numpy(samples), DeviceBuffer(K * 4)
    st = lib.unicycle_rollout_gpu(ct.byref(s), ct.byref(sp), K, T, 0.05, 1.0, alpha, xb.ptr, mb.ptr, sb.ptr, cb.ptr)
    assert st == 0
    synchronize()
    c_gpu, u_gpu = cb.to_numpy(K), sb.to_numpy((K, T, 2))
    c32, u32 = oracle_rollout(lib, s, sp, samples, mean, x0, 0.05, 1.0, alpha)
    c64, _ = oracle_rollout(lib, s, sp, samples, mean, x0, 0.05, 1.0, alpha, f64=True)
    assert np.array_equal(u_gpu, u32)
    e = rel_err(c_gpu, c32)
    print(f"[plugin] rollout K={K} T={T}: gpu-vs-oracle32 {e:.2e}, gpu-vs-fp64 {rel_err(c_gpu, c64):.2e}, "
          f"oracle32-vs-fp64 {rel_err(c32, c64):.2e}")
    assert np.all(np.isfinite(c_gpu)) and e < 1e-4
    assert int(np.argmin(c_gpu)) == int(np.argmin(c32))


def _create(lib, sp, K, T, use_graph=True, keep=False):
    d = capi.ControllerDesc()
    d.sampler = sp
    d.num_rollouts, d.num_timesteps = K, T
    d.dt, d.lambda_, d.alpha = 0.05, 1.0, 0.0
    d.num_iters = 1
    d.use_graph = 1 if use_graph else 0
    d.keep_samples = 1 if keep else 0
    d.slide_control_scale[0] = d.slide_control_scale[1] = 1.0
    h = ct.c_void_p()
    s = setup()
    st = lib.unicycle_controller_create(ct.byref(d), ct.byref(s), ct.byref(h))
    assert st == 0, capi.load().mppi_last_error()
    return h


def _get(name, h, shape):
    out = np.empty(shape, np.float32)
    capi.call(name, h, out.ctypes.data)
    return out


def test_plugin_controller_matches_oracle(hip, engine):
    banner("plugin controller K=2048 T=80, eager + graph")
    lib = plugin_lib()
    s, sp = setup(), sampler()
    K, T, C = 2048, 80, 2
    h = _create(lib, sp, K, T)
    x0 = np.array([0.0, 0.0, 0.0], np.float32)
    for call in range(2):
        nominal = np.zeros((T, C), np.float32) if call == 0 else _get("mppi_controller_get_control_seq", h, (T, C))
        capi.call("mppi_controller_set_control_seq", h, nominal.ctypes.data)
        capi.call("mppi_controller_compute_control", h, x0.ctypes.data, 1)
        raw = _get("mppi_controller_get_raw_mean", h, (T, C))
        st = capi.WeightStats()
        capi.call("mppi_controller_get_weight_stats", h, ct.byref(st))
        eps = gpu_normals(sp.seed, call, K * T * C)
        samples = gpu_sample_gaussian(sp, K, T, C, nominal, 1, 0, eps=eps)
        costs, clamped = oracle_rollout(lib, s, sp, samples, nominal, x0, 0.05, 1.0, 0.0)
        w, eta, b, idx, _ = ol.weights(costs, 1.0)
        want = ol.weighted_update(w, eta, clamped)
        err = float(np.max(np.abs(raw - want))) / max(1e-3, float(np.max(np.abs(want))))
        print(f"[plugin] controller call {call}: mean rel err {err:.2e}, argmin {st.argmin} vs {idx}, "
              f"baseline {st.baseline:.6g} vs {b:.6g}")
        assert err < 1e-4 and st.argmin == idx
        assert st.baseline == pytest.approx(b, rel=1e-4)
        smoothed = ol.smooth(raw)
        np.testing.assert_allclose(_get("mppi_controller_get_control_seq", h, (T, C)), constrain(s, smoothed),
                                   rtol=1e-6, atol=1e-7)
        xs = _get("mppi_controller_get_state_seq", h, (T, 3))
        want_xs = np.empty((T, 3), np.float32)
        lib.unicycle_state_trajectory_f32(ct.byref(s), x0.ctypes.data, smoothed.ctypes.data, T, 0.05,
                                          want_xs.ctypes.data)
        np.testing.assert_allclose(xs, want_xs, rtol=1e-6, atol=1e-6)
    # the device loop replays the plugin's kernels from captured graphs
    capi.call("mppi_controller_step_device", h, 10)
    from mppi_generic_vendor_amd import hip as hipmod
    hipmod.synchronize()
    capi.call("mppi_controller_compute_control", h, x0.ctypes.data, 1)
    assert np.all(np.isfinite(_get("mppi_controller_get_raw_mean", h, (T, C))))
    capi.call("mppi_controller_destroy", h)


def test_plugin_controller_colored_samples(hip, engine):
    """Colored samples drawn by the library's sampler into the sample buffer, read and constrained by the plugin's
    kernel: the stored samples equal the colored sampler's constrained, and the mean matches the oracle chain."""
    from mppi_generic_vendor_amd.hip import D2H, _load

    banner("plugin controller, colored samples, keep_samples")
    lib = plugin_lib()
    s = setup()
    sp = capi.sampler_defaults(capi.SAMPLER_COLORED, 2)
    sp.std_dev[0], sp.std_dev[1] = 0.8, 0.7
    sp.exponents[0], sp.exponents[1] = 1.0, 2.0
    sp.seed = 77
    K, T, C = 1024, 60, 2
    h = _create(lib, sp, K, T, use_graph=False, keep=True)
    x0 = np.array([0.1, 0.0, -0.2], np.float32)
    nominal = np.random.default_rng(4).normal(scale=0.3, size=(T, C)).astype(np.float32)
    capi.call("mppi_controller_set_control_seq", h, nominal.ctypes.data)
    capi.call("mppi_controller_compute_control", h, x0.ctypes.data, 1)
    sptr, _, _ = ct.c_void_p(), ct.c_void_p(), ct.c_void_p()
    capi.call("mppi_controller_device_buffers", h, ct.byref(sptr), None, None)
    got = np.empty((K, T, C), np.float32)
    assert _load().hipMemcpy(got.ctypes.data, sptr, got.nbytes, D2H) == 0
    samples = gpu_sample_colored(sp, K, T, C, nominal, 1, 0, rng_offset=0)
    assert np.array_equal(got, constrain(s, samples))
    costs, clamped = oracle_rollout(lib, s, sp, samples, nominal, x0, 0.05, 1.0, 0.0)
    w, eta, *_ = ol.weights(costs, 1.0)
    want = ol.weighted_update(w, eta, clamped)
    raw = _get("mppi_controller_get_raw_mean", h, (T, C))
    err = float(np.max(np.abs(raw - want))) / max(1e-3, float(np.max(np.abs(want))))
    print(f"[plugin] colored controller: mean rel err {err:.2e}")
    assert err < 1e-4
    capi.call("mppi_controller_destroy", h)


def test_plugin_cpp_controller_closed_loop(hip, engine):
    """VanillaMPPIController over the plugin types (the C++ drop-in) drives the unicycle from the origin, 2.5 from
    the goal (2, 1.5), past the obstacle bump to within 0.5 of it (the same loop on the CPU oracle, lambda 0.05,
    ends 0.28 away: the heading and control terms balance the last stretch)."""
    banner("plugin C++ controller closed loop")
    lib = plugin_lib()
    s = setup()
    x0 = np.zeros(3, np.float32)
    xf = np.zeros(3, np.float32)
    base = ct.c_float()
    st = lib.unicycle_cpp_closed_loop(ct.byref(s), x0.ctypes.data, 150, 0.05, xf.ctypes.data, ct.byref(base))
    assert st == 0
    dist = float(np.hypot(xf[0] - 2.0, xf[1] - 1.5))
    print(f"[plugin] closed loop: final state {xf}, distance to goal {dist:.3f}, baseline {base.value:.4f}")
    assert dist < 0.5 and np.isfinite(base.value)
