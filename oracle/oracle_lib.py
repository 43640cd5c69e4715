"""ctypes access to the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE, the checker of the HIP engine.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use this module. The oracle is built from
oracle/oracle.cpp by oracle/Makefile; this module builds it only when no build exists.
"""
from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np

from mppi_generic_vendor_amd import capi

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "liboracle.so"

P = ct.c_void_p
_i = ct.c_int
_f = ct.c_float
_d = ct.c_double

_SIGS = {
    "oracle_philox4x32_10": ([P, P, P], None),
    "oracle_u32_to_unit": ([ct.c_uint32], _f),
    "oracle_normals": ([ct.c_uint64, ct.c_uint64, ct.c_int64, P], None),
    "oracle_pure_noise_start": ([_i, _f], _i),
    "oracle_gaussian_transform": ([ct.POINTER(capi.SamplerParams), _i, _i, _i, P, P, _i, _i, P], None),
    "oracle_colored_scales": ([ct.POINTER(capi.SamplerParams), _i, _i, P, P, ct.POINTER(_i), ct.POINTER(_i)], None),
    "oracle_colored_samples": ([ct.POINTER(capi.SamplerParams), _i, _i, _i, P, P, _i, _i, P, P], None),
    "oracle_step_f32": ([ct.POINTER(capi.DynamicsParams), P, P, _f, P, P, P], None),
    "oracle_state_deriv_f64": ([ct.POINTER(capi.DynamicsParams), P, P, P], None),
    "oracle_enforce_constraints_f32": ([ct.POINTER(capi.DynamicsParams), P], None),
    "oracle_fnn_forward_f32": ([_i, P, P, P, P], None),
    "oracle_fnn_forward_f64": ([_i, P, P, P, P], None),
    "oracle_fnn_forward_bf16_f64": ([_i, P, P, P, P], None),
    "oracle_bf16_round": ([_f], _f),
    "oracle_proposal_num_params": ([_i, _i, _i], _i),
    "oracle_gelu": ([_d], _d),
    "oracle_dgelu": ([_d], _d),
    "oracle_proposal_input": ([_i, _i, _i, P, P, P, P, P], None),
    "oracle_proposal_forward": ([_i, _i, _i, P, _i, P, _i, P], None),
    "oracle_proposal_loss_grad": ([_i, _i, _i, P, _i, P, P, _i, P], _d),
    "oracle_adam": ([P, P, P, P, _i, _f, _f, _f, _f, _i], None),
    "oracle_is_cost": ([P, P, P, P, P, _i, _i, _i, _d, P], None),
    "oracle_map_query_f32": ([ct.POINTER(capi.CostParams), P, _f, _f], _f),
    "oracle_running_cost_f32": ([ct.POINTER(capi.CostParams), P, P, _i], _f),
    "oracle_colored_weights_f32": ([P, _i, _f, _f, _f, P, P, P, P, P], _f),
    "oracle_enforce_leash": ([ct.POINTER(capi.DynamicsParams), P, P, P, P], None),
    "oracle_terminal_cost_f32": ([ct.POINTER(capi.CostParams), P], _f),
    "oracle_lr_cost_f32": ([ct.POINTER(capi.SamplerParams), _i, P, P, _i, _f, _f, _i], _f),
    "oracle_rollout_f32": ([ct.POINTER(capi.DynamicsParams), ct.POINTER(capi.CostParams), P,
                            ct.POINTER(capi.SamplerParams), _i, _i, _f, _f, _f, P, P, P, P, _i], None),
    "oracle_rollout_f64": ([ct.POINTER(capi.DynamicsParams), ct.POINTER(capi.CostParams), P,
                            ct.POINTER(capi.SamplerParams), _i, _i, _d, _d, _d, P, P, P, P, _i], None),
    "oracle_rollout_two_system_f32": ([ct.POINTER(capi.DynamicsParams), ct.POINTER(capi.CostParams), P, P, _i, _f,
                                       P, _i, _i, _f, _f, _f, P, P, P, P, P, _i], None),
    "oracle_candidate_eval_f32": ([ct.POINTER(capi.DynamicsParams), ct.POINTER(capi.CostParams), P,
                                   ct.POINTER(capi.SamplerParams), _i, _i, _i, _i, _f, _f, _f, P, P, P, P, P], None),
    "oracle_rollout_trace_f32": ([ct.POINTER(capi.DynamicsParams), ct.POINTER(capi.CostParams), P,
                                  ct.POINTER(capi.SamplerParams), _i, _i, _f, _f, _f, P, P, P, P, P, P, P], None),
    "oracle_rollout_trace_f64": ([ct.POINTER(capi.DynamicsParams), ct.POINTER(capi.CostParams), P,
                                  ct.POINTER(capi.SamplerParams), _i, _i, _d, _d, _d, P, P, P, P, P, P, P], None),
    "oracle_best_index": ([P, _i], _i),
    "oracle_weights": ([P, _i, _d, P, ct.POINTER(_f), ct.POINTER(_i), ct.POINTER(_d)], _d),
    "oracle_weights_f32": ([P, _i, _f, P], None),
    "oracle_weighted_update": ([P, _d, P, _i, _i, _i, P], None),
    "oracle_free_energy": ([P, _i, _d, _d, ct.POINTER(_d), ct.POINTER(_d), ct.POINTER(_d)], None),
    "oracle_smooth": ([P, _i, _i], None),
    "oracle_slide": ([P, _i, _i, _i, P, P], None),
    "oracle_state_trajectory": ([ct.POINTER(capi.DynamicsParams), P, P, _i, _f, P], None),
    "oracle_state_trajectory_f64": ([ct.POINTER(capi.DynamicsParams), P, P, _i, _d, P], None),
}

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
    lib = ct.CDLL(str(LIB))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def ptr(a: np.ndarray):
    return a.ctypes.data if a is not None else None


# ------------------------------------------------------------------------------------------- convenience wrappers
def normals(seed: int, offset: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    load().oracle_normals(seed, offset, n, ptr(out))
    return out


def gaussian_transform(sp, K, T, C, mean, eps, stride, iteration, std_dev_t=None) -> np.ndarray:
    out = np.empty((K, T, C), dtype=np.float32)
    _a0 = np.ascontiguousarray(mean, np.float32)
    _a1 = np.ascontiguousarray(eps, np.float32)
    sp = host_sampler(sp, std_dev_t)
    load().oracle_gaussian_transform(ct.byref(sp), K, T, C, ptr(_a0),
                                     ptr(_a1), stride, iteration, ptr(out))
    return out


def colored_scales(sp, T, C):
    scales = np.zeros((2 * T + 2, C), dtype=np.float32)
    sigma = np.zeros(C, dtype=np.float32)
    n, F = _i(), _i()
    load().oracle_colored_scales(ct.byref(sp), T, C, ptr(scales), ptr(sigma), ct.byref(n), ct.byref(F))
    return scales[: F.value], sigma, n.value, F.value


def colored_samples(sp, K, T, C, mean, spectra, stride, iteration, std_dev_t=None):
    eps = np.empty((K, T, C), dtype=np.float32)
    out = np.empty((K, T, C), dtype=np.float32)
    _a0 = np.ascontiguousarray(mean, np.float32)
    _a1 = np.ascontiguousarray(spectra, np.float32)
    sp = host_sampler(sp, std_dev_t)
    load().oracle_colored_samples(ct.byref(sp), K, T, C, ptr(_a0),
                                  ptr(_a1), stride, iteration, ptr(eps),
                                  ptr(out))
    return eps, out


def rollout_f32(prob, samples: np.ndarray, mean: np.ndarray, nthreads: int = 1, x0=None):
    """Returns (costs[K], clamped samples) for a workloads.Problem; `samples` is not modified."""
    K, T = samples.shape[0], samples.shape[1]
    s = np.ascontiguousarray(samples, dtype=np.float32).copy()
    costs = np.empty(K, dtype=np.float32)
    dyn = _host_dyn(prob)
    m = prob.cost_map
    x0 = prob.x0 if x0 is None else x0
    _a0 = np.ascontiguousarray(x0, np.float32)
    _a1 = np.ascontiguousarray(mean, np.float32)
    load().oracle_rollout_f32(ct.byref(dyn), ct.byref(_host_cost(prob)), ptr(m), ct.byref(_host_smp(prob)), K, T, prob.dt,
                              prob.lambda_, prob.alpha, ptr(_a0),
                              ptr(_a1), ptr(s), ptr(costs), nthreads)
    return costs, s


def rollout_two_system(prob, samples: np.ndarray, means: np.ndarray, x0s: np.ndarray, nominal_index: int = -1,
                       vft: float = 0.0, gains: np.ndarray | None = None, std_dev_1=None, std_dev_t_1=None,
                       nthreads: int = 8, parts: bool = False):
    """Two-system rollout (Tube: nominal_index < 0; RMPPI: 0 or 1) of samples [2][K][T][C] around means [2][T][C]
    from x0s [2][S]: returns (costs [2][K], applied controls [2][K][T][C]) and with `parts` also RMPPI's
    (cost_r, track_r, state_n, lr_n) [K][4]. Distribution 1's sampler is prob's with std_dev = std_dev_1 and
    std_dev_t = std_dev_t_1 when given."""
    K, T, C = samples.shape[1], samples.shape[2], samples.shape[3]
    s = np.ascontiguousarray(samples, dtype=np.float32).copy()
    costs = np.empty((2, K), dtype=np.float32)
    pr = np.zeros((K, 4), np.float32)
    smp = (capi.SamplerParams * 2)()
    s0 = _host_smp(prob)
    std_t1 = getattr(prob, "std_dev_t", None) if std_dev_t_1 is None else std_dev_t_1
    s1 = host_sampler(prob.sampler, std_t1)
    if std_dev_1 is not None:
        for c in range(C):
            s1.std_dev[c] = float(std_dev_1[c])
    smp[0], smp[1] = s0, s1
    g = None if gains is None else np.ascontiguousarray(gains, np.float32)
    _a0 = np.ascontiguousarray(x0s, np.float32)
    _a1 = np.ascontiguousarray(means, np.float32)
    load().oracle_rollout_two_system_f32(ct.byref(_host_dyn(prob)), ct.byref(_host_cost(prob)), ptr(prob.cost_map),
                                         smp, int(nominal_index), float(vft), ptr(g), K, T, prob.dt, prob.lambda_,
                                         prob.alpha, ptr(_a0), ptr(_a1), ptr(s), ptr(costs), ptr(pr), nthreads)
    return (costs, s, pr) if parts else (costs, s)


def candidate_eval(prob, samples: np.ndarray, mean: np.ndarray, states: np.ndarray, strides, M: int):
    """RMPPI candidate evaluation over the first M samples of samples [K][T][C]: costs [N][M]."""
    K, T = samples.shape[0], samples.shape[1]
    st = np.ascontiguousarray(states, np.float32)
    N = st.shape[0]
    sr = np.ascontiguousarray(strides, np.int32)
    s = np.ascontiguousarray(samples, np.float32)
    m = np.ascontiguousarray(mean, np.float32)
    out = np.empty((N, M), np.float32)
    load().oracle_candidate_eval_f32(ct.byref(_host_dyn(prob)), ct.byref(_host_cost(prob)), ptr(prob.cost_map),
                                     ct.byref(_host_smp(prob)), N, M, K, T, prob.dt, prob.lambda_, prob.alpha,
                                     ptr(st), ptr(sr), ptr(m), ptr(s), ptr(out))
    return out


def rollout_f64(prob, samples: np.ndarray, mean: np.ndarray, x0=None, nthreads: int = 8):
    K, T = samples.shape[0], samples.shape[1]
    s = np.ascontiguousarray(samples, dtype=np.float32).copy()
    costs = np.empty(K, dtype=np.float64)
    dyn = _host_dyn(prob)
    x0 = prob.x0 if x0 is None else x0
    _a0 = np.ascontiguousarray(x0, np.float32)
    _a1 = np.ascontiguousarray(mean, np.float32)
    load().oracle_rollout_f64(ct.byref(dyn), ct.byref(_host_cost(prob)), ptr(prob.cost_map), ct.byref(_host_smp(prob)), K, T,
                              prob.dt, prob.lambda_, prob.alpha, ptr(_a0),
                              ptr(_a1), ptr(s), ptr(costs), nthreads)
    return costs, s


def rollout_trace(prob, samples: np.ndarray, mean: np.ndarray, x0=None, f64: bool = False):
    """(costs[K], outputs[K][T][O], step costs[K][T], terminal costs[K]) of the oracle loop, FP32 or FP64: outputs are
    the post-step y_{t+1}, step costs (running + likelihood-ratio) / T, terminal / T."""
    K, T = samples.shape[0], samples.shape[1]
    O = prob.dyn.output_dim
    dt = np.float64 if f64 else np.float32
    s = np.ascontiguousarray(samples, dtype=np.float32).copy()
    costs, outs = np.empty(K, dt), np.empty((K, T, O), dt)
    steps, term = np.empty((K, T), dt), np.empty(K, dt)
    dyn = _host_dyn(prob)
    x0 = prob.x0 if x0 is None else x0
    _a0 = np.ascontiguousarray(x0, np.float32)
    _a1 = np.ascontiguousarray(mean, np.float32)
    fn = load().oracle_rollout_trace_f64 if f64 else load().oracle_rollout_trace_f32
    fn(ct.byref(dyn), ct.byref(_host_cost(prob)), ptr(prob.cost_map), ct.byref(_host_smp(prob)), K, T, prob.dt, prob.lambda_,
       prob.alpha, ptr(_a0), ptr(_a1), ptr(s), ptr(costs), ptr(outs), ptr(steps), ptr(term))
    return costs, outs, steps, term


_keep = []


def host_sampler(sp, std_dev_t=None):
    """A copy of the sampler block whose std_dev_t points to host memory holding `std_dev_t` [T][C] (None: unset)."""
    s = capi.SamplerParams.from_buffer_copy(sp)
    if std_dev_t is None:
        s.std_dev_t = None
    else:
        s._hold = np.ascontiguousarray(std_dev_t, dtype=np.float32)  # lives as long as the copy
        s.std_dev_t = s._hold.ctypes.data
    return s


def host_cost(cost, goal_traj=None):
    """A copy of the cost block whose goal_traj points to host memory holding `goal_traj` [H][O] (None: unset)."""
    c = capi.CostParams.from_buffer_copy(cost)
    if goal_traj is None:
        c.goal_traj = None
    else:
        c._hold = np.ascontiguousarray(goal_traj, dtype=np.float32)
        c.goal_traj = c._hold.ctypes.data
        c.goal_traj_len = c._hold.shape[0]
    return c


def _host_smp(prob):
    return host_sampler(prob.sampler, getattr(prob, "std_dev_t", None))


def _host_cost(prob):
    return host_cost(prob.cost, getattr(prob, "goal_traj", None))


def _host_dyn(prob):
    dyn = capi.DynamicsParams.from_buffer_copy(prob.dyn)
    if prob.nn_theta is not None:
        th = np.ascontiguousarray(prob.nn_theta, dtype=np.float32)
        _keep.append(th)
        dyn.nn_theta = th.ctypes.data
    return dyn


def weights(costs: np.ndarray, lam: float):
    K = costs.shape[0]
    c = np.ascontiguousarray(costs, dtype=np.float32)
    w = np.empty(K, dtype=np.float64)
    b, i, s2 = _f(), _i(), _d()
    eta = load().oracle_weights(ptr(c), K, lam, ptr(w), ct.byref(b), ct.byref(i), ct.byref(s2))
    return w, eta, b.value, i.value, s2.value


def weighted_update(w: np.ndarray, eta: float, samples: np.ndarray) -> np.ndarray:
    K, T, C = samples.shape
    out = np.empty((T, C), dtype=np.float64)
    _a0 = np.ascontiguousarray(w, np.float64)
    _a1 = np.ascontiguousarray(samples, np.float32)
    load().oracle_weighted_update(ptr(_a0), eta,
                                  ptr(_a1), K, T, C, ptr(out))
    return out


def smooth(u: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(u, dtype=np.float32).copy()
    T, C = v.shape
    load().oracle_smooth(ptr(v), T, C)
    return v


def state_trajectory(prob, x0, u) -> np.ndarray:
    T, C = u.shape
    out = np.empty((T, prob.S), dtype=np.float32)
    dyn = _host_dyn(prob)
    _a0 = np.ascontiguousarray(x0, np.float32)
    _a1 = np.ascontiguousarray(u, np.float32)
    load().oracle_state_trajectory(ct.byref(dyn), ptr(_a0),
                                   ptr(_a1), T, prob.dt, ptr(out))
    return out


def state_trajectory_f64(prob, x0, u) -> np.ndarray:
    T, C = u.shape
    out = np.empty((T, prob.S), dtype=np.float64)
    dyn = _host_dyn(prob)
    _a0 = np.ascontiguousarray(x0, np.float32)
    _a1 = np.ascontiguousarray(u, np.float32)
    load().oracle_state_trajectory_f64(ct.byref(dyn), ptr(_a0), ptr(_a1), T, float(np.float32(prob.dt)), ptr(out))
    return out


def colored_weights(costs: np.ndarray, lam: float, gamma: float = 0.0, r: float = 0.0):
    """The colored controller's host path: (FP32 weights, float normalizer, baseline, (fe mean, var, modified))."""
    K = costs.shape[0]
    c = np.ascontiguousarray(costs, dtype=np.float32)
    w = np.empty(K, dtype=np.float32)
    b, m, v, mod = _f(), _f(), _f(), _f()
    eta = load().oracle_colored_weights_f32(ptr(c), K, lam, gamma, r, ptr(w), ct.byref(b), ct.byref(m), ct.byref(v),
                                            ct.byref(mod))
    return w, eta, b.value, (m.value, v.value, mod.value)


def enforce_leash(dyn, x_true, x_nominal, leash):
    out = np.zeros(capi.MAX_STATE, np.float32)
    a = [np.ascontiguousarray(np.resize(np.asarray(v, np.float32), capi.MAX_STATE)) for v in (x_true, x_nominal, leash)]
    load().oracle_enforce_leash(ct.byref(dyn), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(out))
    return out[: dyn.state_dim]


def free_energy(w: np.ndarray, baseline: float, lam: float):
    """(mean, variance, modified variance) of core/mppi_common.cu:1101-1117 in FP64 from the weights."""
    m, v, mod = _d(), _d(), _d()
    _a0 = np.ascontiguousarray(w, np.float64)
    load().oracle_free_energy(ptr(_a0), _a0.size, baseline, lam, ct.byref(m), ct.byref(v), ct.byref(mod))
    return m.value, v.value, mod.value


def enforce(prob, u: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(u, dtype=np.float32).copy()
    dyn = _host_dyn(prob)
    for row in v.reshape(-1, v.shape[-1]):
        load().oracle_enforce_constraints_f32(ct.byref(dyn), row.ctypes.data)
    return v


# ------------------------------------------------------------------------------------------- proposal network
def proposal_forward(S, C, T, theta, X, bf16=False):
    X = np.ascontiguousarray(X, np.float32)
    B = X.shape[0]
    Y = np.empty((B, 2 * T * C), np.float64)
    _a0 = np.ascontiguousarray(theta, np.float32)
    load().oracle_proposal_forward(S, C, T, ptr(_a0), B, ptr(X), int(bf16),
                                   ptr(Y))
    return Y


def proposal_loss_grad(S, C, T, theta, X, R, bf16=False):
    X = np.ascontiguousarray(X, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    th = np.ascontiguousarray(theta, np.float32)
    g = np.empty(th.size, np.float64)
    loss = load().oracle_proposal_loss_grad(S, C, T, ptr(th), X.shape[0], ptr(X), ptr(R), int(bf16), ptr(g))
    return loss, g


def is_cost(u, mean, dmu, sig, sig_p, lam):
    u = np.ascontiguousarray(u, np.float32)
    K, T, C = u.shape
    out = np.empty(K, np.float64)
    _a0 = np.ascontiguousarray(mean, np.float32)
    _a1 = np.ascontiguousarray(dmu, np.float32)
    _a2 = np.ascontiguousarray(sig, np.float32)
    _a3 = np.ascontiguousarray(sig_p, np.float32)
    load().oracle_is_cost(ptr(u), ptr(_a0),
                          ptr(_a1), ptr(_a2),
                          ptr(_a3), K, T, C, lam, ptr(out))
    return out
