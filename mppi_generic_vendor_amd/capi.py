"""ctypes binding of the C-ABI declared in include/mppi_amd.h.

The structures below mirror the header field for field; ``check_layout()`` compares their sizes against the
library's own view. Loading fails loudly when libmppi_amd.so has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libmppi_amd.so"

MPPI_OK = 0
MPPI_E_INVALID = -1
MPPI_E_HIP = -2
MPPI_E_UNSUPPORTED = -3
MPPI_E_NOMEM = -4
MPPI_E_SIZE = -5

MAX_STATE = 8
MAX_CONTROL = 4
MAX_OUTPUT = 8
MAX_TIMESTEPS = 256
MAX_NN_LAYERS = 4

MODEL_CARTPOLE = 1
MODEL_DOUBLE_INTEGRATOR = 2
MODEL_RACER_DUBINS = 3
MODEL_RACER_DUBINS_NN = 4

COST_CARTPOLE_QUADRATIC = 1
COST_QUADRATIC_MAP = 2

SAMPLER_GAUSSIAN = 1
SAMPLER_COLORED = 2

FFT_LEN_2T = 0
FFT_LEN_T = 1

CONTROLLER_VANILLA = 0
CONTROLLER_COLORED = 1
TWO_SYSTEM_TUBE = 0
TWO_SYSTEM_RMPPI = 1

NN_FP32 = 0
NN_BF16 = 1

_f = ct.c_float
_i = ct.c_int32


class DynamicsParams(ct.Structure):
    _fields_ = [
        ("model", _i), ("state_dim", _i), ("control_dim", _i), ("output_dim", _i),
        ("control_min", _f * MAX_CONTROL), ("control_max", _f * MAX_CONTROL),
        ("control_deadband", _f * MAX_CONTROL), ("zero_control", _f * MAX_CONTROL),
        ("cart_mass", _f), ("pole_mass", _f), ("pole_length", _f), ("gravity", _f),
        ("c_t", _f), ("c_b", _f), ("c_v", _f), ("c_0", _f), ("steering_constant", _f),
        ("steer_command_angle_scale", _f), ("steer_angle_scale", _f), ("max_steer_angle", _f),
        ("max_steer_rate", _f), ("brake_delay_constant", _f), ("max_brake_rate_neg", _f),
        ("max_brake_rate_pos", _f), ("wheel_base", _f), ("gear_sign", _i),
        ("nn_num_layers", _i), ("nn_layers", _i * MAX_NN_LAYERS), ("nn_output_states", _i * MAX_STATE),
        ("nn_precision", _i), ("nn_num_params", _i), ("nn_theta", ct.c_void_p),
    ]


class CostParams(ct.Structure):
    _fields_ = [
        ("cost", _i), ("output_dim", _i),
        ("coeff", _f * MAX_OUTPUT), ("goal", _f * MAX_OUTPUT), ("terminal_cost_coeff", _f),
        ("map", ct.c_void_p), ("map_width", _i), ("map_height", _i), ("map_origin_x", _f), ("map_origin_y", _f),
        ("map_resolution", _f), ("map_coeff", _f), ("map_x_index", _i), ("map_y_index", _i),
        ("speed_coeff", _f), ("speed_target", _f), ("speed_index", _i),
        ("goal_traj_len", _i), ("goal_traj", ct.c_void_p), ("goal_time", _i),
        ("map_rotated", _i), ("map_rotation", _f * 4),
    ]


class SamplerParams(ct.Structure):
    _fields_ = [
        ("kind", _i), ("control_dim", _i),
        ("std_dev", _f * MAX_CONTROL), ("control_cost_coeff", _f * MAX_CONTROL),
        ("pure_noise_trajectories_percentage", _f), ("std_dev_decay", _f),
        ("exponents", _f * MAX_CONTROL), ("offset_decay_rate", _f), ("fmin", _f),
        ("fft_len_mode", _i), ("_pad", _i), ("seed", ct.c_uint64), ("std_dev_t", ct.c_void_p),
    ]


class TwoSystemParams(ct.Structure):
    _fields_ = [("mode", _i), ("nominal_index", _i), ("value_function_threshold", _f), ("_pad", _i),
                ("feedback_gains", ct.c_void_p)]


class WeightStats(ct.Structure):
    _fields_ = [
        ("baseline", _f), ("argmin", _i), ("normalizer", ct.c_double), ("sum_w2", ct.c_double),
        ("lambda_", _f), ("num_rollouts", _i), ("weighted_cost", ct.c_double),
    ]


class FreeEnergy(ct.Structure):
    _fields_ = [
        ("free_energy_mean", _f), ("free_energy_variance", _f), ("free_energy_modified_variance", _f),
        ("normalizer_percent", _f), ("increase", _f), ("previous_baseline", _f),
    ]


class ControllerDesc(ct.Structure):
    _fields_ = [
        ("dynamics", DynamicsParams), ("cost", CostParams), ("sampler", SamplerParams),
        ("num_rollouts", _i), ("num_timesteps", _i), ("dt", _f), ("lambda_", _f), ("alpha", _f),
        ("num_iters", _i), ("use_graph", _i), ("keep_samples", _i),
        ("slide_control_scale", _f * MAX_CONTROL),
        ("init_control", ct.c_void_p), ("cost_map_host", ct.c_void_p), ("nn_theta_host", ct.c_void_p),
        ("std_dev_t_host", ct.c_void_p), ("goal_traj_host", ct.c_void_p),
        ("controller_kind", _i), ("tsallis_gamma", _f), ("tsallis_r", _f), ("leash_active", _i),
        ("state_leash_dist", _f * MAX_STATE),
    ]


class TubeDesc(ct.Structure):
    _fields_ = [("base", ControllerDesc), ("mode", _i), ("nominal_threshold", _f),
                ("value_function_threshold", _f), ("num_candidates", _i), ("num_eval_samples", _i), ("_pad", _i)]


class InstanceDesc(ct.Structure):
    _fields_ = [
        ("dynamics", DynamicsParams), ("cost", CostParams), ("sampler", SamplerParams),
        ("num_rollouts", _i), ("num_timesteps", _i), ("dt", _f), ("lambda_", _f), ("alpha", _f),
        ("slide_control_scale", _f * MAX_CONTROL), ("init_control", ct.c_void_p),
        ("std_dev_t_host", ct.c_void_p), ("goal_traj_host", ct.c_void_p),
    ]


class BatchConfig(ct.Structure):
    _fields_ = [("threads", _i), ("samples_per_lane", _i), ("chunk", _i), ("blocks_per_cu", _i),
                ("iteration_us", _f)]


P = ct.c_void_p
_SIGS = {
    "mppi_last_error": ([], ct.c_char_p),
    "mppi_status_string": ([ct.c_int], ct.c_char_p),
    "mppi_version": ([], ct.c_char_p),
    "mppi_dynamics_defaults": ([_i, ct.POINTER(DynamicsParams)], ct.c_int),
    "mppi_cost_defaults": ([_i, ct.POINTER(DynamicsParams), ct.POINTER(CostParams)], ct.c_int),
    "mppi_sampler_defaults": ([_i, _i, ct.POINTER(SamplerParams)], ct.c_int),
    "mppi_nn_num_params": ([_i, ct.POINTER(_i), ct.POINTER(_i)], ct.c_int),
    "mppi_pure_noise_start": ([_i, _f, ct.POINTER(_i)], ct.c_int),
    "mppi_host_step": ([ct.POINTER(DynamicsParams), P, P, _f, P, P, P], ct.c_int),
    "mppi_host_enforce_constraints": ([ct.POINTER(DynamicsParams), P], ct.c_int),
    "mppi_host_running_cost": ([ct.POINTER(CostParams), P, ct.POINTER(_f)], ct.c_int),
    "mppi_host_running_cost_at": ([ct.POINTER(CostParams), P, _i, ct.POINTER(_f)], ct.c_int),
    "mppi_host_enforce_leash": ([ct.POINTER(DynamicsParams), P, P, P, P], ct.c_int),
    "mppi_compute_weights_tsallis": ([P, _i, _f, _f, _f, P, P], ct.c_int),
    "mppi_npz_read": ([ct.c_char_p, ct.c_char_p, P, ct.c_int64, P, P, P], ct.c_int),
    "mppi_fnn_load_npz": ([ct.c_char_p, ct.c_char_p, _i, P, P, P, _i, P], ct.c_int),
    "mppi_track_map_load_npz": ([ct.c_char_p, ct.c_char_p, P, ct.c_int64, P, P, P, P, P], ct.c_int),
    "mppi_controller_update_cost_map": ([P, P, _i, _i, _i], ct.c_int),
    "mppi_controller_update_cost_map_frame": ([P, _f, _f, _f, P], ct.c_int),
    "mppi_generate_normals": ([ct.c_uint64, ct.c_uint64, P, ct.c_int64, P], ct.c_int),
    "mppi_sample_gaussian": ([ct.POINTER(SamplerParams), _i, _i, P, P, _i, _i, ct.c_uint64, P, P], ct.c_int),
    "mppi_sample_colored": ([ct.POINTER(SamplerParams), _i, _i, P, P, _i, _i, ct.c_uint64, P, P], ct.c_int),
    "mppi_rollout": ([ct.POINTER(DynamicsParams), ct.POINTER(CostParams), ct.POINTER(SamplerParams), _i, _i, _f, _f,
                      _f, P, P, P, P, P], ct.c_int),
    "mppi_rollout_two_system": ([ct.POINTER(DynamicsParams), ct.POINTER(CostParams), ct.POINTER(SamplerParams),
                                 ct.POINTER(SamplerParams), ct.POINTER(TwoSystemParams), _i, _i, _f, _f, _f, P, P, P,
                                 P, P], ct.c_int),
    "mppi_compute_weights": ([P, _i, _f, P, P], ct.c_int),
    "mppi_weighted_update": ([P, P, P, _i, _i, _i, P, P], ct.c_int),
    "mppi_weighted_update_noalloc": ([P, P, P, _i, _i, _i, P, P], ct.c_int),
    "mppi_controller_create": ([ct.POINTER(ControllerDesc), ct.POINTER(P)], ct.c_int),
    "mppi_controller_destroy": ([P], ct.c_int),
    "mppi_controller_create_plugin": ([ct.POINTER(ControllerDesc), P, ct.POINTER(P)], ct.c_int),
    "mppi_controller_compute_control": ([P, P, _i], ct.c_int),
    "mppi_controller_step_device": ([P, _i], ct.c_int),
    "mppi_controller_get_control_seq": ([P, P], ct.c_int),
    "mppi_controller_get_raw_mean": ([P, P], ct.c_int),
    "mppi_controller_get_state_seq": ([P, P], ct.c_int),
    "mppi_controller_get_weight_stats": ([P, ct.POINTER(WeightStats)], ct.c_int),
    "mppi_controller_get_free_energy": ([P, ct.POINTER(FreeEnergy)], ct.c_int),
    "mppi_controller_set_control_seq": ([P, P], ct.c_int),
    "mppi_controller_slide_control_seq": ([P, _i], ct.c_int),
    "mppi_controller_device_buffers": ([P, ct.POINTER(P), ct.POINTER(P), ct.POINTER(P)], ct.c_int),
    "mppi_controller_set_seed": ([P, ct.c_uint64], ct.c_int),
    "mppi_controller_set_params": ([P, ct.POINTER(ControllerDesc)], ct.c_int),
    "mppi_controller_set_sampled_trajectories": ([P, _i, _i], ct.c_int),
    "mppi_controller_get_sampled_trajectories": ([P, ct.POINTER(_i), P, P, P, P, P, P], ct.c_int),
    "mppi_controller_stream": ([P, ct.POINTER(P)], ct.c_int),
    "mppi_batch_create": ([ct.POINTER(InstanceDesc), _i, P, _i, _i, ct.POINTER(P)], ct.c_int),
    "mppi_batch_destroy": ([P], ct.c_int),
    "mppi_batch_autotune": ([P, _i, ct.POINTER(BatchConfig)], ct.c_int),
    "mppi_batch_set_config": ([P, ct.POINTER(BatchConfig)], ct.c_int),
    "mppi_batch_get_config": ([P, ct.POINTER(BatchConfig)], ct.c_int),
    "mppi_batch_compute_control": ([P, P, _i], ct.c_int),
    "mppi_batch_set_graph_replay": ([P, _i], ct.c_int),
    "mppi_batch_step_device": ([P, _i], ct.c_int),
    "mppi_batch_get_control_seq": ([P, _i, P], ct.c_int),
    "mppi_batch_get_raw_mean": ([P, _i, P], ct.c_int),
    "mppi_batch_get_state_seq": ([P, _i, P], ct.c_int),
    "mppi_batch_get_weight_stats": ([P, _i, ct.POINTER(WeightStats)], ct.c_int),
    "mppi_batch_get_free_energy": ([P, _i, ct.POINTER(FreeEnergy)], ct.c_int),
    "mppi_batch_set_control_seq": ([P, _i, P], ct.c_int),
    "mppi_batch_slide_control_seq": ([P, _i], ct.c_int),
    "mppi_batch_get_samples": ([P, _i, P, P], ct.c_int),
    "mppi_batch_stream": ([P, ct.POINTER(P)], ct.c_int),
    "mppi_proposal_num_params": ([_i, _i, _i, ct.POINTER(_i)], ct.c_int),
    "mppi_proposal_create": ([_i, _i, _i, P, ct.POINTER(P)], ct.c_int),
    "mppi_proposal_destroy": ([P], ct.c_int),
    "mppi_proposal_get_params": ([P, P], ct.c_int),
    "mppi_proposal_set_params": ([P, P], ct.c_int),
    "mppi_proposal_set_input_normalization": ([P, P, P], ct.c_int),
    "mppi_proposal_forward": ([P, P, _i, P, P], ct.c_int),
    "mppi_proposal_infer": ([P, P, P, P, P, P, _i, P], ct.c_int),
    "mppi_proposal_loss_grad": ([P, P, P, _i, P, ct.POINTER(ct.c_double), P], ct.c_int),
    "mppi_proposal_train_step": ([P, P, P, _i, _f, ct.POINTER(ct.c_double), P], ct.c_int),
    "mppi_controller_set_proposal": ([P, P, _i, _f], ct.c_int),
    "mppi_controller_get_proposal_stats": ([P, ct.POINTER(_i), ct.POINTER(_i), ct.POINTER(WeightStats),
                                            ct.POINTER(_f)], ct.c_int),
    "mppi_controller_get_last_call_ms": ([P, ct.POINTER(_f)], ct.c_int),
    "mppi_controller_time_kernels": ([P, _i, ct.POINTER(_f), ct.POINTER(_f), ct.POINTER(_i)], ct.c_int),
    "mppi_tube_create": ([ct.POINTER(TubeDesc), ct.POINTER(P)], ct.c_int),
    "mppi_tube_destroy": ([P], ct.c_int),
    "mppi_tube_compute_control": ([P, P, _i], ct.c_int),
    "mppi_tube_get_control_seq": ([P, _i, P], ct.c_int),
    "mppi_tube_set_control_seq": ([P, _i, P], ct.c_int),
    "mppi_tube_get_state_seq": ([P, _i, P], ct.c_int),
    "mppi_tube_get_weight_stats": ([P, _i, P], ct.c_int),
    "mppi_tube_get_free_energy": ([P, _i, P], ct.c_int),
    "mppi_tube_get_nominal_state_used": ([P, ct.POINTER(_i)], ct.c_int),
    "mppi_tube_slide_control_seq": ([P, _i], ct.c_int),
    "mppi_tube_set_nominal_state": ([P, P], ct.c_int),
    "mppi_tube_get_nominal_state": ([P, P], ct.c_int),
    "mppi_tube_set_feedback_gains": ([P, P], ct.c_int),
    "mppi_tube_get_feedback_control": ([P, P, P, _i, P], ct.c_int),
    "mppi_tube_set_thresholds": ([P, _f, _f], ct.c_int),
    "mppi_tube_get_samples": ([P, P, P], ct.c_int),
    "mppi_tube_time_kernels": ([P, _i, ct.POINTER(_f), ct.POINTER(_f), ct.POINTER(_i)], ct.c_int),
    "mppi_tube_update_importance_sampling_control": ([P, P, _i], ct.c_int),
    "mppi_tube_get_candidates": ([P, P, P, P, P, ct.POINTER(_i)], ct.c_int),
    "mppi_batch_time_kernels": ([P, _i, ct.POINTER(_f), ct.POINTER(_f)], ct.c_int),
}

_lib = None


class MPPIError(RuntimeError):
    def __init__(self, status: int, func: str, msg: str):
        super().__init__(f"{func} failed with status {status}: {msg}")
        self.status = status


def load() -> ct.CDLL:
    """Load libmppi_amd.so (built by mppi_generic_vendor_amd.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m mppi_generic_vendor_amd.build`")
    lib = ct.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS.keys())


def call(name: str, *args) -> int:
    """Call a status-returning entry point; raise MPPIError on a negative status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != MPPI_OK:
        raise MPPIError(st, name, lib.mppi_last_error().decode())
    return st


def dynamics_defaults(model: int) -> DynamicsParams:
    p = DynamicsParams()
    call("mppi_dynamics_defaults", model, ct.byref(p))
    return p


def cost_defaults(cost: int, dyn: DynamicsParams) -> CostParams:
    c = CostParams()
    call("mppi_cost_defaults", cost, ct.byref(dyn), ct.byref(c))
    return c


def sampler_defaults(kind: int, control_dim: int) -> SamplerParams:
    s = SamplerParams()
    call("mppi_sampler_defaults", kind, control_dim, ct.byref(s))
    return s


def pure_noise_start(num_rollouts: int, pct: float) -> int:
    out = _i()
    call("mppi_pure_noise_start", num_rollouts, pct, ct.byref(out))
    return out.value


# ---------------------------------------------------------------------------------------------- weight / map I/O
def npz_read(path: str, key: str):
    """Array `key` of an .npz archive as float32 in C order (mppi_npz_read)."""
    import numpy as np

    nd, cnt = _i(), ct.c_int64()
    shape = (ct.c_int64 * 4)()
    call("mppi_npz_read", str(path).encode(), key.encode(), None, 0, ct.byref(nd), shape, ct.byref(cnt))
    out = np.empty(cnt.value, np.float32)
    call("mppi_npz_read", str(path).encode(), key.encode(), out.ctypes.data, cnt.value, None, None, None)
    return out.reshape([shape[i] for i in range(nd.value)])


def fnn_load_npz(path: str, prefix: str = ""):
    """(layer sizes, FP32 parameter vector) of FNNHelper::loadParams(prefix, npz) (mppi_fnn_load_npz)."""
    import numpy as np

    nl, npar = _i(), _i()
    call("mppi_fnn_load_npz", str(path).encode(), prefix.encode(), 0, ct.byref(nl), None, None, 0, ct.byref(npar))
    layers = np.zeros(nl.value, np.int32)
    theta = np.empty(npar.value, np.float32)
    call("mppi_fnn_load_npz", str(path).encode(), prefix.encode(), nl.value, None, layers.ctypes.data,
         theta.ctypes.data, npar.value, None)
    return [int(v) for v in layers], theta


def track_map_load_npz(path: str, channel: str = "channel0"):
    """(values [height][width], (origin x, origin y), resolution) of a track map (mppi_track_map_load_npz)."""
    import numpy as np

    w, h = _i(), _i()
    ox, oy, res = _f(), _f(), _f()
    call("mppi_track_map_load_npz", str(path).encode(), channel.encode(), None, 0, ct.byref(w), ct.byref(h),
         ct.byref(ox), ct.byref(oy), ct.byref(res))
    vals = np.empty((h.value, w.value), np.float32)
    call("mppi_track_map_load_npz", str(path).encode(), channel.encode(), vals.ctypes.data, vals.size, None, None,
         None, None, None)
    return vals, (ox.value, oy.value), res.value
