"""Python host wrapper of the C-ABI controller, with the reference controller's method names
(controllers/controller.cuh:228-876, controllers/MPPI/mppi_controller.cuh:36-41)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import capi
from .workloads import Problem


def _host_arrays(d, problem: Problem) -> list:
    """Point the descriptor's host-copy fields at the problem's host arrays (cost map, network parameters, per-step
    standard deviations, goal trajectory); returns the arrays, which must stay alive across the C call."""
    keep = []

    def host(name, a, shape=None):
        arr = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None:
            arr = arr.reshape(shape)
        keep.append(arr)
        setattr(d, name, arr.ctypes.data)
        return arr

    if problem.cost_map is not None:
        host("cost_map_host", problem.cost_map)
    if problem.nn_theta is not None:
        host("nn_theta_host", problem.nn_theta)
    if problem.std_dev_t is not None:
        host("std_dev_t_host", problem.std_dev_t, (problem.num_timesteps, problem.C))
    if problem.goal_traj is not None:
        g = host("goal_traj_host", problem.goal_traj, (-1, problem.dyn.output_dim))
        d.cost.goal_traj_len = g.shape[0]
    return keep


class VanillaMPPIController:
    def __init__(self, problem: Problem, num_iters: int = 1, use_graph: bool = False,
                 init_control: np.ndarray | None = None, slide_control_scale=None, keep_samples: bool = False):
        lib = capi.load()
        self.problem = problem
        self.T = problem.num_timesteps
        self.K = problem.num_rollouts
        self.S = problem.S
        self.C = problem.C
        d = capi.ControllerDesc()
        d.dynamics = problem.dyn
        d.cost = problem.cost
        d.sampler = problem.sampler
        d.num_rollouts = self.K
        d.num_timesteps = self.T
        d.dt = problem.dt
        d.lambda_ = problem.lambda_
        d.alpha = problem.alpha
        d.num_iters = num_iters
        d.use_graph = 1 if use_graph else 0
        d.keep_samples = 1 if keep_samples else 0
        if slide_control_scale is not None:
            for i, v in enumerate(slide_control_scale):
                d.slide_control_scale[i] = v
        self._keep = _host_arrays(d, problem)
        if init_control is not None:
            ic = np.ascontiguousarray(init_control, dtype=np.float32).reshape(self.T, self.C)
            self._keep.append(ic)
            d.init_control = ic.ctypes.data
        self._num_iters, self._use_graph, self._keep_samples = num_iters, d.use_graph, d.keep_samples
        self._slide = [d.slide_control_scale[i] for i in range(capi.MAX_CONTROL)]
        self._extend_desc(d)
        h = ct.c_void_p()
        capi.call("mppi_controller_create", ct.byref(d), ct.byref(h))
        self._h = h
        self._lib = lib

    def _extend_desc(self, d):
        """Controller-kind fields of the descriptor (none for the vanilla controller)."""

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mppi_controller_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------------------------------------------- reference API
    def computeControl(self, state: np.ndarray, optimization_stride: int = 1):
        x = np.ascontiguousarray(state, dtype=np.float32)
        capi.call("mppi_controller_compute_control", self._h, x.ctypes.data, int(optimization_stride))

    def stepDevice(self, num_calls: int):
        """Enqueue device-resident iterations (no host transfer, no synchronisation)."""
        capi.call("mppi_controller_step_device", self._h, int(num_calls))

    def getControlSeq(self) -> np.ndarray:
        out = np.empty((self.T, self.C), dtype=np.float32)
        capi.call("mppi_controller_get_control_seq", self._h, out.ctypes.data)
        return out

    def getRawMean(self) -> np.ndarray:
        out = np.empty((self.T, self.C), dtype=np.float32)
        capi.call("mppi_controller_get_raw_mean", self._h, out.ctypes.data)
        return out

    def getTargetStateSeq(self) -> np.ndarray:
        out = np.empty((self.T, self.S), dtype=np.float32)
        capi.call("mppi_controller_get_state_seq", self._h, out.ctypes.data)
        return out

    def getWeightStats(self) -> capi.WeightStats:
        st = capi.WeightStats()
        capi.call("mppi_controller_get_weight_stats", self._h, ct.byref(st))
        return st

    def getBaselineCost(self) -> float:
        return self.getWeightStats().baseline

    def getNormalizerCost(self) -> float:
        return self.getWeightStats().normalizer

    def getFreeEnergyStatistics(self) -> capi.FreeEnergy:
        fe = capi.FreeEnergy()
        capi.call("mppi_controller_get_free_energy", self._h, ct.byref(fe))
        return fe

    def updateImportanceSampler(self, u: np.ndarray):
        arr = np.ascontiguousarray(u, dtype=np.float32).reshape(self.T, self.C)
        capi.call("mppi_controller_set_control_seq", self._h, arr.ctypes.data)

    setControlSeq = updateImportanceSampler

    def slideControlSequence(self, steps: int):
        capi.call("mppi_controller_slide_control_seq", self._h, int(steps))

    def setParams(self, problem: Problem, lambda_: float | None = None, alpha: float | None = None,
                  dt: float | None = None):
        """Swap new parameter blocks (dynamics, cost, sampler, dt, lambda, alpha; the problem's cost map and network
        are uploaded again) into the running controller (mppi_controller_set_params): same sizes and kinds only."""
        d = capi.ControllerDesc()
        d.dynamics = problem.dyn
        d.cost = problem.cost
        d.sampler = problem.sampler
        d.num_rollouts = self.K
        d.num_timesteps = self.T
        d.dt = problem.dt if dt is None else dt
        d.lambda_ = problem.lambda_ if lambda_ is None else lambda_
        d.alpha = problem.alpha if alpha is None else alpha
        d.num_iters = self._num_iters
        d.use_graph = self._use_graph
        d.keep_samples = self._keep_samples
        for i in range(capi.MAX_CONTROL):
            d.slide_control_scale[i] = self._slide[i]
        keep = _host_arrays(d, problem)  # noqa: F841 (alive across the call)
        self._extend_desc(d)
        capi.call("mppi_controller_set_params", self._h, ct.byref(d))
        self.problem = problem

    def updateCostMap(self, values: np.ndarray, column_major: bool = False):
        """Stage new cost-map values (TwoDTextureHelper::updateTexture): [height][width], or [width][height] with
        column_major; applied at the start of the next call (mppi_controller_update_cost_map)."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        h, w = (v.shape[1], v.shape[0]) if column_major else v.shape
        capi.call("mppi_controller_update_cost_map", self._h, v.ctypes.data, int(w), int(h), int(column_major))

    def updateCostMapFrame(self, origin, resolution: float, rotation=None):
        """Stage a new map frame (updateOrigin / updateResolution / updateRotation): origin (x, y), resolution and an
        optional row-major 2 x 2 rotation (mppi_controller_update_cost_map_frame)."""
        rot = None if rotation is None else np.ascontiguousarray(rotation, dtype=np.float32).reshape(4)
        capi.call("mppi_controller_update_cost_map_frame", self._h, float(origin[0]), float(origin[1]),
                  float(resolution), None if rot is None else rot.ctypes.data)

    def setSeedCUDARandomNumberGen(self, seed: int):
        capi.call("mppi_controller_set_seed", self._h, ct.c_uint64(seed))

    def device_buffers(self):
        s, c, m = ct.c_void_p(), ct.c_void_p(), ct.c_void_p()
        capi.call("mppi_controller_device_buffers", self._h, ct.byref(s), ct.byref(c), ct.byref(m))
        return s.value, c.value, m.value

    def stream(self):
        s = ct.c_void_p()
        capi.call("mppi_controller_stream", self._h, ct.byref(s))
        return s.value

    # ------------------------------------------------------------------------------------------ learned proposal
    def setProposal(self, net, num_rollouts: int, fallback_tolerance: float | None = 0.0):
        """Attach a ProposalNet (None detaches): each computeControl then runs one proposal iteration of
        num_rollouts samples, falling back to the plain iterations when its weighted cost exceeds the warm start's
        by more than fallback_tolerance (relative; None disables the fallback)."""
        self._proposal = net
        tol = -1.0 if fallback_tolerance is None else float(fallback_tolerance)
        capi.call("mppi_controller_set_proposal", self._h, None if net is None else net.handle, int(num_rollouts),
                  tol)

    def getProposalStats(self):
        n = ct.c_int()
        f = ct.c_int()
        st = capi.WeightStats()
        c0 = ct.c_float()
        capi.call("mppi_controller_get_proposal_stats", self._h, ct.byref(n), ct.byref(f), ct.byref(st), ct.byref(c0))
        return n.value, f.value, st, c0.value

    def timeKernels(self, reps: int = 50):
        """Average (rollout kernel, combine kernel) durations in ms over `reps` device-loop iterations, from HIP events
        on the controller's stream (advances the device loop like stepDevice), and the number of per-workgroup
        partials the combine reduces."""
        a, b, g = ct.c_float(), ct.c_float(), ct.c_int32()
        capi.call("mppi_controller_time_kernels", self._h, int(reps), ct.byref(a), ct.byref(b), ct.byref(g))
        return a.value, b.value, g.value

    def lastCallMs(self) -> float:
        """Stream time of the last computeControl without the host post-processing."""
        v = ct.c_float()
        capi.call("mppi_controller_get_last_call_ms", self._h, ct.byref(v))
        return v.value


class ColoredMPPIController(VanillaMPPIController):
    """ColoredMPPIController (controllers/ColoredMPPI/colored_mppi_controller.cuh:39-200): the vanilla iteration with
    the Tsallis weights when gamma and r are both nonzero (normExp otherwise), the state leash (x0 pulled towards the
    previous solution's state at the last slide's step count), and only control 1 clamped after smoothing. Parameter
    setters swap the new values in through set_params."""

    def __init__(self, problem: Problem, num_iters: int = 1, use_graph: bool = False,
                 init_control: np.ndarray | None = None, slide_control_scale=None, keep_samples: bool = False,
                 gamma: float = 0.0, r: float = 2.0, leash_active: bool = False, state_leash_dist=None):
        self._gamma, self._r, self._leash_active = float(gamma), float(r), bool(leash_active)
        self._leash = np.zeros(capi.MAX_STATE, np.float32)
        if state_leash_dist is not None:
            self._leash[: len(state_leash_dist)] = state_leash_dist
        super().__init__(problem, num_iters, use_graph, init_control, slide_control_scale, keep_samples)

    def _extend_desc(self, d):
        d.controller_kind = capi.CONTROLLER_COLORED
        d.tsallis_gamma, d.tsallis_r, d.leash_active = self._gamma, self._r, int(self._leash_active)
        for i in range(capi.MAX_STATE):
            d.state_leash_dist[i] = float(self._leash[i])

    def _push(self):
        self.setParams(self.problem)

    def getControllerName(self) -> str:
        return "Colored MPPI"

    def setGamma(self, gamma: float):
        self._gamma = float(gamma)
        self._push()

    def getGamma(self) -> float:
        return self._gamma

    def setRExp(self, r: float):
        self._r = float(r)
        self._push()

    def getRExp(self) -> float:
        return self._r

    def setLeashActive(self, active: bool):
        self._leash_active = bool(active)
        self._push()

    def getLeashActive(self) -> bool:
        return self._leash_active

    def setStateLeashLength(self, length: float, index: int = 0):
        self._leash[index] = length
        self._push()

    def getStateLeashLength(self, index: int) -> float:
        return float(self._leash[index])

    def setNoiseDecay(self, decay: float):
        self.problem.sampler.std_dev_decay = decay
        self._push()

    def setOffsetDecayRate(self, rate: float):
        self.problem.sampler.offset_decay_rate = rate
        self._push()

    def getOffsetDecayRate(self) -> float:
        return self.problem.sampler.offset_decay_rate

    def setColoredNoiseExponents(self, exponents):
        for i, e in enumerate(exponents):
            self.problem.sampler.exponents[i] = e
        self._push()

    def getColoredNoiseExponents(self):
        return [self.problem.sampler.exponents[i] for i in range(self.C)]


class TubeMPPIController:
    """TubeMPPIController (controllers/Tube-MPPI/tube_mppi_controller.cuh:23-116) over the two-system controller of
    the C-ABI (mppi_tube_*): an actual and a nominal system optimised together, the nominal one replaced by the actual
    one when the actual baseline is below the nominal baseline + nominal_threshold."""

    MODE = capi.TWO_SYSTEM_TUBE

    def __init__(self, problem: Problem, num_iters: int = 1, init_control: np.ndarray | None = None,
                 nominal_threshold: float = 20.0, value_function_threshold: float = 1000.0, slide_control_scale=None):
        self._lib = capi.load()
        self.problem = problem
        self.T, self.K, self.S, self.C = problem.num_timesteps, problem.num_rollouts, problem.S, problem.C
        d = capi.TubeDesc()
        b = d.base
        b.dynamics, b.cost, b.sampler = problem.dyn, problem.cost, problem.sampler
        b.num_rollouts, b.num_timesteps = self.K, self.T
        b.dt, b.lambda_, b.alpha, b.num_iters = problem.dt, problem.lambda_, problem.alpha, num_iters
        if slide_control_scale is not None:
            for i, v in enumerate(slide_control_scale):
                b.slide_control_scale[i] = v
        self._keep = _host_arrays(b, problem)
        if init_control is not None:
            ic = np.ascontiguousarray(init_control, dtype=np.float32).reshape(self.T, self.C)
            self._keep.append(ic)
            b.init_control = ic.ctypes.data
        d.mode = self.MODE
        self._extend_tube_desc(d)
        d.nominal_threshold = nominal_threshold
        d.value_function_threshold = value_function_threshold
        self._thresholds = [nominal_threshold, value_function_threshold]
        h = ct.c_void_p()
        capi.call("mppi_tube_create", ct.byref(d), ct.byref(h))
        self._h = h

    def _extend_tube_desc(self, d):
        """RMPPI's candidate-evaluation sizes (unused by Tube-MPPI)."""

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mppi_tube_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, name, nominal, shape):
        out = np.empty(shape, dtype=np.float32)
        capi.call(name, self._h, int(nominal), out.ctypes.data)
        return out

    # --------------------------------------------------------------------------------------------- reference API
    def computeControl(self, state: np.ndarray, optimization_stride: int = 1):
        x = np.ascontiguousarray(state, dtype=np.float32)
        capi.call("mppi_tube_compute_control", self._h, x.ctypes.data, int(optimization_stride))

    def getControlSeq(self) -> np.ndarray:
        """The nominal control sequence (tube_mppi_controller.cuh:58-64)."""
        return self._get("mppi_tube_get_control_seq", 1, (self.T, self.C))

    def getActualControlSeq(self) -> np.ndarray:
        return self._get("mppi_tube_get_control_seq", 0, (self.T, self.C))

    def setControlSeq(self, u: np.ndarray, nominal: bool = True):
        a = np.ascontiguousarray(u, dtype=np.float32).reshape(self.T, self.C)
        capi.call("mppi_tube_set_control_seq", self._h, int(nominal), a.ctypes.data)

    def getTargetStateSeq(self) -> np.ndarray:
        """The nominal state trajectory (tube_mppi_controller.cuh:66-72)."""
        return self._get("mppi_tube_get_state_seq", 1, (self.T, self.S))

    def getActualStateSeq(self) -> np.ndarray:
        return self._get("mppi_tube_get_state_seq", 0, (self.T, self.S))

    def getWeightStats(self, nominal: bool) -> capi.WeightStats:
        st = capi.WeightStats()
        capi.call("mppi_tube_get_weight_stats", self._h, int(nominal), ct.byref(st))
        return st

    def getBaselineCost(self, idx: int = 0) -> float:
        """Baseline of distribution idx in the reference's indexing (Tube: 0 actual, 1 nominal)."""
        return self.getWeightStats(self._is_nominal(idx)).baseline

    def getNormalizerCost(self, idx: int = 0) -> float:
        return self.getWeightStats(self._is_nominal(idx)).normalizer

    def _is_nominal(self, idx: int) -> bool:
        return idx == 1

    def getFreeEnergyStatistics(self, nominal: bool) -> capi.FreeEnergy:
        fe = capi.FreeEnergy()
        capi.call("mppi_tube_get_free_energy", self._h, int(nominal), ct.byref(fe))
        return fe

    def getNominalStateUsed(self) -> int:
        v = ct.c_int32()
        capi.call("mppi_tube_get_nominal_state_used", self._h, ct.byref(v))
        return v.value

    def slideControlSequence(self, steps: int):
        capi.call("mppi_tube_slide_control_seq", self._h, int(steps))

    def setNominalState(self, x: np.ndarray):
        a = np.ascontiguousarray(x, dtype=np.float32)
        capi.call("mppi_tube_set_nominal_state", self._h, a.ctypes.data)

    def getNominalState(self) -> np.ndarray:
        out = np.empty(self.S, np.float32)
        capi.call("mppi_tube_get_nominal_state", self._h, out.ctypes.data)
        return out

    def setFeedbackGains(self, gains: np.ndarray | None):
        if gains is None:
            capi.call("mppi_tube_set_feedback_gains", self._h, None)
            return
        a = np.ascontiguousarray(gains, dtype=np.float32).reshape(self.T, self.S, self.C)
        capi.call("mppi_tube_set_feedback_gains", self._h, a.ctypes.data)

    def getFeedbackControl(self, state: np.ndarray, goal_state: np.ndarray, t: int) -> np.ndarray:
        x = np.ascontiguousarray(state, np.float32)
        g = np.ascontiguousarray(goal_state, np.float32)
        out = np.empty(self.C, np.float32)
        capi.call("mppi_tube_get_feedback_control", self._h, x.ctypes.data, g.ctypes.data, int(t), out.ctypes.data)
        return out

    def getNominalThreshold(self) -> float:
        return self._thresholds[0]

    def setNominalThreshold(self, v: float):
        self._thresholds[0] = float(v)
        capi.call("mppi_tube_set_thresholds", self._h, *self._thresholds)

    def getValueFunctionThreshold(self) -> float:
        return self._thresholds[1]

    def setValueFunctionThreshold(self, v: float):
        self._thresholds[1] = float(v)
        capi.call("mppi_tube_set_thresholds", self._h, *self._thresholds)

    def getSamples(self):
        """The last iteration's applied controls [2][K][T][C] and costs [2][K] (buffer system order)."""
        u = np.empty((2, self.K, self.T, self.C), np.float32)
        c = np.empty((2, self.K), np.float32)
        capi.call("mppi_tube_get_samples", self._h, u.ctypes.data, c.ctypes.data)
        return u, c

    def timeKernels(self, reps: int = 50):
        """Average (fused two-system rollout kernel, two-system combine) durations in ms over `reps` repeats of the last
        iteration (Gaussian sampler), from HIP events on the controller's stream, and the number of per-workgroup
        partials each combine reduces."""
        a, b, g = ct.c_float(), ct.c_float(), ct.c_int32()
        capi.call("mppi_tube_time_kernels", self._h, int(reps), ct.byref(a), ct.byref(b), ct.byref(g))
        return a.value, b.value, g.value


class RobustMPPIController(TubeMPPIController):
    """The optimisation of RobustMPPIController (controllers/R-MPPI/robust_mppi_controller.cu:635-756): nominal
    (distribution 0) and real (distribution 1) systems sampled around the nominal control, the real one tracking the
    nominal with the feedback gains; the nominal state is set by the caller (setNominalState)."""

    MODE = capi.TWO_SYSTEM_RMPPI

    def __init__(self, problem: Problem, num_iters: int = 1, init_control: np.ndarray | None = None,
                 value_function_threshold: float = 1000.0, slide_control_scale=None, num_candidates: int = 9,
                 num_eval_samples: int = 64):
        self._cand = (num_candidates, num_eval_samples)
        super().__init__(problem, num_iters, init_control, 20.0, value_function_threshold, slide_control_scale)

    def _extend_tube_desc(self, d):
        d.num_candidates, d.num_eval_samples = self._cand

    def updateImportanceSamplingControl(self, state: np.ndarray, stride: int = 1):
        """The candidate evaluation that picks the nominal state and slides the nominal control
        (robust_mppi_controller.cu:548-628)."""
        x = np.ascontiguousarray(state, dtype=np.float32)
        capi.call("mppi_tube_update_importance_sampling_control", self._h, x.ctypes.data, int(stride))

    def getCandidates(self):
        """The last evaluation: candidate states [N][S], strides [N], costs [N][M], free energies [N], chosen index."""
        N, M = self._cand
        st = np.empty((N, self.S), np.float32)
        sr = np.empty(N, np.int32)
        co = np.empty((N, M), np.float32)
        fe = np.empty(N, np.float32)
        bi = ct.c_int32()
        capi.call("mppi_tube_get_candidates", self._h, st.ctypes.data, sr.ctypes.data, co.ctypes.data, fe.ctypes.data,
                  ct.byref(bi))
        return st, sr, co, fe, bi.value

    def _is_nominal(self, idx: int) -> bool:
        return idx == 0

    def getControlSeq(self) -> np.ndarray:
        """The real system's control sequence."""
        return self._get("mppi_tube_get_control_seq", 0, (self.T, self.C))

    def getNominalControlSeq(self) -> np.ndarray:
        return self._get("mppi_tube_get_control_seq", 1, (self.T, self.C))
