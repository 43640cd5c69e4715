"""Python host wrapper of the C-ABI batch object: many independent MPPI instances advanced by one persistent launch
per iteration (mppi_batch_* in include/mppi_amd.h)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import capi
from .workloads import Problem


class MPPIBatch:
    def __init__(self, problems: list[Problem], cost_map: np.ndarray | None = None):
        capi.load()
        self.problems = problems
        n = len(problems)
        descs = (capi.InstanceDesc * n)()
        self._keep = []
        for d, p in zip(descs, problems):
            d.dynamics = p.dyn
            d.cost = p.cost
            d.cost.map = None
            d.sampler = p.sampler
            d.num_rollouts = p.num_rollouts
            d.num_timesteps = p.num_timesteps
            d.dt = p.dt
            d.lambda_ = p.lambda_
            d.alpha = p.alpha
            d.init_control = None
            if p.std_dev_t is not None:
                a = np.ascontiguousarray(p.std_dev_t, dtype=np.float32).reshape(p.num_timesteps, p.C)
                self._keep.append(a)
                d.std_dev_t_host = a.ctypes.data
            if p.goal_traj is not None:
                g = np.ascontiguousarray(p.goal_traj, dtype=np.float32).reshape(-1, p.dyn.output_dim)
                self._keep.append(g)
                d.goal_traj_host = g.ctypes.data
                d.cost.goal_traj_len = g.shape[0]
        mp = None
        h = w = 0
        if cost_map is not None:
            m = np.ascontiguousarray(cost_map, dtype=np.float32)
            self._keep.append(m)
            mp = m.ctypes.data
            h, w = m.shape
        handle = ct.c_void_p()
        capi.call("mppi_batch_create", descs, n, mp, h, w, ct.byref(handle))
        self._h = handle
        self.n = n

    # ------------------------------------------------------------------------------------------------ control
    def packStates(self, x0) -> np.ndarray:
        """The instances' states (a list of per-instance vectors) as the [n][MAX_STATE] array computeControl takes."""
        x = np.zeros((self.n, capi.MAX_STATE), np.float32)
        for i, p in enumerate(self.problems):
            x[i, :p.S] = np.asarray(x0[i], np.float32)[:p.S]
        return x

    def computeControl(self, x0, stride: int = 1):
        """x0: a float32 [n][MAX_STATE] array (used as is), or a list of per-instance state vectors."""
        if isinstance(x0, np.ndarray) and x0.dtype == np.float32 and x0.shape == (self.n, capi.MAX_STATE) \
                and x0.flags.c_contiguous:
            x = x0
        else:
            x = self.packStates(x0)
        capi.call("mppi_batch_compute_control", self._h, x.ctypes.data, stride)

    def setGraphReplay(self, enable: bool):
        """Replay computeControl's launches from a captured hipGraph (default: direct launches)."""
        capi.call("mppi_batch_set_graph_replay", self._h, 1 if enable else 0)

    def stepDevice(self, num_calls: int):
        capi.call("mppi_batch_step_device", self._h, num_calls)

    def autotune(self, reps: int = 3) -> capi.BatchConfig:
        c = capi.BatchConfig()
        capi.call("mppi_batch_autotune", self._h, reps, ct.byref(c))
        return c

    def setConfig(self, threads: int, samples_per_lane: int, chunk: int):
        c = capi.BatchConfig(threads, samples_per_lane, chunk, 0, 0.0)
        capi.call("mppi_batch_set_config", self._h, ct.byref(c))

    def getConfig(self) -> capi.BatchConfig:
        c = capi.BatchConfig()
        capi.call("mppi_batch_get_config", self._h, ct.byref(c))
        return c

    def _seq(self, name: str, i: int, shape):
        out = np.zeros(shape, np.float32)
        capi.call(name, self._h, i, out.ctypes.data)
        return out

    def getControlSeq(self, i: int) -> np.ndarray:
        p = self.problems[i]
        return self._seq("mppi_batch_get_control_seq", i, (p.num_timesteps, p.C))

    def getRawMean(self, i: int) -> np.ndarray:
        p = self.problems[i]
        return self._seq("mppi_batch_get_raw_mean", i, (p.num_timesteps, p.C))

    def getTargetStateSeq(self, i: int) -> np.ndarray:
        p = self.problems[i]
        return self._seq("mppi_batch_get_state_seq", i, (p.num_timesteps, p.S))

    def getWeightStats(self, i: int) -> capi.WeightStats:
        st = capi.WeightStats()
        capi.call("mppi_batch_get_weight_stats", self._h, i, ct.byref(st))
        return st

    def getFreeEnergyStatistics(self, i: int) -> capi.FreeEnergy:
        fe = capi.FreeEnergy()
        capi.call("mppi_batch_get_free_energy", self._h, i, ct.byref(fe))
        return fe

    def setControlSeq(self, i: int, u: np.ndarray):
        p = self.problems[i]
        a = np.ascontiguousarray(u, dtype=np.float32).reshape(p.num_timesteps, p.C)
        capi.call("mppi_batch_set_control_seq", self._h, i, a.ctypes.data)

    def slideControlSequence(self, steps: int):
        capi.call("mppi_batch_slide_control_seq", self._h, steps)

    def getSamples(self, i: int):
        p = self.problems[i]
        u = np.zeros((p.num_rollouts, p.num_timesteps, p.C), np.float32)
        c = np.zeros(p.num_rollouts, np.float32)
        capi.call("mppi_batch_get_samples", self._h, i, u.ctypes.data, c.ctypes.data)
        return u, c

    def stream(self) -> int:
        s = ct.c_void_p()
        capi.call("mppi_batch_stream", self._h, ct.byref(s))
        return s.value

    def timeKernels(self, reps: int = 20):
        """Average (persistent rollout kernel, combine kernel) durations in ms over `reps` device-loop iterations,
        from HIP events on the batch's stream."""
        a, b = ct.c_float(), ct.c_float()
        capi.call("mppi_batch_time_kernels", self._h, int(reps), ct.byref(a), ct.byref(b))
        return a.value, b.value

    def close(self):
        if self._h:
            capi.call("mppi_batch_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
