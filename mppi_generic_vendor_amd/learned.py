"""Learned-proposal MPPI on the bicycle + residual-network task: training data from plain MPPI, training of the
proposal network on the GPU, and closed-loop comparison with plain MPPI.

- Data: closed-loop plain MPPI (K samples, `iters` iterations per control step) from seeded random starts; every
  control step gives one pair (state, previous control sequence) -> optimized sequence (the raw weighted mean).
- Training target: the residual r = (u* - u_prev) / sigma_base per (t, c); the network outputs a mean shift and a
  log-std in the same units (mppi_proposal_* in include/mppi_amd.h).
- Evaluation: the cost of the executed trajectory (sum of running costs of the visited states) from held-out starts
  under (a) plain MPPI at its converged setting, (b) the learned proposal with a fallback to (a), (c) plain MPPI with
  the proposal's sample budget and one iteration (the method without the network), plus each controller's stream
  time per control step.
"""
from __future__ import annotations

import ctypes as ct
import dataclasses
import statistics
import time
import threading

import numpy as np

from . import capi, workloads
from .controller import VanillaMPPIController
from .proposal import ProposalNet, init_params, network_inputs


def task(num_rollouts: int, cost_map: np.ndarray | None = None, seed: int = 42) -> workloads.Problem:
    """The bicycle with the 9-64-64-4 BF16 residual network, the blob cost map and the 8 m/s speed term."""
    return workloads.bicycle_nn(num_rollouts=num_rollouts, num_timesteps=100, seed=seed, cost_map=cost_map,
                                precision=capi.NN_BF16)


def random_start(rng: np.random.Generator) -> np.ndarray:
    x = np.zeros(7, np.float32)
    x[0] = rng.uniform(3.0, 8.0)
    x[1] = rng.uniform(-np.pi, np.pi)
    x[2:4] = rng.uniform(-30.0, 30.0, 2)
    x[4] = rng.uniform(-0.2, 0.2)
    return x


def kick(rng: np.random.Generator, x: np.ndarray) -> np.ndarray:
    """A disturbance of the running system: new speed ~ U(2, 10) m/s, heading turned by ~ U(-1, 1) rad and steering
    angle ~ U(-0.3, 0.3). Right after one the warm start (the previous plan, slid by a step) no longer fits the state."""
    y = x.copy()
    y[0] = rng.uniform(2.0, 10.0)
    y[1] = y[1] + rng.uniform(-1.0, 1.0)
    y[4] = rng.uniform(-0.3, 0.3)
    return y.astype(np.float32)


class Plant:
    """The true system for closed-loop runs: the same model stepped on the host (mppi_host_step)."""

    def __init__(self, prob: workloads.Problem):
        self.dyn = capi.DynamicsParams.from_buffer_copy(prob.dyn)
        self._theta = None
        if prob.nn_theta is not None:
            self._theta = np.ascontiguousarray(prob.nn_theta, np.float32)
            self.dyn.nn_theta = self._theta.ctypes.data
        self.cost = capi.CostParams.from_buffer_copy(prob.cost)
        self._map = np.ascontiguousarray(prob.cost_map, np.float32) if prob.cost_map is not None else None
        if self._map is not None:
            self.cost.map = self._map.ctypes.data
        self.dt = prob.dt

    def step(self, x: np.ndarray, u: np.ndarray) -> np.ndarray:
        xn = np.zeros(7, np.float32)
        uc = np.ascontiguousarray(u, np.float32).copy()
        capi.call("mppi_host_enforce_constraints", ct.byref(self.dyn), uc.ctypes.data)
        xa = np.ascontiguousarray(x, np.float32)
        capi.call("mppi_host_step", ct.byref(self.dyn), xa.ctypes.data, uc.ctypes.data, self.dt, xn.ctypes.data,
                  None, None)
        return xn

    def cost_of(self, y: np.ndarray) -> float:
        out = ct.c_float()
        ya = np.ascontiguousarray(y, np.float32)
        capi.call("mppi_host_running_cost", ct.byref(self.cost), ya.ctypes.data, ct.byref(out))
        return out.value


@dataclasses.dataclass
class Dataset:
    x: np.ndarray      # [N][S] states
    prev: np.ndarray   # [N][T][C] warm-start sequences
    opt: np.ndarray    # [N][T][C] optimized sequences

    def inputs(self) -> np.ndarray:
        return network_inputs(self.x, self.prev)

    def targets(self, sigma: np.ndarray) -> np.ndarray:
        r = (self.opt - self.prev) / sigma[None, None, :]
        return r.reshape(r.shape[0], -1).astype(np.float32)


def generate_pairs(num_pairs: int, seed: int = 1, num_rollouts: int = 8192, iters: int = 3,
                   steps_per_episode: int = 50, workers: int = 4, cost_map: np.ndarray | None = None,
                   log=None, kick_every: int = 0) -> Dataset:
    """Closed-loop plain MPPI from seeded random starts, `workers` controllers on their own streams; with kick_every
    > 0 the system is disturbed (`kick`) every kick_every control steps."""
    m = workloads.gaussian_blob_map() if cost_map is None else cost_map
    per = [num_pairs // workers + (1 if w < num_pairs % workers else 0) for w in range(workers)]
    out = [None] * workers
    errors = []

    def run(w):
        try:
            prob = task(num_rollouts, cost_map=m, seed=1000 * seed + w)
            plant = Plant(prob)
            ctrl = VanillaMPPIController(prob, num_iters=iters)
            rng = np.random.default_rng(seed * 7919 + w)
            T, C = prob.num_timesteps, prob.C
            xs, ps, us = [], [], []
            while len(xs) < per[w]:
                x = random_start(rng)
                ctrl.updateImportanceSampler(np.zeros((T, C), np.float32))
                for step in range(steps_per_episode):
                    if len(xs) >= per[w]:
                        break
                    if kick_every > 0 and step > 0 and step % kick_every == 0:
                        x = kick(rng, x)
                    prev = ctrl.getControlSeq()
                    ctrl.computeControl(x, 1)
                    xs.append(x.copy())
                    ps.append(prev)
                    us.append(ctrl.getRawMean())
                    x = plant.step(x, ctrl.getControlSeq()[0])
                    ctrl.slideControlSequence(1)
            ctrl.close()
            out[w] = (np.array(xs, np.float32), np.array(ps, np.float32), np.array(us, np.float32))
        except Exception as e:  # surfaced after join
            errors.append(e)

    threads = [threading.Thread(target=run, args=(w,)) for w in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    ds = Dataset(*(np.concatenate([o[i] for o in out]) for i in range(3)))
    if log:
        log(f"[learned] {ds.x.shape[0]} pairs from plain MPPI (K={num_rollouts}, {iters} iterations)")
    return ds


def train(ds: Dataset, sigma: np.ndarray, epochs: int = 30, batch: int = 512, lr: float = 1e-3, seed: int = 7,
          log=None) -> ProposalNet:
    """Normalizes the inputs (per-feature mean / std of the data) and trains the network on the GPU."""
    X = ds.inputs()
    shift = X.mean(axis=0).astype(np.float32)
    scale = (1.0 / (X.std(axis=0) + 1e-3)).astype(np.float32)
    Xn = ((X - shift) * scale).astype(np.float32)
    R = ds.targets(sigma)
    S, C, T = ds.x.shape[1], ds.prev.shape[2], ds.prev.shape[1]
    net = ProposalNet(S, C, T, init_params(S, C, T, seed))
    net.set_normalization(shift, scale)
    net.train(Xn, R, epochs=epochs, batch=min(batch, Xn.shape[0]), lr=lr, seed=seed, log=log)
    return net


def closed_loop(prob: workloads.Problem, ctrl: VanillaMPPIController, starts: list[np.ndarray], steps: int,
                wall: list | None = None, kick_every: int = 0, kick_seed: int = 5):
    """Executed-trajectory cost per episode and the stream time of every control step (appended to `wall`: the
    host wall time of every computeControl call). With kick_every > 0 the system is disturbed every kick_every
    steps by the same seeded sequence of kicks for every controller."""
    plant = Plant(prob)
    T, C = prob.num_timesteps, prob.C
    costs, ms = [], []
    krng = np.random.default_rng(kick_seed)
    for x0 in starts:
        x = x0.copy()
        ctrl.updateImportanceSampler(np.zeros((T, C), np.float32))
        total = 0.0
        for step in range(steps):
            if kick_every > 0 and step > 0 and step % kick_every == 0:
                x = kick(krng, x)
            t0 = time.perf_counter()
            ctrl.computeControl(x, 1)
            if wall is not None:
                wall.append((time.perf_counter() - t0) * 1e3)
            ms.append(ctrl.lastCallMs())
            x = plant.step(x, ctrl.getControlSeq()[0])
            total += plant.cost_of(x)
            ctrl.slideControlSequence(1)
        costs.append(total / steps)
    return np.array(costs), ms


def evaluate(net: ProposalNet | None, episodes: int = 10, steps: int = 100, seed: int = 99,
             plain_rollouts: int = 8192, plain_iters: int = 3, proposal_rollouts: int = 2048,
             fallback_tolerance: float = 0.005, cost_map: np.ndarray | None = None, log=None,
             use_graph: bool = False, kick_every: int = 0) -> dict:
    m = workloads.gaussian_blob_map() if cost_map is None else cost_map
    rng = np.random.default_rng(seed)
    starts = [random_start(rng) for _ in range(episodes)]
    res = {}
    modes = [("plain", plain_rollouts, plain_iters, False), ("plain_small", proposal_rollouts, 1, False)]
    if net is not None:
        modes.insert(1, ("learned", plain_rollouts, plain_iters, True))
    for name, K, iters, learned in modes:
        prob = task(K, cost_map=m, seed=4242)
        ctrl = VanillaMPPIController(prob, num_iters=iters, use_graph=use_graph)
        if learned:
            ctrl.setProposal(net, proposal_rollouts, fallback_tolerance=fallback_tolerance)
        ctrl.computeControl(starts[0], 1)  # warm-up (graph capture)
        ctrl.computeControl(starts[0], 1)
        wall = []
        c, ms = closed_loop(prob, ctrl, starts, steps, wall, kick_every=kick_every)
        entry = {"cost": float(c.mean()), "costs": c.tolist(), "ms_per_call": float(np.mean(ms)),
                 "ms_median": float(statistics.median(ms)), "hz": 1e3 / float(np.mean(ms)),
                 "ms_host_median": float(statistics.median(wall)), "hz_host": 1e3 / float(statistics.median(wall)),
                 "samples": proposal_rollouts if learned else K, "iterations": 1 if learned else iters}
        if learned:
            n, f, _, _ = ctrl.getProposalStats()
            entry["fallback_fraction"] = f / max(1, n)
        res[name] = entry
        ctrl.close()
        if log:
            log(f"[learned] {name}: cost {entry['cost']:.4f}, {entry['ms_per_call']:.4f} ms per control step")
    return res


def sequence_cost(prob: workloads.Problem, plant: Plant, x0: np.ndarray, u: np.ndarray) -> float:
    """Cost of one control sequence from x0 through the model: mean running cost of the visited states."""
    x = x0.copy()
    total = 0.0
    for t in range(u.shape[0]):
        x = plant.step(x, u[t])
        total += plant.cost_of(x)
    return total / u.shape[0]


def cold_start_quality(net: ProposalNet | None, starts: int = 32, seed: int = 11, plain_rollouts: int = 8192,
                       plain_iters: int = 3, proposal_rollouts: int = 2048, cost_map: np.ndarray | None = None,
                       log=None) -> dict:
    """One control step from a cold start (zero warm start) per state: the cost of the returned sequence for plain
    MPPI at its converged setting, the learned proposal (one iteration, no fallback) and plain MPPI with the
    proposal's samples and one iteration. Isolates the optimizer's quality from closed-loop divergence."""
    m = workloads.gaussian_blob_map() if cost_map is None else cost_map
    rng = np.random.default_rng(seed)
    xs = [random_start(rng) for _ in range(starts)]
    res = {}
    modes = [("plain", plain_rollouts, plain_iters, False), ("plain_small", proposal_rollouts, 1, False)]
    if net is not None:
        modes.insert(1, ("learned", plain_rollouts, plain_iters, True))
    for name, K, iters, learned_mode in modes:
        prob = task(K, cost_map=m, seed=4243)
        plant = Plant(prob)
        ctrl = VanillaMPPIController(prob, num_iters=iters)
        if learned_mode:
            ctrl.setProposal(net, proposal_rollouts, fallback_tolerance=None)
        T, C = prob.num_timesteps, prob.C
        js = []
        for x in xs:
            ctrl.updateImportanceSampler(np.zeros((T, C), np.float32))
            ctrl.computeControl(x, 1)
            js.append(sequence_cost(prob, plant, x, ctrl.getControlSeq()))
        ctrl.close()
        res[name] = {"cost": float(np.mean(js)), "costs": js}
        if log:
            log(f"[learned] cold start {name}: sequence cost {res[name]['cost']:.4f}")
    return res

