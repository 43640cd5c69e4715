"""Synthetic, seeded MPPI problems: cart-pole, double integrator on a cost map, kinematic bicycle with a residual MLP.

Each builder returns a ``Problem`` holding the C-ABI parameter blocks, the sizes and the host arrays (initial state,
cost map, MLP parameters). Seeds follow the reference tests: 7 for states, maps and weights, 42 for the noise.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import capi


@dataclasses.dataclass
class Problem:
    name: str
    dyn: capi.DynamicsParams
    cost: capi.CostParams
    sampler: capi.SamplerParams
    num_rollouts: int
    num_timesteps: int
    dt: float
    lambda_: float
    alpha: float
    x0: np.ndarray
    cost_map: np.ndarray | None = None
    nn_theta: np.ndarray | None = None
    # per-step standard deviations [T][C] (time_specific_std_dev) and a goal trajectory [H][O] for the quadratic
    # cost (QuadraticCostTrajectoryParams); host arrays, copied by the controller
    std_dev_t: np.ndarray | None = None
    goal_traj: np.ndarray | None = None

    @property
    def S(self) -> int:
        return self.dyn.state_dim

    @property
    def C(self) -> int:
        return self.dyn.control_dim


def gaussian_blob_map(width: int = 1024, height: int = 1024, blobs: int = 64, seed: int = 7) -> np.ndarray:
    """Sum of `blobs` Gaussian bumps: centres ~ U(map), sigma ~ U(5, 40) cells, height ~ U(0, 50)."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, width, blobs)
    cy = rng.uniform(0, height, blobs)
    sig = rng.uniform(5.0, 40.0, blobs)
    hgt = rng.uniform(0.0, 50.0, blobs)
    m = np.zeros((height, width), dtype=np.float64)
    xs = np.arange(width, dtype=np.float64) + 0.5
    ys = np.arange(height, dtype=np.float64) + 0.5
    for i in range(blobs):
        r = int(math.ceil(5 * sig[i]))
        x0, x1 = max(0, int(cx[i]) - r), min(width, int(cx[i]) + r + 1)
        y0, y1 = max(0, int(cy[i]) - r), min(height, int(cy[i]) + r + 1)
        gx = np.exp(-((xs[x0:x1] - cx[i]) ** 2) / (2 * sig[i] ** 2))
        gy = np.exp(-((ys[y0:y1] - cy[i]) ** 2) / (2 * sig[i] ** 2))
        m[y0:y1, x0:x1] += hgt[i] * np.outer(gy, gx)
    return m.astype(np.float32)


def xavier_mlp(layers=(9, 64, 64, 4), seed: int = 7) -> np.ndarray:
    """FNN parameter vector (per layer: row-major W[out][in], then bias[out]) with Xavier-normal weights, zero bias."""
    rng = np.random.default_rng(seed)
    parts = []
    for nin, nout in zip(layers[:-1], layers[1:]):
        std = math.sqrt(2.0 / (nin + nout))
        parts.append(rng.normal(0.0, std, size=(nout, nin)).astype(np.float32).ravel())
        parts.append(np.zeros(nout, dtype=np.float32))
    return np.concatenate(parts).astype(np.float32)


def cartpole(num_rollouts: int = 2048, num_timesteps: int = 100, seed: int = 42) -> Problem:
    """Cart-pole swing-up: Q = diag(10, 1, 10, 1) to (0, 0, pi, 0), terminal 100 x running, 0.1 on the control
    through the sampler's likelihood-ratio cost (alpha = 0), sigma 1, lambda 1, dt 0.02, x0 ~ U(+-0.5)^4 (seed 7)."""
    dyn = capi.dynamics_defaults(capi.MODEL_CARTPOLE)
    cost = capi.cost_defaults(capi.COST_CARTPOLE_QUADRATIC, dyn)
    for i, q in enumerate((10.0, 1.0, 10.0, 1.0)):
        cost.coeff[i] = q
    cost.goal[2] = math.pi
    cost.terminal_cost_coeff = 100.0
    smp = capi.sampler_defaults(capi.SAMPLER_GAUSSIAN, 1)
    smp.std_dev[0] = 1.0
    smp.control_cost_coeff[0] = 0.1
    smp.seed = seed
    x0 = np.random.default_rng(7).uniform(-0.5, 0.5, 4).astype(np.float32)
    return Problem("cartpole", dyn, cost, smp, num_rollouts, num_timesteps, 0.02, 1.0, 0.0, x0)


def double_integrator_map(num_rollouts: int = 8192, num_timesteps: int = 100, seed: int = 42,
                          colored: bool = True, fft_len_mode: int = capi.FFT_LEN_2T,
                          cost_map: np.ndarray | None = None, terminal_cost_coeff: float = 0.0) -> Problem:
    """2-D double integrator with power-law noise (exponents 1, 2) and a 1024^2 blob cost map at 0.05 m/cell
    centred on the origin; quadratic goal term to (2, 2) plus the bilinear map lookup. The terminal cost is 0, as
    QuadraticCost::terminalCost returns (cost_functions/quadratic_cost/quadratic_cost.cu:51-54), unless
    terminal_cost_coeff asks for terminal_cost_coeff x the quadratic term."""
    dyn = capi.dynamics_defaults(capi.MODEL_DOUBLE_INTEGRATOR)
    cost = capi.cost_defaults(capi.COST_QUADRATIC_MAP, dyn)
    for i, q in enumerate((2.0, 2.0, 0.5, 0.5)):
        cost.coeff[i] = q
    cost.goal[0] = 2.0
    cost.goal[1] = 2.0
    cost.terminal_cost_coeff = terminal_cost_coeff
    m = gaussian_blob_map() if cost_map is None else cost_map
    cost.map_height, cost.map_width = m.shape
    cost.map_resolution = 0.05
    cost.map_origin_x = -0.5 * cost.map_width * cost.map_resolution
    cost.map_origin_y = -0.5 * cost.map_height * cost.map_resolution
    cost.map_coeff = 1.0
    cost.map_x_index, cost.map_y_index = 0, 1
    smp = capi.sampler_defaults(capi.SAMPLER_COLORED if colored else capi.SAMPLER_GAUSSIAN, 2)
    smp.exponents[0], smp.exponents[1] = 1.0, 2.0
    smp.fft_len_mode = fft_len_mode
    if fft_len_mode == capi.FFT_LEN_T:
        smp.offset_decay_rate = 0.0
    smp.seed = seed
    x0 = np.random.default_rng(7).uniform(-0.5, 0.5, 4).astype(np.float32)
    return Problem("double_integrator_map", dyn, cost, smp, num_rollouts, num_timesteps, 0.02, 1.0, 0.0, x0,
                   cost_map=m)


def bicycle_nn(num_rollouts: int = 16384, num_timesteps: int = 100, seed: int = 42,
               precision: int = capi.NN_FP32, cost_map: np.ndarray | None = None, residual: bool = True) -> Problem:
    """Kinematic bicycle (RacerDubins parameters) with a 9-64-64-4 tanh residual MLP (Xavier-normal, seed 7), the
    1024^2 blob map at 0.1 m/cell and a speed term toward 8 m/s."""
    model = capi.MODEL_RACER_DUBINS_NN if residual else capi.MODEL_RACER_DUBINS
    dyn = capi.dynamics_defaults(model)
    if residual:
        dyn.nn_precision = precision
    cost = capi.cost_defaults(capi.COST_QUADRATIC_MAP, dyn)
    for i in range(capi.MAX_OUTPUT):
        cost.coeff[i] = 0.0
    m = gaussian_blob_map() if cost_map is None else cost_map
    cost.map_height, cost.map_width = m.shape
    cost.map_resolution = 0.1
    cost.map_origin_x = -0.5 * cost.map_width * cost.map_resolution
    cost.map_origin_y = -0.5 * cost.map_height * cost.map_resolution
    cost.map_coeff = 1.0
    cost.map_x_index, cost.map_y_index = 2, 3
    cost.speed_coeff = 1.0
    cost.speed_target = 8.0
    cost.speed_index = 0
    smp = capi.sampler_defaults(capi.SAMPLER_GAUSSIAN, 2)
    smp.std_dev[0], smp.std_dev[1] = 0.5, 0.3
    smp.seed = seed
    rng = np.random.default_rng(7)
    x0 = np.zeros(7, dtype=np.float32)
    x0[0] = 5.0 + rng.uniform(-0.5, 0.5)
    x0[1] = rng.uniform(-0.5, 0.5)
    x0[2:4] = rng.uniform(-0.5, 0.5, 2)
    theta = xavier_mlp() if residual else None
    return Problem("bicycle_nn" if residual else "bicycle", dyn, cost, smp, num_rollouts, num_timesteps, 0.02, 1.0,
                   0.0, x0, cost_map=m, nn_theta=theta)


def batch_mix(num_instances: int = 256, seed: int = 7, k_range=(512, 8192), t_range=(50, 150),
              cost_map: np.ndarray | None = None) -> tuple[list[Problem], np.ndarray]:
    """Independent instances for the batched launch: a random 50/50 mix of double integrators and (plain)
    bicycles, sample counts ~ U[k_range] rounded to a multiple of 64, horizons ~ U[t_range], and per-instance
    seeded initial states, goals and noise seeds. All instances share one cost map (returned with the list)."""
    rng = np.random.default_rng(seed)
    m = gaussian_blob_map() if cost_map is None else cost_map
    out = []
    for i in range(num_instances):
        K = int(round(rng.uniform(*k_range) / 64.0)) * 64
        K = int(min(max(K, 64), k_range[1]))
        T = int(rng.integers(t_range[0], t_range[1] + 1))
        if rng.random() < 0.5:
            p = double_integrator_map(num_rollouts=K, num_timesteps=T, colored=False, cost_map=m,
                                      seed=int(1000 + i))
            p.x0 = rng.uniform(-2.0, 2.0, 4).astype(np.float32)
            p.cost.goal[0], p.cost.goal[1] = rng.uniform(-5.0, 5.0, 2)
        else:
            p = bicycle_nn(num_rollouts=K, num_timesteps=T, residual=False, cost_map=m, seed=int(1000 + i))
            x0 = np.zeros(7, np.float32)
            x0[0] = rng.uniform(2.0, 8.0)
            x0[1] = rng.uniform(-np.pi, np.pi)
            x0[2:4] = rng.uniform(-10.0, 10.0, 2)
            p.x0 = x0
            p.cost.speed_target = float(rng.uniform(4.0, 10.0))
        p.name = f"{p.name}#{i}"
        out.append(p)
    return out, m
