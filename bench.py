#!/usr/bin/env python3
"""MPPI benchmark on one MI355X.

Headline workload (default, `--workload batch256`): 256 independent MPPI instances (double integrator / kinematic
bicycle 50/50, K ~ U[512, 8192] in multiples of 64, T ~ U[50, 150], the 1024 x 1024 blob cost map), advanced by one
persistent rollout launch + one combine launch per control call, launch shape from the autotuner. One step is one
host-inclusive computeControl(x, 1) of the whole batch: the 256 states uploaded, the iteration, the 256 new control
sequences downloaded and smoothed on the host. `value` = steps / wall time of exactly --steps such calls (Hz);
rollout_steps_per_s = value * sum_i K_i T_i.

The same line carries, from the same run:
  - the device-resident rate (graph replays with the inputs already in HBM, no host transfer): device_loop_hz;
  - per-kernel durations from HIP events on the launching stream, and the roofline of the dominant kernel;
  - the CPU restatement of the reference's rollout loop (+ weights + update) on the same noise, timed on this host on
    a bounded sample, on one core and on all cores;
  - `configs`: the single-controller configurations (cart-pole K=2048; double integrator + colored noise + map
    K=8192; bicycle + BF16 residual network K=16384), each with its host-inclusive and device-resident control rate,
    its fused kernel's roofline and combine-kernel bandwidth, and its CPU baseline;
  - `learned_vs_plain`: the learned-proposal comparison on the bicycle + network task.
`--workload cartpole|di_colored_map|bicycle_nn` makes that controller configuration the headline instead.
"""
from __future__ import annotations

import argparse
import ctypes as ct
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MPPI control Hz + rollout-steps/sec; learned vs plain cost at equal time"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
# FLOPs of the 9-64-64-4 residual network per sample-step: 2 (9*64 + 64*64 + 64*4)
NN_FLOP_PER_STEP = 2 * (9 * 64 + 64 * 64 + 64 * 4)
PARTIAL_BYTES = 32  # sizeof(Partial): baseline, argmin, three FP64 sums


def host_threads() -> int:
    """Host cores this process may use: the scheduler's affinity set, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fused_alg_bytes(K: int, T: int, C: int) -> int:
    """Algorithmic HBM bytes of one fused iteration over K samples: the Gaussian/colored sampler's write (4 K T C),
    the rollout's read and write of the samples and its costs (8 K T C + 4 K) and the weighted update's read of the
    samples and weights and write of the mean (4 K T C + 4 K + 4 T C) -- the three per-unit figures of the kernels
    the fused kernel replaces."""
    return 16 * K * T * C + 8 * K + 4 * T * C


def rollout_alg_bytes(K: int, T: int, C: int) -> int:
    return 8 * K * T * C + 4 * K


def measured_traffic(workload: str, kernel_prefix: str):
    """HBM bytes per dispatch of the kernel whose name starts with `kernel_prefix`, from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known size), or None."""
    path = ROOT / "profiles" / f"mppi_iteration_traffic_{workload}.json"
    if not path.exists():
        return None
    tab = json.loads(path.read_text())["traffic"]
    hits = [v for k, v in tab.items() if k.startswith(kernel_prefix)]
    return hits[0]["total_bytes"] if len(hits) == 1 else None


def measured_utilisation(workload: str, prefixes):
    """VALU and MFMA utilisation of this workload's kernels (committed rocprofv3 derived counters VALUBusy,
    VALUUtilization, MfmaUtil), or None."""
    path = ROOT / "profiles" / f"utilisation_{workload}.json"
    if not path.exists():
        return None
    tab = json.loads(path.read_text())
    out = {k: {"valu_busy_pct": round(v["VALUBusy"], 1), "valu_lanes_pct": round(v["VALUUtilization"], 1),
               "mfma_busy_pct": round(v["MfmaUtil"], 1)}
           for k, v in sorted(tab.items()) if any(k.startswith(p) for p in prefixes)}
    if not out:
        return None
    out["source"] = f"profiles/utilisation_{workload}.json (rocprofv3 VALUBusy, VALUUtilization, MfmaUtil)"
    return out


def build_problem(name: str):
    from mppi_generic_vendor_amd import capi, workloads

    if name == "di_colored_map":
        return workloads.double_integrator_map(num_rollouts=8192, num_timesteps=100, colored=True,
                                               fft_len_mode=capi.FFT_LEN_2T)
    if name == "cartpole":
        return workloads.cartpole(num_rollouts=2048, num_timesteps=100)
    if name == "bicycle_nn":
        return workloads.bicycle_nn(num_rollouts=16384, num_timesteps=100, precision=capi.NN_BF16)
    raise SystemExit(f"unknown workload {name}")


def samples_for(prob, k: int, mean: np.ndarray) -> np.ndarray:
    """The first k control samples of the GPU sampler's buffer for this problem (same seed, rng offset 0)."""
    import ctypes as ct

    from mppi_generic_vendor_amd import capi
    from mppi_generic_vendor_amd.hip import DeviceBuffer, synchronize

    T, C = prob.num_timesteps, prob.C
    mb = DeviceBuffer.from_numpy(mean)
    out = DeviceBuffer(k * T * C * 4)
    name = "mppi_sample_colored" if prob.sampler.kind == capi.SAMPLER_COLORED else "mppi_sample_gaussian"
    capi.call(name, ct.byref(prob.sampler), k, T, mb.ptr, out.ptr, 1, 0, ct.c_uint64(0), None, None)
    synchronize()
    return out.to_numpy((k, T, C))


def cpu_baseline(items, total_steps: int, seconds: float, threads: int, what: str):
    """The oracle's FP32 rollout loop (the reference's launchCPURolloutKernel restated) + weights + weighted update
    over `items` = [(problem, samples, mean)], repeated until `seconds` have passed (at least once), on `threads`
    host threads; the measured seconds per sample-step give the rate of a full iteration of `total_steps`
    sample-steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_lib as ol

    steps = sum(u.shape[0] * u.shape[1] for _, u, _ in items)
    reps = 0
    t0 = time.perf_counter()
    while True:
        for prob, u, mean in items:
            costs, clamped = ol.rollout_f32(prob, u, mean, nthreads=threads)
            w, eta, *_ = ol.weights(costs, prob.lambda_)
            ol.weighted_update(w, eta, clamped)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds or reps >= 1000:
            break
    per_step = el / (reps * steps)
    return {"value": round(1.0 / (per_step * total_steps), 4), "unit": "Hz", "cores": threads, "kind": "port",
            "sample": f"{what}, {reps} pass(es) of {steps} sample-steps in {el:.1f} s, scaled to the full "
                      f"{total_steps} sample-steps",
            "rollout_steps_per_s": round(1.0 / per_step, 1)}


def cpu_baselines(items, total_steps: int, seconds: float, what: str):
    """Both modes of the reference's CPU loop: all host cores (the object's main fields) and one core."""
    n = host_threads()
    allc = cpu_baseline(items, total_steps, seconds * 0.4, n, what)
    one = cpu_baseline(items, total_steps, seconds * 0.6, 1, what)
    allc["single_core"] = {k: one[k] for k in ("value", "rollout_steps_per_s", "sample")}
    return allc


# ------------------------------------------------------------------------------------------------ batch (C4)
def run_batch(args):
    from mppi_generic_vendor_amd import hip, workloads
    from mppi_generic_vendor_amd.batch import MPPIBatch

    probs, m = workloads.batch_mix(256, seed=7)
    KT = sum(p.num_rollouts * p.num_timesteps for p in probs)
    b = MPPIBatch(probs, cost_map=m)
    if args.batch_shape:
        # a fixed launch shape (profiled runs: the autotuner's timing is perturbed under a profiler)
        b.setConfig(*[int(v) for v in args.batch_shape.split(",")])
        cfg = b.getConfig()
    else:
        cfg = b.autotune(reps=3)
    log(f"batch256: {sum(p.num_rollouts for p in probs)} samples, {KT} sample-steps per call; "
        f"threads={cfg.threads} samples/lane={cfg.samples_per_lane} chunk={cfg.chunk} blocks/CU={cfg.blocks_per_cu}")
    x0 = b.packStates([p.x0 for p in probs])

    def host_calls():
        for _ in range(args.warmup):
            b.computeControl(x0, 1)
        hip.synchronize()
        per = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t1 = time.perf_counter()
            b.computeControl(x0, 1)
            per.append(time.perf_counter() - t1)
        hip.synchronize()
        return time.perf_counter() - t0, per

    # the same calls replayed from a captured hipGraph, for comparison with the direct launches
    b.setGraphReplay(True)
    graph_wall, _ = host_calls()
    b.setGraphReplay(False)
    # timed region of `value`: exactly --steps host-inclusive control calls
    wall, per = host_calls()
    value = args.steps / wall
    log(f"host-inclusive batched computeControl: {wall * 1e3 / args.steps:.4f} ms/call")
    for i in range(len(probs)):
        st = b.getWeightStats(i)
        if not np.isfinite(st.normalizer) or st.normalizer <= 0:
            raise SystemExit(f"batch instance {i} produced an invalid normalizer")
    # device-resident iterations (inputs in HBM, graph replays)
    b.stepDevice(args.warmup)
    hip.synchronize()
    n_dev = max(args.steps, 20)
    t0 = time.perf_counter()
    b.stepDevice(n_dev)
    hip.synchronize()
    dev_wall = time.perf_counter() - t0
    roll_ms, comb_ms = b.timeKernels(args.kernel_reps)
    log(f"device loop {dev_wall * 1e3 / n_dev:.4f} ms/call; kernels: rollout {roll_ms:.4f} ms, combine "
        f"{comb_ms:.4f} ms")
    spl = cfg.samples_per_lane
    tiles = [-(-p.num_rollouts // (64 * spl)) for p in probs]
    alg = sum(fused_alg_bytes(p.num_rollouts, p.num_timesteps, p.C) for p in probs)
    alg_roll = sum(rollout_alg_bytes(p.num_rollouts, p.num_timesteps, p.C) for p in probs)
    comb_bytes = sum(g * (4 * p.num_timesteps * p.C + PARTIAL_BYTES) + 4 * p.num_timesteps * p.C + 40
                     for g, p in zip(tiles, probs))
    kname = f"batch_rollout_kernel<{cfg.threads}, {spl}>"
    achieved = alg / (roll_ms * 1e-3) / 1e9
    roofline = {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic("batch256", kname),
                "algorithmic_bytes": alg, "avg_ms": round(roll_ms, 5),
                "rollout_only": {"algorithmic_bytes": alg_roll,
                                 "frac": round(alg_roll / (roll_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}}
    cpu = None
    if not args.no_cpu:
        sample = probs[::16]
        items = [(p, samples_for(p, p.num_rollouts, np.zeros((p.num_timesteps, p.C), np.float32)),
                  np.zeros((p.num_timesteps, p.C), np.float32)) for p in sample]
        cpu = cpu_baselines(items, KT, args.cpu_seconds,
                            f"{len(sample)} of the 256 instances (every 16th, full K and T)")
    b.close()
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Hz", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded map, states, goals; Philox noise)",
        "config": {"workload": "batch256: 256 MPPI instances (double integrator / bicycle 50/50), "
                               "K~U[512,8192]/64, T~U[50,150], 1024x1024 cost map; one step = one host-inclusive "
                               "batched computeControl (states from host memory, persistent rollout + combine with "
                               "smoothing and constraints, controls back to host memory)",
                   "instances": len(probs), "sample_steps_per_call": KT,
                   "launch_shape": {"threads": cfg.threads, "samples_per_lane": cfg.samples_per_lane,
                                    "chunk": cfg.chunk, "blocks_per_cu": cfg.blocks_per_cu}},
        "rollout_steps_per_s": round(value * KT, 1),
        "instance_controls_per_s": round(value * len(probs), 1),
        "control_hz_host_median": round(1.0 / statistics.median(per), 3),
        "control_hz_host_graph": round(args.steps / graph_wall, 3),
        "device_loop_hz": round(n_dev / dev_wall, 3),
        "device_loop_rollout_steps_per_s": round(n_dev / dev_wall * KT, 1),
        "kernels_ms": {kname: round(roll_ms, 5), "batch_combine_kernel": round(comb_ms, 5)},
        "reduction_kernels": {"batch_combine_kernel": {
            "algorithmic_bytes": comb_bytes, "avg_ms": round(comb_ms, 5),
            "achieved_gbs": round(comb_bytes / (comb_ms * 1e-3) / 1e9, 2)}},
        "roofline": roofline,
        "utilisation": measured_utilisation("batch256", [kname]),
        "cpu_baseline": cpu,
    }
    return line


# ------------------------------------------------------------------------------------------ one controller
def run_controller(name: str, args, steps: int, warmup: int, cpu_seconds: float):
    """Host-inclusive computeControl rate over exactly `steps` calls, the device-resident rate, the fused kernel's
    and the combine kernel's durations and rooflines, and the CPU baseline, for one controller configuration."""
    from mppi_generic_vendor_amd import capi, hip
    from mppi_generic_vendor_amd.controller import VanillaMPPIController

    prob = build_problem(name)
    K, T, C = prob.num_rollouts, prob.num_timesteps, prob.C
    x = prob.x0.copy()

    def host_rate(ctrl):
        for _ in range(warmup):
            ctrl.computeControl(x, 1)
        hip.synchronize()
        per = []
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            ctrl.computeControl(x, 1)
            per.append(time.perf_counter() - t1)
        return time.perf_counter() - t0, per

    # the same calls replayed from a captured hipGraph, for comparison with the direct launches
    gctrl = VanillaMPPIController(prob, num_iters=1, use_graph=True)
    graph_wall, _ = host_rate(gctrl)
    gctrl.close()
    ctrl = VanillaMPPIController(prob, num_iters=1, use_graph=False)
    wall, per = host_rate(ctrl)
    st = ctrl.getWeightStats()
    if not np.isfinite(st.normalizer) or st.normalizer <= 0:
        raise SystemExit(f"{name}: invalid normalizer")
    ctrl.stepDevice(warmup)
    hip.synchronize()
    n_dev = max(steps, 200)
    t0 = time.perf_counter()
    ctrl.stepDevice(n_dev)
    hip.synchronize()
    dev_wall = time.perf_counter() - t0
    roll_ms, comb_ms, G = ctrl.timeKernels(args.kernel_reps)
    ctrl.close()
    log(f"{name}: host-inclusive {wall * 1e3 / steps:.4f} ms/call, device loop {dev_wall * 1e3 / n_dev:.4f} ms/call, "
        f"fused kernel {roll_ms:.4f} ms, combine {comb_ms:.4f} ms over {G} partials")
    kname = "rollout_update_kernel"
    alg = fused_alg_bytes(K, T, C)
    bf16 = prob.dyn.model == capi.MODEL_RACER_DUBINS_NN and prob.dyn.nn_precision == capi.NN_BF16
    if bf16:
        flops = NN_FLOP_PER_STEP * K * T
        ach = flops / (roll_ms * 1e-3) / 1e12
        roofline = {"kernel": kname, "bound": "mfma", "achieved": round(ach, 3), "peak": MFMA_BF16_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(ach / MFMA_BF16_PEAK_TFLOPS, 5),
                    "traffic": measured_traffic(name, kname), "algorithmic_flops": flops, "avg_ms": round(roll_ms, 5)}
    else:
        ach = alg / (roll_ms * 1e-3) / 1e9
        roofline = {"kernel": kname, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": measured_traffic(name, kname),
                    "algorithmic_bytes": alg, "avg_ms": round(roll_ms, 5),
                    "rollout_only": {"algorithmic_bytes": rollout_alg_bytes(K, T, C),
                                     "frac": round(rollout_alg_bytes(K, T, C) / (roll_ms * 1e-3) / 1e9
                                                   / HBM_PEAK_GBS, 5)}}
    comb_bytes = G * (4 * T * C + PARTIAL_BYTES) + 4 * T * C + 40
    cpu = None
    if cpu_seconds > 0:
        mean = np.zeros((T, C), np.float32)
        k_s = min(K, 2048)
        cpu = cpu_baselines([(prob, samples_for(prob, k_s, mean), mean)], K * T, cpu_seconds,
                            f"{k_s} of {K} samples x {T} steps")
    return {
        "workload": f"{name}: K={K}, T={T}, S={prob.S}, C={C}, dt={prob.dt}, num_iters=1",
        "control_hz_host": round(steps / wall, 3),
        "control_hz_host_median": round(1.0 / statistics.median(per), 3),
        "control_hz_host_graph": round(steps / graph_wall, 3),
        "rollout_steps_per_s": round(steps / wall * K * T, 1),
        "device_loop_hz": round(n_dev / dev_wall, 3),
        "device_loop_rollout_steps_per_s": round(n_dev / dev_wall * K * T, 1),
        "kernels_ms": {kname: round(roll_ms, 5), "combine_kernel": round(comb_ms, 5)},
        "roofline": roofline,
        "reduction_kernels": {"combine_kernel": {"partials": G, "algorithmic_bytes": comb_bytes,
                                                 "avg_ms": round(comb_ms, 5),
                                                 "achieved_gbs": round(comb_bytes / (comb_ms * 1e-3) / 1e9, 2)}},
        "utilisation": measured_utilisation(name, [kname]),
        "cpu_baseline": cpu,
        "dtype": "f32+bf16" if bf16 else "f32",
        "_wall": wall,
    }


def learned_comparison(episodes: int, use_graph: bool = False):
    """Learned vs plain MPPI on the bicycle + residual-network task with the committed trained proposal, closed loop
    from held-out starts (executed-trajectory cost, stream and host-inclusive rates per control step):
      - plain MPPI's convergence curve at 8192 samples and 1, 2, 3, 5 iterations;
      - the learned proposal at 2048 and at 256 samples (one iteration, fallback above 0.5 %) beside plain MPPI with
        the same samples and one iteration;
      - the cheapest plain setting within 1 % of the converged cost (8192 x 5), and the learned rate against it;
      - the same closed loop disturbed every 10 control steps (learned.kick): plain 8192 x 1, 8192 x 5, 256 x 1 and
        the learned proposal at 256 samples."""
    from mppi_generic_vendor_amd import learned, train_proposal, workloads

    if episodes <= 0 or not train_proposal.DEFAULT_OUT.exists():
        return None
    net, summary = train_proposal.load()
    m = workloads.gaussian_blob_map()
    keep = ("cost", "hz", "hz_host", "samples", "iterations", "fallback_fraction")

    def ev(**kw):
        r = learned.evaluate(kw.pop("net", None), episodes=episodes, steps=100, seed=99, cost_map=m,
                             use_graph=use_graph, **kw)
        return {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items() if kk in keep}
                for k, v in r.items()}

    curve = {}
    for iters in (1, 2, 3, 5):
        curve[f"8192x{iters}"] = ev(plain_rollouts=8192, plain_iters=iters)["plain"]
    regimes = {}
    for P in (2048, 256):
        r = ev(net=net, plain_rollouts=8192, plain_iters=1, proposal_rollouts=P)
        regimes[str(P)] = {"learned": r["learned"], "plain_same_samples": r["plain_small"]}
        curve[f"{P}x1"] = r["plain_small"]
    # a disturbed closed loop: every 10 control steps the system is kicked (new speed, heading turned, steering reset),
    # so the warm start no longer fits the state right after each kick
    kicked = {}
    for K, iters in ((8192, 1), (8192, 5), (256, 1)):
        kicked[f"{K}x{iters}"] = ev(plain_rollouts=K, plain_iters=iters, kick_every=10)["plain"]
    kicked["learned_256"] = ev(net=net, plain_rollouts=8192, plain_iters=1, proposal_rollouts=256,
                               kick_every=10)["learned"]
    net.close()
    converged = curve["8192x5"]["cost"]
    within = {k: v for k, v in curve.items() if v["cost"] <= converged * 1.01}
    cheapest = max(within.items(), key=lambda kv: kv[1]["hz_host"])
    l256 = regimes["256"]["learned"]
    log(f"learned vs plain: converged plain {converged:.4f}; learned at 256 samples {l256['cost']:.4f} at "
        f"{l256['hz_host']:.0f} Hz host; cheapest plain within 1 %: {cheapest[0]} at {cheapest[1]['hz_host']:.0f} Hz")
    return {"task": "bicycle + BF16 residual MLP, T=100, closed loop from held-out starts", "episodes": episodes,
            "trained_on_pairs": summary.get("pairs"), "plain_convergence": curve, "learned": regimes,
            "converged_cost": converged,
            "cheapest_plain_within_1pct": {"setting": cheapest[0], **cheapest[1]},
            "learned_256_cost_vs_converged": round(l256["cost"] / converged, 6),
            "learned_256_cost_gain_over_same_samples":
                round(1 - l256["cost"] / regimes["256"]["plain_same_samples"]["cost"], 6),
            "learned_256_rate_vs_plain_8192x3_host": round(l256["hz_host"] / curve["8192x3"]["hz_host"], 4),
            "learned_256_rate_vs_cheapest_within_1pct_host": round(l256["hz_host"] / cheapest[1]["hz_host"], 4),
            "disturbed_every_10_steps": kicked}


# -------------------------------------------------------------------------------------- two-system (Tube / RMPPI)
def two_system_problem():
    from mppi_generic_vendor_amd import workloads

    return workloads.double_integrator_map(num_rollouts=8192, num_timesteps=100, colored=False)


def pd_gains(T: int, kp: float = 4.0, kd: float = 2.0) -> np.ndarray:
    """Feedback gains [T][S][C] of a PD law on the double integrator: a_x = -kp e_px - kd e_vx, a_y likewise."""
    g = np.zeros((T, 4, 2), np.float32)
    g[:, 0, 0] = g[:, 1, 1] = -kp
    g[:, 2, 0] = g[:, 3, 1] = -kd
    return g


def run_two_system(args, host_calls: int, cpu_seconds: float):
    """Tube-MPPI and RMPPI on the double integrator over the 1024x1024 64-blob map, K=8192, T=100: host-inclusive control
    rates of both controllers (RMPPI: candidate evaluation + optimisation per call); the durations (HIP events on the
    controller's stream) of the RMPPI controller's fused two-system iteration kernel and its combine launch, with the
    fused kernel's roofline over its algorithmic bytes per launch -- the applied controls and costs of both systems
    written, 2 (4 K T C + 4 K), plus the per-workgroup partials, 2 G (4 T C + 32); the kernel-level two-system rollout
    entry point in RMPPI mode on device buffers (samples read and written, 2 (8 K T C + 4 K)); and the oracle's CPU
    loop."""
    from mppi_generic_vendor_amd import capi, hip
    from mppi_generic_vendor_amd.controller import RobustMPPIController, TubeMPPIController

    prob = two_system_problem()
    K, T, C, S = prob.num_rollouts, prob.num_timesteps, prob.C, prob.S
    gains = pd_gains(T)
    rates = {}
    for name in ("tube", "rmppi"):
        ctrl = TubeMPPIController(prob) if name == "tube" else RobustMPPIController(prob, value_function_threshold=50.0)
        if name == "rmppi":
            ctrl.setFeedbackGains(gains)
        x = prob.x0.copy()

        def call():
            if name == "rmppi":
                ctrl.updateImportanceSamplingControl(x, 1)
            ctrl.computeControl(x, 1)
            if name == "tube":
                ctrl.slideControlSequence(1)

        for _ in range(10):
            call()
        t0 = time.perf_counter()
        for _ in range(host_calls):
            call()
        rates[name] = host_calls / (time.perf_counter() - t0)
        if name == "rmppi":
            fused_ms, comb_ms, G = ctrl.timeKernels(max(args.kernel_reps, 5))
        ctrl.close()
    # the kernel alone, RMPPI mode, on device buffers
    rng = np.random.default_rng(3)
    u = rng.normal(size=(2, K, T, C)).astype(np.float32)
    means = np.zeros((2, T, C), np.float32)
    x0s = np.stack([prob.x0, prob.x0 + np.float32(0.1)]).astype(np.float32)
    from mppi_generic_vendor_amd.hip import DeviceBuffer

    mb_map = DeviceBuffer.from_numpy(prob.cost_map.astype(np.float32))
    cost = capi.CostParams.from_buffer_copy(prob.cost)
    cost.map = mb_map.ptr
    gb, xb = DeviceBuffer.from_numpy(gains), DeviceBuffer.from_numpy(x0s)
    mb, sb, cb = DeviceBuffer.from_numpy(means), DeviceBuffer.from_numpy(u), DeviceBuffer(2 * K * 4)
    tp = capi.TwoSystemParams(capi.TWO_SYSTEM_RMPPI, 0, 50.0, 0, gb.ptr)
    st = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()

    def launch():
        capi.call("mppi_rollout_two_system", ct.byref(prob.dyn), ct.byref(cost), ct.byref(prob.sampler), None,
                  ct.byref(tp), K, T, prob.dt, prob.lambda_, prob.alpha, xb.ptr, mb.ptr, sb.ptr, cb.ptr, st.ptr)

    for _ in range(3):
        launch()
    reps = max(args.kernel_reps, 5)
    e0.record(st)
    for _ in range(reps):
        launch()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    alg = 2 * (8 * K * T * C + 4 * K)
    ach = alg / (ms * 1e-3) / 1e9
    alg_f = 2 * (4 * K * T * C + 4 * K) + 2 * G * (4 * T * C + 32)
    ach_f = alg_f / (fused_ms * 1e-3) / 1e9
    log(f"two-system: Tube {rates['tube']:.1f} Hz, RMPPI {rates['rmppi']:.1f} Hz host-inclusive; RMPPI fused iteration "
        f"kernel {fused_ms:.4f} ms = {ach_f:.1f} GB/s + combine {comb_ms:.4f} ms ({G} partials per system); "
        f"kernel-level rollout {ms:.4f} ms = {ach:.1f} GB/s")
    cpu = None
    if cpu_seconds > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle_lib as ol

        Ks = 512
        sub = np.ascontiguousarray(u[:, :Ks])
        res = {}
        for n in (host_threads(), 1):
            reps_c, t0 = 0, time.perf_counter()
            while True:
                ol.rollout_two_system(prob, sub, means, x0s, 0, 50.0, gains, nthreads=n)
                reps_c += 1
                el = time.perf_counter() - t0
                if el >= cpu_seconds / 2 or reps_c >= 5000:
                    break
            per = el / (reps_c * Ks)
            res[n] = {"value": round(1.0 / (per * K), 4), "unit": "Hz", "cores": n, "kind": "port",
                      "sample": f"RMPPI two-system rollout of {Ks} of the {K} samples (both systems, full T), "
                                f"{reps_c} pass(es) in {el:.1f} s, scaled to K = {K}"}
        cpu = res[host_threads()]
        cpu["single_core"] = res[1]
    mb_map.free()
    return {"workload": "Tube-MPPI / RMPPI, double integrator over the 1024x1024 64-blob map, K=8192, T=100 (two "
                        "systems per sample), PD feedback gains (RMPPI)",
            "tube_control_hz_host": round(rates["tube"], 3), "rmppi_control_hz_host": round(rates["rmppi"], 3),
            "kernels_ms": {"two_system_rollout_kernel<2, 2, 0, true>": round(fused_ms, 5),
                           "combine_kernel (two systems, one launch)": round(comb_ms, 5),
                           "two_system_rollout_kernel<2, 2, 0, false>": round(ms, 5)},
            "roofline": {"kernel": "two_system_rollout_kernel<2, 2, 0, true>", "bound": "hbm",
                         "achieved": round(ach_f, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach_f / HBM_PEAK_GBS, 5), "algorithmic_bytes": alg_f, "avg_ms": round(fused_ms, 5),
                         "partials_per_system": G,
                         "traffic": measured_traffic("tube_di", "two_system_rollout_kernel<2, 2, 0, true>")},
            "kernel_level_rollout": {"kernel": "two_system_rollout_kernel<2, 2, 0, false>", "achieved": round(ach, 2),
                                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "algorithmic_bytes": alg,
                                     "avg_ms": round(ms, 5),
                                     "traffic": measured_traffic("tube_di", "two_system_rollout_kernel<2, 2, 0, false>")},
            "utilisation": measured_utilisation("tube_di", ["two_system_rollout_kernel", "candidate_eval_kernel"]),
            "cpu_baseline_rollout": cpu}


CONTROLLER_WORKLOADS = ("cartpole", "di_colored_map", "bicycle_nn")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="batch256", choices=["batch256", "tube_di", *CONTROLLER_WORKLOADS])
    ap.add_argument("--host-calls", type=int, default=200,
                    help="timed computeControl calls of each configuration in `configs`")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--batch-shape", default="",
                    help="threads,samples_per_lane,chunk of the batched kernel instead of the autotuner's choice")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU-baseline time of the headline workload")
    ap.add_argument("--config-cpu-seconds", type=float, default=3.0,
                    help="CPU-baseline time of each configuration in `configs`")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations")
    ap.add_argument("--learned-episodes", type=int, default=4,
                    help="closed-loop episodes of the learned-proposal comparison (0: skip)")
    ap.add_argument("--learned-graph", action="store_true",
                    help="run the learned-proposal comparison from captured hipGraphs instead of direct launches")
    ap.add_argument("--calibration-mib", type=int, default=256,
                    help="device-to-device copy of this size, a plain copy kernel of known traffic for counter scaling")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every GPU box of this setup has one GPU and the engine has no "
              "multi-GPU path (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from mppi_generic_vendor_amd import capi, hip

    capi.load()
    if hip.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.calibration_mib > 0:
        # a plain copy kernel of known size (reads and writes calibration_mib MiB), for scaling HBM counters
        nb = args.calibration_mib << 20
        a_buf, b_buf = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
        hip.check(hip._load().hipMemcpy(b_buf.ptr, a_buf.ptr, nb, hip.D2D), "calibration copy")
        hip.synchronize()
        a_buf.free()
        b_buf.free()

    if args.workload == "batch256":
        line = run_batch(args)
    elif args.workload == "tube_di":
        r = run_two_system(args, args.steps, 0 if args.no_cpu else args.cpu_seconds)
        line = {"metric": METRIC, "value": r["tube_control_hz_host"], "unit": "Hz", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic (seeded cost map, Philox noise)",
                "config": {"workload": r["workload"] + "; one step = one host-inclusive Tube computeControl(x, 1)"}}
        line.update({k: v for k, v in r.items() if k != "workload"})
    else:
        r = run_controller(args.workload, args, args.steps, args.warmup, 0 if args.no_cpu else args.cpu_seconds)
        prob = build_problem(args.workload)
        KT = prob.num_rollouts * prob.num_timesteps
        line = {
            "metric": METRIC, "value": r["control_hz_host"], "unit": "Hz", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(r["_wall"] * 1e3 / args.steps, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": r["dtype"],
            "data": "synthetic (seeded cost map, Philox noise)",
            "config": {"workload": r["workload"] + "; one step = one host-inclusive computeControl(x, 1)"},
            "sample_steps_per_call": KT,
        }
        line.update({k: v for k, v in r.items() if k not in ("workload", "control_hz_host", "_wall", "dtype")})
    if not args.no_configs:
        line["configs"] = {}
        for name in CONTROLLER_WORKLOADS:
            if name == args.workload:
                continue
            r = run_controller(name, args, args.host_calls, 20, 0 if args.no_cpu else args.config_cpu_seconds)
            r.pop("_wall")
            line["configs"][name] = r
        if args.workload != "tube_di":
            line["two_system"] = run_two_system(args, min(args.host_calls, 100),
                                                0 if args.no_cpu else args.config_cpu_seconds)
    line["learned_vs_plain"] = learned_comparison(args.learned_episodes, args.learned_graph)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
