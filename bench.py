#!/usr/bin/env python3
"""MPPI benchmark on one MI355X.

Workload (default): 2-D double integrator with power-law (colored) noise and a 1024x1024 cost map, K = 8192 samples,
T = 100, C = 2 -- one step is the full GPU work of one computeControl call with num_iters = 1 (colored sampling,
rollout with cost-map lookups, weights, weighted update). The timed region replays device-resident steps (x0, the
nominal sequence and the map already in HBM) from a captured hipGraph, bracketed by device synchronisations.

Also reported: the host-inclusive control rate (median wall time of computeControl(x, 1), which adds the 16 B x0
upload, the 800 B control download, SG smoothing and the host state propagation), per-kernel HIP-event durations
with the roofline of the dominant kernel, and the CPU restatement of the reference rollout loop timed on this host.
"""
from __future__ import annotations

import argparse
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
# kernels behind each timed entry point, as named in rocprofv3 traces
KERNEL_NAMES = {"sample_colored": ["colored_mfma_kernel"], "sample_gaussian": ["sample_gaussian_kernel"],
                "rollout": ["rollout_kernel"], "weights": ["weights_kernel"],
                "weighted_update": ["update_partial_kernel", "update_columns_kernel"]}


def measured_traffic(workload: str, entry: str):
    """HBM bytes per launch of `entry` from the committed rocprofv3 counter summary of this workload (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size), or None when there is none."""
    path = ROOT / "profiles" / f"mppi_iteration_traffic_{workload}.json"
    if not path.exists():
        return None
    tab = json.loads(path.read_text())["traffic"]
    total = 0
    for prefix in KERNEL_NAMES[entry]:
        hits = [v for k, v in tab.items() if k.startswith(prefix + "<") or k == prefix]
        if len(hits) != 1:
            return None
        total += hits[0]["total_bytes"]
    return total


def measured_utilisation(workload: str, prefixes):
    """VALU and MFMA utilisation of this workload's kernels whose names start with one of `prefixes`, from the
    committed rocprofv3 counter summary (derived counters VALUBusy: VALU issue time as a percentage of the CUs'
    capacity; VALUUtilization: active lanes per VALU instruction; MfmaUtil: matrix-core busy cycles as a percentage
    of SIMD cycles), or None when there is none."""
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


def time_kernels(prob, reps: int):
    """Average duration of each kernel of one iteration, HIP events on the stream the kernel is launched on."""
    import ctypes as ct

    from mppi_generic_vendor_amd import capi
    from mppi_generic_vendor_amd.hip import DeviceBuffer, Event, Stream

    K, T, C = prob.num_rollouts, prob.num_timesteps, prob.C
    s = Stream()
    dyn = capi.DynamicsParams.from_buffer_copy(prob.dyn)
    cost = capi.CostParams.from_buffer_copy(prob.cost)
    keep = []
    if prob.cost_map is not None:
        mb = DeviceBuffer.from_numpy(prob.cost_map)
        keep.append(mb)
        cost.map = mb.ptr
    if prob.nn_theta is not None:
        tb = DeviceBuffer.from_numpy(prob.nn_theta)
        keep.append(tb)
        dyn.nn_theta = tb.ptr
    x0 = DeviceBuffer.from_numpy(prob.x0)
    mean = DeviceBuffer.zeros((T, C))
    samples = DeviceBuffer(K * T * C * 4)
    costs = DeviceBuffer(K * 4)
    stats = DeviceBuffer(64)
    out = DeviceBuffer(T * C * 4)
    sampler_name = "sample_colored" if prob.sampler.kind == capi.SAMPLER_COLORED else "sample_gaussian"

    def sample(i):
        capi.call("mppi_" + sampler_name, ct.byref(prob.sampler), K, T, mean.ptr, samples.ptr, 1, 0,
                  ct.c_uint64(i), None, s.ptr)

    def rollout(i):
        capi.call("mppi_rollout", ct.byref(dyn), ct.byref(cost), ct.byref(prob.sampler), K, T, prob.dt, prob.lambda_,
                  prob.alpha, x0.ptr, mean.ptr, samples.ptr, costs.ptr, s.ptr)

    def weights(i):
        capi.call("mppi_compute_weights", costs.ptr, K, prob.lambda_, stats.ptr, s.ptr)

    def update(i):
        capi.call("mppi_weighted_update", costs.ptr, stats.ptr, samples.ptr, K, T, C, out.ptr, s.ptr)

    res = {}
    sample(0)
    for name, fn in ((sampler_name, sample), ("rollout", rollout), ("weights", weights), ("weighted_update", update)):
        # back-to-back launches between two events on the launching stream (the rollout's clamping is idempotent,
        # so repeated rollouts of one sample buffer do identical work)
        for i in range(3):
            fn(i)
        s.synchronize()
        e0, e1 = Event(), Event()
        e0.record(s)
        for i in range(reps):
            fn(i)
        e1.record(s)
        e1.synchronize()
        res[name] = e0.elapsed_ms(e1) / reps
    return res


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


def cpu_baseline(prob, seconds: float, threads: int):
    """The CPU restatement of the reference rollout loop (+ weights and update) on a bounded sample of the same
    workload and noise, timed on this host; scaled to iterations/s of the full K."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_lib as ol

    K, T, C = prob.num_rollouts, prob.num_timesteps, prob.C
    mean = np.zeros((T, C), np.float32)
    k_probe = 256
    u = samples_for(prob, k_probe, mean)
    t0 = time.perf_counter()
    costs, clamped = ol.rollout_f32(prob, u, mean, nthreads=threads)
    w, eta, *_ = ol.weights(costs, prob.lambda_)
    ol.weighted_update(w, eta, clamped)
    dt_probe = max(time.perf_counter() - t0, 1e-6)
    k_sample = int(min(K, max(k_probe, k_probe * seconds / dt_probe / 3)))
    u = samples_for(prob, k_sample, mean)
    reps = 0
    t0 = time.perf_counter()
    while True:
        costs, clamped = ol.rollout_f32(prob, u, mean, nthreads=threads)
        w, eta, *_ = ol.weights(costs, prob.lambda_)
        ol.weighted_update(w, eta, clamped)
        reps += 1
        el = time.perf_counter() - t0
        if el > seconds * 0.5 or reps >= 50:
            break
    per_iter_full = el / reps * (K / k_sample)
    return {
        "value": 1.0 / per_iter_full,
        "unit": "Hz",
        "cores": threads,
        "kind": "port",
        "sample": f"{k_sample} of {K} samples x {T} steps, {reps} reps, rollout+weights+update, scaled to K={K}",
        "rollout_steps_per_s": k_sample * T * reps / el,
    }


def run_batch(args):
    """256 independent instances (50/50 double integrator / bicycle, K ~ U[512, 8192] in multiples of 64,
    T ~ U[50, 150]) advanced by one persistent launch + one combine launch per iteration, shape from the autotuner."""
    from mppi_generic_vendor_amd import hip, workloads
    from mppi_generic_vendor_amd.batch import MPPIBatch

    probs, m = workloads.batch_mix(256, seed=7)
    ks = sum(p.num_rollouts * p.num_timesteps for p in probs)
    b = MPPIBatch(probs, cost_map=m)
    cfg = b.autotune(reps=3)
    print(f"[bench] batch256: {sum(p.num_rollouts for p in probs)} samples, {ks} sample-steps per iteration; "
          f"autotuned threads={cfg.threads} samples/lane={cfg.samples_per_lane} chunk={cfg.chunk} "
          f"blocks/CU={cfg.blocks_per_cu} ({cfg.iteration_us:.1f} us)", file=sys.stderr, flush=True)
    x0 = [p.x0 for p in probs]
    for _ in range(3):
        b.computeControl(x0, 1)
    host_ms = []
    for _ in range(max(5, args.host_calls // 10)):
        t0 = time.perf_counter()
        b.computeControl(x0, 1)
        host_ms.append((time.perf_counter() - t0) * 1e3)
    b.stepDevice(args.warmup)
    hip.synchronize()
    stream = hip.Stream.__new__(hip.Stream)
    stream.ptr = b.stream()
    e0, e1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    b.stepDevice(args.steps)
    e1.record(stream)
    hip.synchronize()
    wall = time.perf_counter() - t0
    stream.ptr = None
    ev_ms = e0.elapsed_ms(e1) / args.steps
    value = args.steps / wall
    for i in range(len(probs)):
        st = b.getWeightStats(i)
        if not np.isfinite(st.normalizer) or st.normalizer <= 0:
            raise SystemExit(f"batch instance {i} produced an invalid normalizer")
    b.close()
    # algorithmic HBM bytes of one iteration: clamped samples written (4 K T C), costs (4 K), and the per-tile
    # partial rows (T C floats per 64-sample tile) written and read back by the combine
    alg = sum(4 * p.num_rollouts * p.num_timesteps * p.C + 4 * p.num_rollouts
              + 2 * 4 * (p.num_rollouts // 64) * p.num_timesteps * p.C for p in probs)
    achieved = alg / (ev_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Hz", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded map, states, goals)",
        "config": {"workload": "batch256: 256 instances (DI / bicycle 50/50), K~U[512,8192]/64, T~U[50,150], "
                               "one persistent launch + combine per iteration, hipGraph replay",
                   "launch_shape": {"threads": cfg.threads, "samples_per_lane": cfg.samples_per_lane,
                                    "chunk": cfg.chunk, "blocks_per_cu": cfg.blocks_per_cu}},
        "rollout_steps_per_s": round(value * ks, 1),
        "control_hz_host": round(1e3 / statistics.median(host_ms), 3),
        "roofline": {"kernel": "batch_rollout_kernel+batch_combine_kernel", "bound": "hbm",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "algorithmic_bytes": alg,
                     "avg_ms": round(ev_ms, 5)},
        "utilisation": measured_utilisation(
            "batch256", [f"batch_rollout_kernel<{cfg.threads}, {cfg.samples_per_lane}>"]),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="di_colored_map",
                    choices=["di_colored_map", "cartpole", "bicycle_nn", "batch256"])
    ap.add_argument("--host-calls", type=int, default=200)
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--learned-episodes", type=int, default=4,
                    help="closed-loop episodes of the learned-proposal comparison (0: skip)")
    ap.add_argument("--learned-eager", action="store_true",
                    help="run the learned-proposal comparison with eager launches instead of graph replay "
                         "(for kernel-trace profiles of that path)")
    ap.add_argument("--calibration-mib", type=int, default=256,
                    help="device-to-device copy of this size, a plain copy kernel of known traffic for counter scaling")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (pass --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from mppi_generic_vendor_amd import capi, hip
    from mppi_generic_vendor_amd.controller import VanillaMPPIController

    capi.load()
    if hip.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.workload == "batch256":
        run_batch(args)
        return
    prob = build_problem(args.workload)
    K, T, C = prob.num_rollouts, prob.num_timesteps, prob.C
    print(f"[bench] workload {args.workload} K={K} T={T} C={C}", file=sys.stderr, flush=True)

    ctrl = VanillaMPPIController(prob, num_iters=1, use_graph=True)
    x = prob.x0.copy()
    # host-inclusive control rate: computeControl(x, 1) from the x0 upload to the smoothed host control
    for _ in range(args.warmup):
        ctrl.computeControl(x, 1)
    host_ms = []
    for _ in range(args.host_calls):
        t0 = time.perf_counter()
        ctrl.computeControl(x, 1)
        host_ms.append((time.perf_counter() - t0) * 1e3)
    control_hz_host = 1e3 / statistics.median(host_ms)
    print(f"[bench] host-inclusive computeControl median {statistics.median(host_ms):.4f} ms", file=sys.stderr,
          flush=True)

    # device-resident steps (the timed region of `value`)
    ctrl.stepDevice(args.warmup)
    hip.synchronize()
    stream = hip.Stream.__new__(hip.Stream)
    stream.ptr = ctrl.stream()
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    ctrl.stepDevice(args.steps)
    e1.record(stream)
    hip.synchronize()
    wall = time.perf_counter() - t0
    stream.ptr = None
    ev_ms = e0.elapsed_ms(e1)
    value = args.steps / wall
    ms_per_step = wall * 1e3 / args.steps
    print(f"[bench] device-resident {ms_per_step:.4f} ms/step (events {ev_ms / args.steps:.4f})", file=sys.stderr,
          flush=True)
    st = ctrl.getWeightStats()
    if not np.isfinite(st.normalizer) or st.normalizer <= 0:
        raise SystemExit("benchmark iteration produced an invalid normalizer")
    ctrl.close()

    if args.calibration_mib > 0:
        # a plain copy kernel of known size (reads and writes calibration_mib MiB), for scaling HBM counters
        nb = args.calibration_mib << 20
        a_buf, b_buf = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
        hip.check(hip._load().hipMemcpy(b_buf.ptr, a_buf.ptr, nb, hip.D2D), "calibration copy")
        hip.synchronize()
        a_buf.free()
        b_buf.free()
    kern = time_kernels(prob, args.kernel_reps)
    print(f"[bench] kernel ms {kern}", file=sys.stderr, flush=True)
    # algorithmic bytes per launch (4 B per FP32): sampler writes K*T*C, rollout reads + writes K*T*C and writes K
    alg = {
        "sample_colored": 4 * K * T * C, "sample_gaussian": 4 * K * T * C,
        "rollout": 8 * K * T * C + 4 * K,
        "weights": 8 * K + 16,
        "weighted_update": 4 * K * T * C + 4 * K + 4 * T * C,
    }
    dom = max((k for k in kern if k != "weights"), key=lambda k: kern[k])
    if "sample_colored" in kern:
        # the colored stage is the larger part of the controller's fused kernel (about 27 K of 50 K cycles per
        # workgroup), and the kernel-level sampler and rollout times lie within noise of each other
        dom = "sample_colored"
    if prob.dyn.model == capi.MODEL_RACER_DUBINS_NN and dom == "rollout":
        # the network on the matrix cores bounds the rollout: algorithmic FLOPs per launch over its duration
        flops = NN_FLOP_PER_STEP * K * T
        achieved = flops / (kern[dom] * 1e-3) / 1e12
        roofline = {"kernel": "rollout_kernel", "entry_point": "mppi_rollout", "bound": "mfma",
                    "achieved": round(achieved, 3), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 5),
                    "traffic": measured_traffic(args.workload, dom), "algorithmic_flops": flops,
                    "avg_ms": round(kern[dom], 5)}
    else:
        achieved = alg[dom] / (kern[dom] * 1e-3) / 1e9
        roofline = {"kernel": "+".join(KERNEL_NAMES[dom]), "entry_point": "mppi_" + dom, "bound": "hbm",
                    "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.workload, dom),
                    "algorithmic_bytes": alg[dom], "avg_ms": round(kern[dom], 5)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, args.cpu_seconds, 1)
    learned_cmp = (learned_comparison(args.learned_episodes, not args.learned_eager)
                   if args.workload == "di_colored_map" else None)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Hz",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16" if prob.dyn.model == capi.MODEL_RACER_DUBINS_NN and prob.dyn.nn_precision == capi.NN_BF16
                 else "f32",
        "data": "synthetic (seeded cost map, Philox noise)",
        "config": {"workload": f"{args.workload}: K={K} samples, T={T}, S={prob.S}, C={C}, dt={prob.dt}, "
                               f"num_iters=1, device-resident graph replay"},
        "rollout_steps_per_s": round(value * K * T, 1),
        "control_hz_host": round(control_hz_host, 3),
        "kernels_ms": {k: round(v, 5) for k, v in kern.items()},
        "roofline": roofline,
        # the controller iteration's fused kernel and the kernel behind the roofline entry point
        "utilisation": measured_utilisation(args.workload, ["rollout_update_kernel<"] + KERNEL_NAMES[dom]),
        "cpu_baseline": cpu,
        "learned_vs_plain": learned_cmp,
    }
    print(json.dumps(line), flush=True)


def learned_comparison(episodes: int, use_graph: bool = True):
    """Learned vs plain MPPI on the bicycle + residual-network task with the committed trained proposal: closed-loop
    executed-trajectory cost from held-out starts, and the stream time (from the event before the call's first
    launch to the one recorded after its last) and host wall time per control step, of plain MPPI (8192 samples,
    3 iterations), the learned proposal (2048 samples, 1 iteration, fallback above 0.5 %), and plain MPPI with the
    proposal's samples and one iteration (about the learned step's time)."""
    from mppi_generic_vendor_amd import learned, train_proposal, workloads

    if episodes <= 0 or not train_proposal.DEFAULT_OUT.exists():
        return None
    net, summary = train_proposal.load()
    res = learned.evaluate(net, episodes=episodes, steps=100, seed=99, proposal_rollouts=2048,
                           cost_map=workloads.gaussian_blob_map(), use_graph=use_graph)
    net.close()
    for v in res.values():
        v.pop("costs", None)
    p, l, s = res["plain"], res["learned"], res["plain_small"]
    print(f"[bench] learned vs plain: cost {l['cost']:.4f} vs {p['cost']:.4f} (same samples, one iteration: "
          f"{s['cost']:.4f}), {l['hz']:.0f} vs {p['hz']:.0f} control Hz (stream time), {l['hz_host']:.0f} vs "
          f"{p['hz_host']:.0f} (host wall time)", file=sys.stderr, flush=True)
    return {"task": "bicycle + BF16 residual MLP, K=8192 T=100, closed loop from held-out starts",
            "episodes": episodes, "trained_on_pairs": summary.get("pairs"),
            "plain": p, "learned": l, "plain_same_samples": s,
            "cost_ratio_learned_to_plain": round(l["cost"] / p["cost"], 6),
            "rate_ratio_learned_to_plain": round(l["hz"] / p["hz"], 4),
            "rate_ratio_host_learned_to_plain": round(l["hz_host"] / p["hz_host"], 4),
            "cost_gain_over_same_samples": round((s["cost"] - l["cost"]) / s["cost"], 6)}


if __name__ == "__main__":
    main()
