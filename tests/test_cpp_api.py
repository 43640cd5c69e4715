"""The C++ plugin API (include/mppi_amd/mppi.hpp) compiles against the C-ABI library with a plain host compiler, and
its host methods agree with the oracle; the closed-loop program runs on the GPU."""
import ctypes as ct
import subprocess

import numpy as np
import pytest

from cpp_api import ROOT, build_program

PREBUILT = ROOT / "tests" / "cpp" / "_build" / "cartpole_closed_loop"


def test_cpp_api_compiles_with_gcc(tmp_path, engine):
    for name in ("cartpole_closed_loop", "plant_loop", "tracking_loop", "colored_loop", "map_io", "tube_loop"):
        assert build_program(name, tmp_path).exists()


def test_host_step_matches_oracle(engine, oracle):
    from mppi_generic_vendor_amd import capi

    rng = np.random.default_rng(0)
    for model in (capi.MODEL_CARTPOLE, capi.MODEL_DOUBLE_INTEGRATOR, capi.MODEL_RACER_DUBINS):
        dyn = capi.dynamics_defaults(model)
        S, C = dyn.state_dim, dyn.control_dim
        for _ in range(20):
            x = rng.normal(size=S).astype(np.float32)
            u = rng.uniform(-1, 1, size=C).astype(np.float32)
            xn, xd, y = (np.zeros(S, np.float32) for _ in range(3))
            capi.call("mppi_host_step", ct.byref(dyn), x.ctypes.data, u.ctypes.data, 0.02, xn.ctypes.data,
                      xd.ctypes.data, y.ctypes.data)
            xn2, xd2, y2 = (np.zeros(8, np.float32) for _ in range(3))
            oracle.oracle_step_f32(ct.byref(dyn), x.ctypes.data, u.ctypes.data, 0.02, xn2.ctypes.data,
                                   xd2.ctypes.data, y2.ctypes.data)
            np.testing.assert_allclose(xn, xn2[:S], rtol=2e-6, atol=1e-6)


def test_host_step_nn_matches_oracle(engine, oracle):
    """Bicycle + residual network, FP32 and BF16 network arithmetic (the host twin of the matrix-core path)."""
    from mppi_generic_vendor_amd import capi, workloads

    rng = np.random.default_rng(1)
    theta = workloads.xavier_mlp([9, 64, 64, 4])
    for precision in (capi.NN_FP32, capi.NN_BF16):
        dyn = capi.dynamics_defaults(capi.MODEL_RACER_DUBINS_NN)
        dyn.nn_precision = precision
        dyn.nn_theta = theta.ctypes.data
        for _ in range(20):
            x = rng.normal(size=7).astype(np.float32)
            u = rng.uniform(-1, 1, size=2).astype(np.float32)
            xn = np.zeros(7, np.float32)
            capi.call("mppi_host_step", ct.byref(dyn), x.ctypes.data, u.ctypes.data, 0.02, xn.ctypes.data, None,
                      None)
            xn2, xd2, y2 = (np.zeros(8, np.float32) for _ in range(3))
            oracle.oracle_step_f32(ct.byref(dyn), x.ctypes.data, u.ctypes.data, 0.02, xn2.ctypes.data,
                                   xd2.ctypes.data, y2.ctypes.data)
            np.testing.assert_allclose(xn, xn2[:7], rtol=2e-6, atol=1e-6)


@pytest.mark.gpu
def test_cpp_closed_loop_swing_up(tmp_path, hip, engine):
    exe = PREBUILT if PREBUILT.exists() else build_program("cartpole_closed_loop", tmp_path)
    print("[gpu] C++ cartpole closed loop", flush=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr


PLANT = ROOT / "tests" / "cpp" / "_build" / "plant_loop"


@pytest.mark.gpu
def test_cpp_plant_loop(tmp_path, hip, engine):
    """BasePlant (include/mppi_amd/plant.hpp) around the cart-pole controller: interpolated controls between
    solutions, swing-up, a cost-parameter swap mid-run, the threaded runControlLoop with a dynamics-parameter swap
    queued from the robot thread (tests/cpp/plant_loop.cpp prints each check)."""
    exe = PLANT if PLANT.exists() else build_program("plant_loop", tmp_path)
    print("[gpu] C++ plant loop", flush=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr


TRACKING = ROOT / "tests" / "cpp" / "_build" / "tracking_loop"


@pytest.mark.gpu
def test_cpp_goal_trajectory_tracking(tmp_path, hip, engine):
    """QuadraticCostTrajectory with setCurrentTime advanced every step and per-step standard deviations
    (tests/cpp/tracking_loop.cpp): the double integrator tracks a point moving on the unit circle, and does not when
    current_time is held at 0."""
    exe = TRACKING if TRACKING.exists() else build_program("tracking_loop", tmp_path)
    print("[gpu] C++ goal-trajectory tracking", flush=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr


COLORED = ROOT / "tests" / "cpp" / "_build" / "colored_loop"


@pytest.mark.gpu
def test_cpp_colored_controller(tmp_path, hip, engine):
    """ColoredMPPIController (tests/cpp/colored_loop.cpp): with normExp weights and no leash it reproduces the
    vanilla controller bit for bit; with Tsallis weights and the leash it drives the double integrator to its goal."""
    exe = COLORED if COLORED.exists() else build_program("colored_loop", tmp_path)
    print("[gpu] C++ colored controller", flush=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr


MAP_IO = ROOT / "tests" / "cpp" / "_build" / "map_io"


@pytest.mark.gpu
def test_cpp_weight_and_map_io(tmp_path, hip, engine):
    """RacerDubinsNN::loadParams and QuadraticMapCost::loadTrackData from .npz files numpy writes here, and in-place
    updateTexture under a running controller (tests/cpp/map_io.cpp)."""
    from test_io import _reference_track_map

    d, off = {}, 0
    for i, (nin, nout) in enumerate(zip([9, 64, 64], [64, 64, 4]), start=1):
        d[f"dynamics_W{i}"] = np.arange(off, off + nin * nout, dtype=np.float64).reshape(nout, nin)
        off += nin * nout
        d[f"dynamics_b{i}"] = np.arange(off, off + nout, dtype=np.float64)
        off += nout
    np.savez_compressed(tmp_path / "net.npz", **d)
    np.savez(tmp_path / "track.npz", **_reference_track_map())
    exe = MAP_IO if MAP_IO.exists() else build_program("map_io", tmp_path)
    print("[gpu] C++ weight and map I/O", flush=True)
    res = subprocess.run([str(exe), str(tmp_path / "net.npz"), str(tmp_path / "track.npz")], capture_output=True,
                         text=True, timeout=240)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr


def test_host_running_cost_matches_oracle(engine, oracle):
    from mppi_generic_vendor_amd import capi, workloads

    m = workloads.gaussian_blob_map(128, 96, blobs=8, seed=3)
    prob = workloads.bicycle_nn(num_rollouts=64, num_timesteps=10, residual=False, cost_map=m)
    cost = capi.CostParams.from_buffer_copy(prob.cost)
    cost.map = m.ctypes.data
    rng = np.random.default_rng(4)
    for _ in range(50):
        y = rng.normal(scale=4.0, size=7).astype(np.float32)
        got = ct.c_float()
        capi.call("mppi_host_running_cost", ct.byref(cost), y.ctypes.data, ct.byref(got))
        want = oracle.oracle_running_cost_f32(ct.byref(cost), m.ctypes.data, y.ctypes.data, 0)
        assert got.value == pytest.approx(want, rel=2e-6, abs=1e-6)


TUBE = ROOT / "tests" / "cpp" / "_build" / "tube_loop"


@pytest.mark.gpu
def test_cpp_tube_and_rmppi_loops(tmp_path, hip, engine):
    """TubeMPPIController and RobustMPPIController (include/mppi_amd/mppi.hpp) in disturbed closed loops on the double
    integrator: goal reached, nominal replacement seen (Tube), PD feedback law and a bounded tube (RMPPI);
    tests/cpp/tube_loop.cpp prints each check."""
    exe = TUBE if TUBE.exists() else build_program("tube_loop", tmp_path)
    print("[gpu] C++ Tube-MPPI / RMPPI loops", flush=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
