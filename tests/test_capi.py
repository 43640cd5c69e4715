"""Host-side checks of the C-ABI (no GPU needed): every declared symbol is exported, the ctypes mirrors have the C
layout, defaults match the reference's parameter structs, and argument validation returns error codes."""
import ctypes as ct
import re
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest

from mppi_generic_vendor_amd import capi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mppi_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:const\s+)?[a-z_0-9]+\s*\*?\s*(mppi_[a-z_0-9]+)\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("mppi_rollout", "mppi_sample_gaussian", "mppi_sample_colored", "mppi_compute_weights",
                 "mppi_weighted_update", "mppi_controller_create", "mppi_controller_compute_control"):
        assert must in names


def test_library_exports_every_declared_symbol(engine):
    missing = [n for n in declared_functions() if not hasattr(engine, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (mppi_\w+)", out))
    assert set(declared_functions()) <= exported
    assert set(capi.exported_symbols()) <= exported


def test_ctypes_layout_matches_c(tmp_path):
    structs = {"mppi_dynamics_params": capi.DynamicsParams, "mppi_cost_params": capi.CostParams,
               "mppi_sampler_params": capi.SamplerParams, "mppi_weight_stats": capi.WeightStats,
               "mppi_free_energy": capi.FreeEnergy, "mppi_controller_desc": capi.ControllerDesc,
               "mppi_instance_desc": capi.InstanceDesc, "mppi_batch_config": capi.BatchConfig}
    fields = []
    for cname, py in structs.items():
        for f in py._fields_:
            fname = f[0] if f[0] != "lambda_" else "lambda"
            fields.append((cname, fname))
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "mppi_amd.h"', "int main(void){"]
    for cname in structs:
        src.append(f'printf("%s size %zu\\n", "{cname}", sizeof({cname}));')
    for cname, f in fields:
        src.append(f'printf("%s %s %zu\\n", "{cname}", "{f}", offsetof({cname}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(HEADER.parent), str(c), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    got = {}
    for ln in lines:
        parts = ln.split()
        got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in structs.items():
        assert got[(cname, "size")] == ct.sizeof(py), cname
        for f in py._fields_:
            fname = f[0] if f[0] != "lambda_" else "lambda"
            assert got[(cname, fname)] == getattr(py, f[0]).offset, (cname, fname)


def test_defaults_follow_reference_params(engine):
    cp = capi.dynamics_defaults(capi.MODEL_CARTPOLE)
    assert (cp.state_dim, cp.control_dim, cp.output_dim) == (4, 1, 4)
    assert (cp.cart_mass, cp.pole_mass, cp.pole_length) == (1.0, 1.0, 1.0)
    assert cp.gravity == pytest.approx(9.81)
    assert cp.control_min[0] == -np.finfo(np.float32).max and cp.control_max[0] == np.finfo(np.float32).max
    rd = capi.dynamics_defaults(capi.MODEL_RACER_DUBINS)
    assert (rd.state_dim, rd.control_dim) == (7, 2)
    assert rd.c_t == pytest.approx(1.3) and rd.c_0 == pytest.approx(4.9) and rd.wheel_base == pytest.approx(0.3)
    assert rd.steer_angle_scale == pytest.approx(-9.1) and rd.max_brake_rate_pos == pytest.approx(0.33)
    nn = capi.dynamics_defaults(capi.MODEL_RACER_DUBINS_NN)
    assert list(nn.nn_layers) == [9, 64, 64, 4] and nn.nn_num_params == 10 * 64 + 65 * 64 + 65 * 4
    cc = capi.cost_defaults(capi.COST_CARTPOLE_QUADRATIC, cp)
    assert list(cc.coeff)[:4] == [1000.0, 100.0, 2000.0, 100.0]
    assert cc.goal[2] == pytest.approx(np.pi)
    sp = capi.sampler_defaults(capi.SAMPLER_COLORED, 2)
    assert sp.pure_noise_trajectories_percentage == pytest.approx(0.01)
    assert sp.offset_decay_rate == pytest.approx(0.97) and sp.std_dev_decay == 1.0
    n = ct.c_int32()
    layers = (ct.c_int32 * 4)(6, 32, 32, 4)
    capi.call("mppi_nn_num_params", 4, layers, ct.byref(n))
    assert n.value == 1412  # the reference FNN test network (tests/nn_helpers/fnn_helper_test.cu:409)


def test_argument_validation_returns_codes(engine):
    dyn = capi.dynamics_defaults(capi.MODEL_CARTPOLE)
    cost = capi.cost_defaults(capi.COST_CARTPOLE_QUADRATIC, dyn)
    smp = capi.sampler_defaults(capi.SAMPLER_GAUSSIAN, 1)
    P = ct.c_void_p(16)
    st = engine.mppi_rollout(ct.byref(dyn), ct.byref(cost), ct.byref(smp), 0, 100, 0.02, 1.0, 0.0, P, P, P, P, None)
    assert st == capi.MPPI_E_INVALID and b"num_rollouts" in engine.mppi_last_error()
    st = engine.mppi_rollout(ct.byref(dyn), ct.byref(cost), ct.byref(smp), 64, 1000, 0.02, 1.0, 0.0, P, P, P, P, None)
    assert st == capi.MPPI_E_SIZE
    bad = capi.DynamicsParams.from_buffer_copy(dyn)
    bad.model = 99
    st = engine.mppi_rollout(ct.byref(bad), ct.byref(cost), ct.byref(smp), 64, 10, 0.02, 1.0, 0.0, P, P, P, P, None)
    assert st == capi.MPPI_E_UNSUPPORTED
    smp2 = capi.sampler_defaults(capi.SAMPLER_GAUSSIAN, 2)
    st = engine.mppi_rollout(ct.byref(dyn), ct.byref(cost), ct.byref(smp2), 64, 10, 0.02, 1.0, 0.0, P, P, P, P, None)
    assert st == capi.MPPI_E_INVALID
    st = engine.mppi_rollout(ct.byref(dyn), ct.byref(cost), ct.byref(smp), 64, 10, 0.02, 1.0, 0.0, None, P, P, P, None)
    assert st == capi.MPPI_E_INVALID
    st = engine.mppi_compute_weights(P, 10, -1.0, P, None)
    assert st == capi.MPPI_E_INVALID
    smpc = capi.sampler_defaults(capi.SAMPLER_COLORED, 1)
    st = engine.mppi_sample_colored(ct.byref(smpc), 64, 10, P, P, 20, 0, 0, None, None)
    assert st == capi.MPPI_E_INVALID  # stride outside [0, 2T)
    d = capi.ControllerDesc()
    d.dynamics, d.cost, d.sampler = dyn, cost, smp
    d.num_rollouts, d.num_timesteps, d.dt, d.lambda_, d.num_iters = 128, 50, 0.02, 0.0, 1
    h = ct.c_void_p()
    st = engine.mppi_controller_create(ct.byref(d), ct.byref(h))
    assert st == capi.MPPI_E_INVALID and not h.value
    # the colored controller's parameters: kind, Tsallis gamma > 0 with r != 1, leash distances >= 0
    d.lambda_ = 1.0
    for kind, gamma, r, leash in ((7, 0.0, 2.0, 0.0), (1, -1.0, 2.0, 0.0), (1, 1.0, 1.0, 0.0), (1, 0.0, 2.0, -1.0)):
        d.controller_kind, d.tsallis_gamma, d.tsallis_r, d.state_leash_dist[0] = kind, gamma, r, leash
        st = engine.mppi_controller_create(ct.byref(d), ct.byref(h))
        assert st == capi.MPPI_E_INVALID and not h.value, (kind, gamma, r, leash)
    assert engine.mppi_compute_weights_tsallis(P, 10, 1.0, 0.0, 2.0, P, None) == capi.MPPI_E_INVALID
    assert engine.mppi_compute_weights_tsallis(P, 10, 1.0, 1.0, 1.0, P, None) == capi.MPPI_E_INVALID
    assert engine.mppi_compute_weights_tsallis(P, 10, 1.0, -2.0, 2.0, P, None) == capi.MPPI_E_INVALID
    assert engine.mppi_status_string(capi.MPPI_E_SIZE) == b"size exceeds a compiled limit"
