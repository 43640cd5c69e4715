"""Two-system rollouts on the GPU (mppi_rollout_two_system: Tube-MPPI's z = 2 rollout,
controllers/Tube-MPPI/tube_mppi_controller.cu:158-236, and RMPPI's nominal + feedback-tracking real system,
core/rmppi_kernels.cu:359-798) against the oracle's two_system_loop on the same samples. Bars: costs within 1e-5
relative (Tube) / 1e-4 relative (RMPPI: the feedback term amplifies the last-bit state differences of FMA-contracted
device arithmetic); applied controls bit-exact where no feedback enters (both Tube systems, the RMPPI nominal), within
1e-5 of the control scale for the RMPPI real system. Ragged K (a partial last workgroup), every workgroup shape (64,
32, 16 samples: small K and the T = 256 horizon), both nominal indices, the cart-pole, the double integrator over a
cost map and the kinematic bicycle."""
import ctypes as ct

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import banner, device_problem, gpu_rollout, rel_err
from mppi_generic_vendor_amd import capi, workloads
from mppi_generic_vendor_amd.hip import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu


def gpu_two_system(prob, u, means, x0s, nom=-1, vft=0.0, gains=None, std1=None):
    K, T, C = u.shape[1], u.shape[2], u.shape[3]
    dyn, cost, sp0, keep = device_problem(prob, with_sampler=True)
    sp1 = capi.SamplerParams.from_buffer_copy(sp0)
    if std1 is not None:
        for c in range(C):
            sp1.std_dev[c] = float(std1[c])
    gb = DeviceBuffer.from_numpy(np.ascontiguousarray(gains, np.float32)) if gains is not None else None
    tp = capi.TwoSystemParams(capi.TWO_SYSTEM_TUBE if nom < 0 else capi.TWO_SYSTEM_RMPPI, max(nom, 0), vft, 0,
                              gb.ptr if gb else None)
    xb = DeviceBuffer.from_numpy(np.ascontiguousarray(x0s, np.float32))
    mb = DeviceBuffer.from_numpy(np.ascontiguousarray(means, np.float32))
    sb = DeviceBuffer.from_numpy(np.ascontiguousarray(u, np.float32))
    cb = DeviceBuffer(2 * K * 4)
    capi.call("mppi_rollout_two_system", ct.byref(dyn), ct.byref(cost), ct.byref(sp0), ct.byref(sp1), ct.byref(tp),
              K, T, prob.dt, prob.lambda_, prob.alpha, xb.ptr, mb.ptr, sb.ptr, cb.ptr, None)
    synchronize()
    return cb.to_numpy((2, K)), sb.to_numpy((2, K, T, C))


def _problem(name, K, T):
    if name == "cartpole":
        p = workloads.cartpole(num_rollouts=K, num_timesteps=T)
    elif name == "di_map":
        p = workloads.double_integrator_map(num_rollouts=K, num_timesteps=T, colored=False,
                                            cost_map=workloads.gaussian_blob_map(256, 256, blobs=16, seed=7))
    else:
        p = workloads.bicycle_nn(num_rollouts=K, num_timesteps=T, residual=False,
                                 cost_map=workloads.gaussian_blob_map(256, 256, blobs=16, seed=7))
    p.sampler.control_cost_coeff[0] = 0.3
    if p.C > 1:
        p.sampler.control_cost_coeff[1] = 0.6
    p.sampler.pure_noise_trajectories_percentage = 0.05
    return p


def _inputs(prob, seed, scale=1.0):
    K, T, C, S = prob.num_rollouts, prob.num_timesteps, prob.C, prob.S
    rng = np.random.default_rng(seed)
    u = (rng.normal(size=(2, K, T, C)) * scale).astype(np.float32)
    means = (rng.normal(size=(2, T, C)) * 0.3 * scale).astype(np.float32)
    x0s = np.stack([prob.x0, prob.x0 + rng.normal(size=S).astype(np.float32) * 0.2]).astype(np.float32)
    return u, means, x0s


# K = 9000 fills the CUs with 64-sample workgroups (one lane per sample carrying both systems); the smaller K run
# 16- or 32-sample workgroups, whose producer lanes carry one system each
@pytest.mark.parametrize("name,K,T", [("di_map", 2000, 60), ("cartpole", 300, 100), ("bicycle", 1000, 50),
                                      ("di_map", 512, 256), ("di_map", 9000, 40)])
def test_tube_two_system_matches_oracle(hip, engine, name, K, T):
    prob = _problem(name, K, T)
    banner(f"two-system Tube rollout {name} K={K} T={T}")
    u, means, x0s = _inputs(prob, 11)
    std1 = np.array([prob.sampler.std_dev[c] * 1.5 for c in range(prob.C)], np.float32)
    got, got_u = gpu_two_system(prob, u, means, x0s, -1, std1=std1)
    want, want_u = ol.rollout_two_system(prob, u, means, x0s, -1, std_dev_1=std1)
    e = rel_err(got, want)
    print(f"[two] tube {name}: cost rel err {e:.2e}")
    assert e < 1e-5
    assert np.array_equal(got_u, want_u)
    # system 0 alone is the single-system rollout
    single, _ = gpu_rollout(prob, u[0], means[0], x0=x0s[0])
    assert rel_err(got[0], single) < 1e-6


@pytest.mark.parametrize("name,K,T,nom", [("di_map", 2000, 60, 0), ("cartpole", 700, 100, 1), ("bicycle", 1000, 50, 0),
                                          ("bicycle", 96, 40, 1), ("di_map", 256, 256, 1), ("di_map", 9000, 40, 1),
                                          ("bicycle", 9000, 30, 0)])
def test_rmppi_two_system_matches_oracle(hip, engine, name, K, T, nom):
    prob = _problem(name, K, T)
    banner(f"two-system RMPPI rollout {name} K={K} T={T} nominal={nom}")
    u, means, x0s = _inputs(prob, 12 + nom)
    rng = np.random.default_rng(5)
    gains = (rng.normal(size=(T, prob.S, prob.C)) * 0.3).astype(np.float32)
    # the value-function threshold at the median tracking cost: both branches of min(track_r, threshold) occur
    vft = float(np.median(ol.rollout_two_system(prob, u, means, x0s, nom, 0.0, gains, parts=True)[2][:, 1]))
    got, got_u = gpu_two_system(prob, u, means, x0s, nom, vft=vft, gains=gains)
    want, want_u, parts = ol.rollout_two_system(prob, u, means, x0s, nom, vft=vft, gains=gains, parts=True)
    e = rel_err(got, want)
    real = 1 - nom
    du = float(np.max(np.abs(got_u[real] - want_u[real]))) / max(1.0, float(np.max(np.abs(want_u[real]))))
    clipped = float(np.mean(parts[:, 1] > vft))
    print(f"[two] rmppi {name} nominal={nom}: cost rel err {e:.2e}, real controls {du:.2e}, "
          f"track > threshold in {clipped:.0%} of samples")
    assert e < 1e-4
    assert np.array_equal(got_u[nom], want_u[nom])
    assert du < 1e-5
    assert not np.array_equal(got_u[real], got_u[nom])  # the feedback acted
    assert 0.2 < clipped < 0.8


def test_rmppi_without_feedback_equals_mppi_rollout(hip, engine):
    """ValidateCombinedRMPPIRolloutKernelAgainstMPPIRollout (tests/mppi_core/rmppi_kernel_tests.cu:506-591): no
    gains, both systems from the same state over the same samples and mean -> both costs are the MPPI rollout's
    (there within 1e-3; here 1e-5)."""
    prob = _problem("di_map", 2048, 50)
    banner("two-system RMPPI without feedback vs mppi_rollout")
    u, means, x0s = _inputs(prob, 21)
    u[1], means[1], x0s[1] = u[0], means[0], x0s[0]
    got, got_u = gpu_two_system(prob, u, means, x0s, 0, vft=20.0, gains=None)
    single, single_u = gpu_rollout(prob, u[0], means[0], x0=x0s[0])
    for d in range(2):
        assert rel_err(got[d], single) < 1e-5
        assert np.array_equal(got_u[d], single_u)
