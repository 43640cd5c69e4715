"""What the trained proposal network gains over the method without it, on the bicycle + residual-network task.

The network shipped in mppi_generic_vendor_amd/data/proposal_bicycle_nn.npz was trained by the project's kernels
(python -m mppi_generic_vendor_amd.train_proposal: 200k pairs from closed-loop plain MPPI with 8192 samples and 3
iterations). Closed-loop episodes from held-out starts compare the executed-trajectory cost of (a) plain MPPI at that
converged setting, (b) the learned proposal with 2048 samples and one iteration, (c) plain MPPI with the same 2048
samples and one iteration -- the method without the network -- on cold starts (the cost of the returned sequence)
and in closed loop. A second test runs the data-generation and training pipeline on a small set and checks the
held-out loss falls.
"""
import numpy as np
import pytest

from gpu_util import banner

pytestmark = pytest.mark.gpu


def test_trained_proposal_beats_plain_at_equal_samples(hip, engine):
    """Cold starts (zero warm start) from 48 seeded states: the sequence one learned iteration returns costs less
    than the one a plain iteration with the same 2048 samples returns, and lies within 1 % of plain MPPI at 8192
    samples and 3 iterations. Then closed loop from 4 held-out starts: within 1 % of plain at its converged
    setting."""
    from mppi_generic_vendor_amd import learned, train_proposal, workloads

    banner("learned proposal gain (shipped network)")
    net, summary = train_proposal.load()
    m = workloads.gaussian_blob_map()
    q = learned.cold_start_quality(net, starts=48, seed=11, proposal_rollouts=2048, cost_map=m, log=print)
    plain, lrn, small = q["plain"]["cost"], q["learned"]["cost"], q["plain_small"]["cost"]
    print(f"[learned] cold-start sequence cost: plain {plain:.4f}, learned {lrn:.4f}, same samples {small:.4f}: "
          f"gain {(small - lrn) / small * 100:.3f} %, gap to plain {(lrn - plain) / plain * 100:.3f} %")
    assert lrn < small
    assert lrn <= plain * 1.01
    res = learned.evaluate(net, episodes=4, steps=100, seed=7, proposal_rollouts=2048, cost_map=m, log=print)
    print(f"[learned] closed loop: plain {res['plain']['cost']:.4f}, learned {res['learned']['cost']:.4f}, same "
          f"samples {res['plain_small']['cost']:.4f}; rate {res['learned']['hz']:.0f} vs {res['plain']['hz']:.0f} Hz")
    assert res["learned"]["cost"] <= res["plain"]["cost"] * 1.01
    net.close()


def test_learned_at_256_samples_beats_same_budget_plain(hip, engine):
    """Closed loop from 6 held-out starts with 256 proposal samples (1/32 of plain's): the learned proposal's executed
    cost is at least 0.5 % below plain MPPI with the same 256 samples and one iteration (measured: 1.2 %), within 1 %
    of plain MPPI converged at 8192 samples and 5 iterations (measured: 0.96 % below it), and its host-inclusive rate is
    at least 2x that of plain at 8192 samples and 3 iterations (measured: 2.9x)."""
    from mppi_generic_vendor_amd import learned, train_proposal, workloads

    banner("learned proposal at 256 samples vs plain")
    net, _ = train_proposal.load()
    m = workloads.gaussian_blob_map()
    conv = learned.evaluate(None, episodes=6, steps=100, seed=99, plain_rollouts=8192, plain_iters=5, cost_map=m)
    ref3 = learned.evaluate(None, episodes=6, steps=100, seed=99, plain_rollouts=8192, plain_iters=3, cost_map=m)
    res = learned.evaluate(net, episodes=6, steps=100, seed=99, plain_rollouts=8192, plain_iters=1,
                           proposal_rollouts=256, cost_map=m)
    lrn, small, c5 = res["learned"], res["plain_small"], conv["plain"]
    print(f"[learned] 256 samples: learned {lrn['cost']:.4f} ({lrn['hz_host']:.0f} Hz host), same budget "
          f"{small['cost']:.4f}, plain 8192x5 {c5['cost']:.4f}, plain 8192x3 {ref3['plain']['hz_host']:.0f} Hz host")
    assert lrn["cost"] <= small["cost"] * (1 - 0.005)
    assert lrn["cost"] <= c5["cost"] * 1.01
    assert lrn["hz_host"] >= 2.0 * ref3["plain"]["hz_host"]
    net.close()


def test_pipeline_trains_on_mppi_pairs(hip, engine):
    from mppi_generic_vendor_amd import learned, workloads
    from mppi_generic_vendor_amd.proposal import ProposalNet, init_params

    banner("learned proposal pipeline (4000 pairs)")
    m = workloads.gaussian_blob_map()
    ds = learned.generate_pairs(4000, seed=3, num_rollouts=8192, iters=3, cost_map=m, log=print)
    prob = learned.task(8192, cost_map=m)
    sigma = np.array([prob.sampler.std_dev[c] for c in range(prob.C)], np.float32)
    assert ds.x.shape == (4000, 7) and ds.prev.shape == (4000, 100, 2) and np.all(np.isfinite(ds.opt))
    net = learned.train(learned.Dataset(ds.x[:3600], ds.prev[:3600], ds.opt[:3600]), sigma, epochs=10)
    hold = learned.Dataset(ds.x[3600:], ds.prev[3600:], ds.opt[3600:])
    X = ((hold.inputs() - net.shift) * net.scale).astype(np.float32)
    R = hold.targets(sigma)
    fresh = ProposalNet(7, 2, 100, init_params(7, 2, 100, 7))
    l0, _ = fresh.loss_grad(X, R)
    l1, _ = net.loss_grad(X, R)
    print(f"[learned] held-out loss {l0:.4f} (untrained) -> {l1:.4f} (trained)")
    assert l1 < l0
    net.close()
    fresh.close()
