"""What the trained proposal network gains over the method without it, on the bicycle + residual-network task.

The network shipped in mppi_generic_vendor_amd/data/proposal_bicycle_nn.npz was trained by the project's kernels
(python -m mppi_generic_vendor_amd.train_proposal: 200k pairs from closed-loop plain MPPI with 8192 samples and 3
iterations). Closed-loop episodes from held-out starts compare the executed-trajectory cost of (a) plain MPPI at that
converged setting, (b) the learned proposal with 2048 samples and one iteration (fallback to (a) above a 0.5 % cost
tolerance), (c) plain MPPI with the same 2048 samples and one iteration -- the method without the network. A second
test runs the data-generation and training pipeline on a small set and checks the held-out loss falls.
"""
import numpy as np
import pytest

from gpu_util import banner

pytestmark = pytest.mark.gpu


def test_trained_proposal_beats_plain_at_equal_samples(hip, engine):
    from mppi_generic_vendor_amd import learned, train_proposal, workloads

    banner("learned proposal gain (shipped network)")
    net, summary = train_proposal.load()
    m = workloads.gaussian_blob_map()
    res = learned.evaluate(net, episodes=4, steps=100, seed=5, proposal_rollouts=2048, cost_map=m, log=print)
    plain, lrn, small = res["plain"]["cost"], res["learned"]["cost"], res["plain_small"]["cost"]
    print(f"[learned] cost: plain {plain:.4f}, learned {lrn:.4f} (fallback {res['learned']['fallback_fraction']:.2f})"
          f", plain with the same samples {small:.4f}; gain {(small - lrn) / small * 100:.3f} %, "
          f"gap to plain {(lrn - plain) / plain * 100:.3f} %; rate {res['learned']['hz']:.0f} vs "
          f"{res['plain']['hz']:.0f} Hz")
    assert lrn < small
    assert lrn <= plain * 1.01
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
