"""The proposal network's GPU kernels against the oracle (oracle/proposal.cpp): forward outputs, loss and the
gradient of every parameter (BF16 operand rounding emulated in the oracle, FP64 otherwise), one Adam step, and a
training run that must lower the loss."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import banner
from mppi_generic_vendor_amd.proposal import ProposalNet, init_params

pytestmark = pytest.mark.gpu


def _random_params(S, C, T, seed):
    th = init_params(S, C, T, seed)
    rng = np.random.default_rng(seed + 1)
    din, dout = S + 16 * C, 2 * T * C
    w3 = din * 128 + 128 + 128 * 128 + 128
    th[w3:w3 + dout * 128] = rng.normal(scale=0.05, size=dout * 128)
    th[-dout:] = rng.normal(scale=0.1, size=dout)
    return th


def _nrel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("S,C,T,B", [(7, 2, 100, 64), (3, 1, 9, 5), (7, 2, 100, 300)])
def test_forward_and_gradients_match_oracle(hip, engine, S, C, T, B):
    banner(f"proposal forward/backward S={S} C={C} T={T} B={B}")
    th = _random_params(S, C, T, 3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(B, S + 16 * C)).astype(np.float32)
    R = rng.normal(scale=0.5, size=(B, T * C)).astype(np.float32)
    net = ProposalNet(S, C, T, th)
    Y = net.forward(X)
    Yb = ol.proposal_forward(S, C, T, th, X, bf16=True)
    Y64 = ol.proposal_forward(S, C, T, th, X, bf16=False)
    ef, ef64 = _nrel(Y, Yb), _nrel(Y, Y64)
    loss, g = net.loss_grad(X, R)
    lb, gb = ol.proposal_loss_grad(S, C, T, th, X, R, bf16=True)
    l64, g64 = ol.proposal_loss_grad(S, C, T, th, X, R, bf16=False)
    eg, eg64 = _nrel(g, gb), _nrel(g, g64)
    print(f"[proposal] forward err vs bf16-oracle {ef:.2e} vs fp64 {ef64:.2e}; loss {loss:.6f} vs {lb:.6f} "
          f"(fp64 {l64:.6f}); grad err vs bf16-oracle {eg:.2e} vs fp64 {eg64:.2e}")
    assert ef < 1e-3
    assert loss == pytest.approx(lb, rel=1e-4)
    assert eg < 5e-3
    # every parameter block individually
    din, dout = S + 16 * C, 2 * T * C
    offs = np.cumsum([0, din * 128, 128, 128 * 128, 128, dout * 128, dout])
    for a, b in zip(offs[:-1], offs[1:]):
        assert _nrel(g[a:b], gb[a:b]) < 1e-2, (a, b)
    net.close()


def test_adam_step_matches_oracle(hip, engine):
    banner("proposal Adam step")
    S, C, T, B = 7, 2, 100, 128
    th = _random_params(S, C, T, 5)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(B, S + 16 * C)).astype(np.float32)
    R = rng.normal(scale=0.5, size=(B, T * C)).astype(np.float32)
    net = ProposalNet(S, C, T, th)
    _, g = net.loss_grad(X, R)
    net.train(X, R, epochs=1, batch=B, lr=1e-3)
    got = net.params()
    want = th.copy()
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    ol.load().oracle_adam(want.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data, th.size, 1e-3, 0.9, 0.999,
                          1e-8, 1)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
    net.close()


def test_training_lowers_the_loss(hip, engine):
    banner("proposal training on a synthetic regression")
    S, C, T = 7, 2, 100
    rng = np.random.default_rng(9)
    N = 4096
    X = rng.normal(size=(N, S + 16 * C)).astype(np.float32)
    Wt = rng.normal(scale=0.3, size=(S + 16 * C, T * C))
    R = np.tanh(X @ Wt).astype(np.float32) + rng.normal(scale=0.05, size=(N, T * C)).astype(np.float32)
    net = ProposalNet(S, C, T, init_params(S, C, T, 1))
    l0, _ = net.loss_grad(X[:512], R[:512])
    hist = net.train(X, R, epochs=15, batch=256, lr=2e-3)
    l1, _ = net.loss_grad(X[:512], R[:512])
    print(f"[proposal] loss {l0:.4f} -> {l1:.4f} ({hist[0]:.4f} .. {hist[-1]:.4f})")
    assert l1 < 0.5 * l0
    net.close()
