"""Pins of the learned-proposal oracle (oracle/proposal.cpp): GELU and its derivative against closed forms, the FP64
backward pass against central finite differences of the FP64 loss, Adam's first step against its closed form, and
the input downsampling against hand-listed knots."""
import math

import numpy as np
import pytest

import oracle_lib as ol


def test_gelu_closed_forms(oracle):
    assert oracle.oracle_gelu(0.0) == 0.0
    # 0.5 (1 + erf(1/sqrt 2)) = Phi(1) = 0.8413447460685429
    assert oracle.oracle_gelu(1.0) == pytest.approx(0.8413447460685429, abs=1e-15)
    assert oracle.oracle_gelu(-1.0) == pytest.approx(-1.0 * (1 - 0.8413447460685429), abs=1e-15)
    assert oracle.oracle_dgelu(0.0) == pytest.approx(0.5, abs=1e-15)
    for z in (-3.0, -0.7, 0.2, 1.3, 4.0):
        h = 1e-6
        fd = (oracle.oracle_gelu(z + h) - oracle.oracle_gelu(z - h)) / (2 * h)
        assert oracle.oracle_dgelu(z) == pytest.approx(fd, abs=1e-8)


def _net(S, C, T, seed):
    P = ol.load().oracle_proposal_num_params(S, C, T)
    rng = np.random.default_rng(seed)
    din, dout = S + 16 * C, 2 * T * C
    theta = np.zeros(P, np.float32)
    off = 0
    for ni, no in ((din, 128), (128, 128), (128, dout)):
        theta[off:off + ni * no] = rng.normal(scale=math.sqrt(2.0 / (ni + no)), size=ni * no)
        off += ni * no
        theta[off:off + no] = rng.normal(scale=0.1, size=no)
        off += no
    assert off == P
    return theta


def test_forward_matches_numpy(oracle):
    """The FP64 forward against an independent numpy formulation: X W1^T + b1 -> GELU (scipy erf) -> W2^T + b2 ->
    GELU -> W3^T + b3, with W_l read as [out][in] row-major blocks followed by b_l (the FNN layout). Non-square
    layers make a transposed weight block fail, and the distinct biases catch a dropped or misplaced bias."""
    from scipy import special

    S, C, T, B = 7, 2, 11, 6
    theta = _net(S, C, T, 5)
    din, dout = S + 16 * C, 2 * T * C
    rng = np.random.default_rng(3)
    X = rng.normal(size=(B, din)).astype(np.float32)
    th = theta.astype(np.float64)
    off = 0
    h = X.astype(np.float64)
    for li, (ni, no) in enumerate(((din, 128), (128, 128), (128, dout))):
        W = th[off:off + ni * no].reshape(no, ni)
        off += ni * no
        b = th[off:off + no]
        off += no
        h = h @ W.T + b
        if li < 2:
            h = 0.5 * h * (1.0 + special.erf(h / math.sqrt(2.0)))
    got = ol.proposal_forward(S, C, T, theta, X)
    np.testing.assert_allclose(got, h, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences(oracle):
    S, C, T, B = 3, 1, 5, 4
    theta = _net(S, C, T, 0)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(B, S + 16 * C)).astype(np.float32)
    R = rng.normal(size=(B, T * C)).astype(np.float32)
    loss, g = ol.proposal_loss_grad(S, C, T, theta, X, R)
    assert np.isfinite(loss)
    P = theta.size
    din = S + 16 * C
    # parameters from every block: W1, b1, W2, b2, W3 (mean and log-std rows), b3
    idx = list(rng.choice(128 * din, 8, replace=False)) + [128 * din + 3, 128 * din + 77]
    w2 = 128 * din + 128
    idx += list(w2 + rng.choice(128 * 128, 8, replace=False)) + [w2 + 128 * 128 + 5]
    w3 = w2 + 128 * 128 + 128
    idx += list(w3 + rng.choice(2 * T * C * 128, 10, replace=False)) + [P - 1, P - 2 * T * C]
    for i in idx:
        h = 2e-3
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        hp = float(tp[i]) - float(theta[i])
        hm = float(theta[i]) - float(tm[i])
        lp, _ = ol.proposal_loss_grad(S, C, T, tp, X, R)
        lm, _ = ol.proposal_loss_grad(S, C, T, tm, X, R)
        fd = (lp - lm) / (hp + hm)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=2e-7), f"param {i}"


def test_adam_first_step(oracle):
    P = 7
    th = np.linspace(-1, 1, P).astype(np.float32)
    g = np.array([0.3, -2.0, 1e-3, 0.0, 5.0, -0.01, 0.7], np.float32)
    m = np.zeros(P, np.float32)
    v = np.zeros(P, np.float32)
    t0 = th.copy()
    oracle.oracle_adam(th.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data, P, 1e-2, 0.9, 0.999, 1e-8, 1)
    # first step: m_hat = g, v_hat = g^2 -> theta -= lr g / (|g| + eps)
    want = t0 - 1e-2 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(th, want, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)


def test_input_downsampling_knots(oracle):
    S, C, T = 7, 2, 100
    x = np.arange(1, 8, dtype=np.float32)
    prev = np.array([[1000.0 * c + t for c in range(C)] for t in range(T)], np.float32)
    shift = np.zeros(S + 16 * C, np.float32)
    scale = np.ones(S + 16 * C, np.float32)
    out = np.zeros(S + 16 * C, np.float32)
    oracle.oracle_proposal_input(S, C, T, x.ctypes.data, prev.ctypes.data, shift.ctypes.data, scale.ctypes.data,
                                 out.ctypes.data)
    knots = [0, 6, 13, 19, 26, 33, 39, 46, 52, 59, 66, 72, 79, 85, 92, 99]
    want = list(range(1, 8)) + [v for t in knots for v in (t, 1000 + t)]
    np.testing.assert_array_equal(out, np.array(want, np.float32))
    shift[:] = 1.0
    scale[:] = 0.5
    oracle.oracle_proposal_input(S, C, T, x.ctypes.data, prev.ctypes.data, shift.ctypes.data, scale.ctypes.data,
                                 out.ctypes.data)
    np.testing.assert_array_equal(out, (np.array(want, np.float32) - 1.0) * 0.5)


def test_importance_correction_closed_forms(oracle):
    """No shift gives zero correction for any u and spread; one step by hand: u = 1, mu = 0, dmu = 0.5, sp = 1,
    lambda = 2: 2 (1 - 0.25) / 2 = 0.75; and the term is linear in u: (2 d0 dmu - dmu^2) / (2 sp^2)."""
    rng = np.random.default_rng(2)
    K, T, C = 5, 7, 2
    u = rng.normal(size=(K, T, C)).astype(np.float32)
    mean = rng.normal(size=(T, C)).astype(np.float32)
    sp = np.array([0.7, 1.3], np.float32)
    z = ol.is_cost(u, mean, np.zeros((T, C)), rng.uniform(0.1, 2.0, (T, C)), sp, 1.5)
    np.testing.assert_allclose(z, 0.0, atol=1e-12)
    v = ol.is_cost(np.ones((1, 1, 1)), np.zeros((1, 1)), np.full((1, 1), 0.5), np.full((1, 1), 0.5),
                   np.ones(1, np.float32), 2.0)
    assert v[0] == pytest.approx(0.75, abs=1e-12)
    dmu = rng.normal(size=(T, C)).astype(np.float32)
    got = ol.is_cost(u, mean, dmu, np.ones((T, C)), sp, 1.0)
    d0 = u.astype(np.float64) - mean
    want = ((2 * d0 * dmu - dmu.astype(np.float64) ** 2) / (2 * sp.astype(np.float64) ** 2)).sum(axis=(1, 2))
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)
