"""Python host wrapper of the learned proposal network (mppi_proposal_* in include/mppi_amd.h)."""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np

from . import capi
from .hip import DeviceBuffer

HIDDEN = 128
KNOTS = 16


def num_params(S: int, C: int, T: int) -> int:
    n = ct.c_int()
    capi.call("mppi_proposal_num_params", S, C, T, ct.byref(n))
    return n.value


def init_params(S: int, C: int, T: int, seed: int = 7) -> np.ndarray:
    """Xavier-normal hidden layers, zero output layer: the untrained proposal is the base distribution."""
    rng = np.random.default_rng(seed)
    din, dout = S + KNOTS * C, 2 * T * C
    parts = []
    for ni, no in ((din, HIDDEN), (HIDDEN, HIDDEN)):
        parts.append(rng.normal(scale=math.sqrt(2.0 / (ni + no)), size=ni * no))
        parts.append(np.zeros(no))
    parts.append(np.zeros(dout * HIDDEN))
    parts.append(np.zeros(dout))
    theta = np.concatenate(parts).astype(np.float32)
    assert theta.size == num_params(S, C, T)
    return theta


def knot_indices(T: int) -> np.ndarray:
    return np.array([j * (T - 1) // 15 for j in range(KNOTS)])


def network_inputs(x: np.ndarray, prev: np.ndarray) -> np.ndarray:
    """Raw (unnormalized) inputs [N][S + 16 C] from states [N][S] and previous sequences [N][T][C]."""
    x = np.asarray(x, np.float32)
    prev = np.asarray(prev, np.float32)
    T = prev.shape[1]
    k = prev[:, knot_indices(T), :].reshape(prev.shape[0], -1)
    return np.concatenate([x, k], axis=1).astype(np.float32)


class ProposalNet:
    def __init__(self, S: int, C: int, T: int, params: np.ndarray | None = None):
        capi.load()
        self.S, self.C, self.T = S, C, T
        self.din, self.dout = S + KNOTS * C, 2 * T * C
        self.P = num_params(S, C, T)
        h = ct.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        capi.call("mppi_proposal_create", S, C, T, None if p is None else p.ctypes.data, ct.byref(h))
        self._h = h
        self.shift = np.zeros(self.din, np.float32)
        self.scale = np.ones(self.din, np.float32)

    @property
    def handle(self):
        return self._h

    def params(self) -> np.ndarray:
        out = np.empty(self.P, np.float32)
        capi.call("mppi_proposal_get_params", self._h, out.ctypes.data)
        return out

    def set_params(self, theta: np.ndarray):
        t = np.ascontiguousarray(theta, np.float32)
        capi.call("mppi_proposal_set_params", self._h, t.ctypes.data)

    def set_normalization(self, shift: np.ndarray, scale: np.ndarray):
        self.shift = np.ascontiguousarray(shift, np.float32)
        self.scale = np.ascontiguousarray(scale, np.float32)
        capi.call("mppi_proposal_set_input_normalization", self._h, self.shift.ctypes.data, self.scale.ctypes.data)

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        xb = DeviceBuffer.from_numpy(X)
        yb = DeviceBuffer(X.shape[0] * self.dout * 4)
        capi.call("mppi_proposal_forward", self._h, xb.ptr, X.shape[0], yb.ptr, None)
        return yb.to_numpy((X.shape[0], self.dout))

    def infer(self, x0: np.ndarray, nominal: np.ndarray, sigma_base: np.ndarray, fused: bool = True):
        """(dmu [T][C], sigma_q [T][C]) for one state, as the controller computes them."""
        xb = DeviceBuffer.from_numpy(np.ascontiguousarray(x0, np.float32))
        nb = DeviceBuffer.from_numpy(np.ascontiguousarray(nominal, np.float32))
        sb = DeviceBuffer.from_numpy(np.ascontiguousarray(sigma_base, np.float32))
        db = DeviceBuffer(self.T * self.C * 4)
        gb = DeviceBuffer(self.T * self.C * 4)
        capi.call("mppi_proposal_infer", self._h, xb.ptr, nb.ptr, sb.ptr, db.ptr, gb.ptr, 1 if fused else 0, None)
        return db.to_numpy((self.T, self.C)), gb.to_numpy((self.T, self.C))

    def loss_grad(self, X: np.ndarray, R: np.ndarray):
        xb = DeviceBuffer.from_numpy(np.ascontiguousarray(X, np.float32))
        rb = DeviceBuffer.from_numpy(np.ascontiguousarray(R, np.float32))
        gb = DeviceBuffer(self.P * 4)
        loss = ct.c_double()
        capi.call("mppi_proposal_loss_grad", self._h, xb.ptr, rb.ptr, X.shape[0], gb.ptr, ct.byref(loss), None)
        return loss.value, gb.to_numpy((self.P,))

    def train(self, X: np.ndarray, R: np.ndarray, epochs: int, batch: int = 512, lr: float = 1e-3, seed: int = 0,
              log=None) -> list[float]:
        """Minibatch Adam over device-resident data; returns the mean loss of every epoch."""
        X = np.ascontiguousarray(X, np.float32)
        R = np.ascontiguousarray(R, np.float32)
        N = X.shape[0]
        xb = DeviceBuffer.from_numpy(X)
        rb = DeviceBuffer.from_numpy(R)
        rng = np.random.default_rng(seed)
        hist = []
        xs = DeviceBuffer(batch * self.din * 4)
        rs = DeviceBuffer(batch * self.T * self.C * 4)
        loss = ct.c_double()
        for ep in range(epochs):
            perm = rng.permutation(N)
            tot, nb = 0.0, 0
            Xp = X[perm]
            Rp = R[perm]
            xb.copy_from(Xp)
            rb.copy_from(Rp)
            for b0 in range(0, N - batch + 1, batch):
                xp = xb.ptr + b0 * self.din * 4
                rp = rb.ptr + b0 * self.T * self.C * 4
                last = b0 + 2 * batch > N
                capi.call("mppi_proposal_train_step", self._h, xp, rp, batch, lr,
                          ct.byref(loss) if last else None, None)
                nb += 1
            hist.append(loss.value)
            if log:
                log(f"[proposal] epoch {ep}: loss {loss.value:.5f}")
        return hist

    def close(self):
        if self._h:
            capi.call("mppi_proposal_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
