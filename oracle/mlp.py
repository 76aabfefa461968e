"""CPU restatement of the learned-proposal network (test infrastructure only).

The reference has no neural network: its MCMC proposals are random walks and
adaptive Metropolis (src/MCMC.jl:393-449, 1306-1633).  This module restates,
in numpy, the network that libmessi trains on the device so that the device
forward pass, loss, gradients and Adam step can be checked value by value:

  x (din) -> tanh(x W1 + b1) (dh) -> tanh(. W2 + b2) (dh) -> . W3 + b3 (2 din)
  output = (mu, ell);  loss = mean over rows of
           sum_d ell_d + 0.5 ((t_d - x_d - mu_d) exp(-ell_d))^2

With ``emulate=True`` every operand that enters a product on the device is
rounded to bf16 (round to nearest even) at the same places, and products are
accumulated in float64 then stored as float32, as the device stores its FP32
accumulators.  With ``emulate=False`` the arithmetic is plain float64 and the
gradient is the exact derivative of the loss, which the tests pin by central
finite differences.
"""
from __future__ import annotations

import numpy as np


def ceil16(x: int) -> int:
    return (x + 15) // 16 * 16


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest bf16 (ties to even); returns float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


class Layout:
    """Offsets of the flat parameter vector (the device's layout)."""

    def __init__(self, din: int, dh: int):
        self.din, self.dh = din, dh
        self.D0, self.D3 = ceil16(din), ceil16(2 * din)
        D0, H, D3 = self.D0, dh, self.D3
        self.oW1 = 0
        self.ob1 = self.oW1 + D0 * H
        self.oW2 = self.ob1 + H
        self.ob2 = self.oW2 + H * H
        self.oW3 = self.ob2 + H
        self.ob3 = self.oW3 + H * D3
        self.n = self.ob3 + D3

    def unpack(self, w):
        D0, H, D3 = self.D0, self.dh, self.D3
        w = np.asarray(w)
        return (w[self.oW1:self.ob1].reshape(D0, H), w[self.ob1:self.oW2],
                w[self.oW2:self.ob2].reshape(H, H), w[self.ob2:self.oW3],
                w[self.oW3:self.ob3].reshape(H, D3), w[self.ob3:self.n])

    def mask(self) -> np.ndarray:
        """1 where a parameter is live, 0 for padding."""
        m = np.zeros(self.n, dtype=bool)
        m[self.oW1:self.ob1] = (np.arange(self.D0)[:, None] < self.din).repeat(self.dh, 1).reshape(-1)
        m[self.ob1:self.oW3] = True
        m[self.oW3:self.ob3] = (np.arange(self.D3)[None, :] < 2 * self.din).repeat(self.dh, 0).reshape(-1)
        m[self.ob3:self.n] = np.arange(self.D3) < 2 * self.din
        return m


def _ops(emulate: bool):
    if emulate:
        rnd = bf16

        def mm(a, b):
            return (np.asarray(a, np.float64) @ np.asarray(b, np.float64)).astype(np.float32)

        def act(x):
            return np.tanh(np.asarray(x, np.float32))

        cast = np.float32
    else:
        def rnd(x):
            return np.asarray(x, np.float64)

        def mm(a, b):
            return np.asarray(a, np.float64) @ np.asarray(b, np.float64)

        def act(x):
            return np.tanh(np.asarray(x, np.float64))

        cast = np.float64
    return rnd, mm, act, cast


def forward(w, x, din: int, dh: int, emulate: bool = True):
    """Rows x (n, din) -> (n, 2 din) outputs and the activations of the pass."""
    lay = Layout(din, dh)
    rnd, mm, act, cast = _ops(emulate)
    W1, b1, W2, b2, W3, b3 = lay.unpack(w)
    n = x.shape[0]
    xp = np.zeros((n, lay.D0), dtype=cast)
    xp[:, :din] = x
    xb = rnd(xp)
    H1 = act(mm(xb, rnd(W1)) + b1.astype(cast))
    H1b = rnd(H1)
    H2 = act(mm(H1b, rnd(W2)) + b2.astype(cast))
    H2b = rnd(H2)
    O = mm(H2b, rnd(W3)) + b3.astype(cast)
    cache = dict(xb=xb, H1=H1, H1b=H1b, H2=H2, H2b=H2b, O=O)
    return O[:, :2 * din], cache


def loss_grad(w, x, t, din: int, dh: int, emulate: bool = True):
    """Mean negative log density and its gradient (flat, the parameter layout)."""
    lay = Layout(din, dh)
    rnd, mm, act, cast = _ops(emulate)
    W1, b1, W2, b2, W3, b3 = lay.unpack(w)
    n = x.shape[0]
    _, c = forward(w, x, din, dh, emulate)
    O = c["O"]
    mu, ell = O[:, :din].astype(np.float64), O[:, din:2 * din].astype(np.float64)
    s = np.exp(-ell)
    r = (np.asarray(t, np.float64) - np.asarray(x, np.float64) - mu) * s
    loss = float(np.sum(ell + 0.5 * r * r) / n)
    dO = np.zeros((n, lay.D3), dtype=cast)
    dO[:, :din] = -r * s / n
    dO[:, din:2 * din] = (1.0 - r * r) / n
    dOb = rnd(dO)
    g = np.zeros(lay.n, dtype=np.float64)
    g[lay.oW3:lay.ob3] = mm(c["H2b"].T, dOb).reshape(-1)
    g[lay.ob3:lay.n] = dO.astype(np.float64).sum(0)
    dA2 = (mm(dOb, rnd(W3).T) * (1.0 - c["H2"] * c["H2"])).astype(cast)
    dA2b = rnd(dA2)
    g[lay.oW2:lay.ob2] = mm(c["H1b"].T, dA2b).reshape(-1)
    g[lay.ob2:lay.oW3] = dA2.astype(np.float64).sum(0)
    dA1 = (mm(dA2b, rnd(W2).T) * (1.0 - c["H1"] * c["H1"])).astype(cast)
    dA1b = rnd(dA1)
    g[lay.oW1:lay.ob1] = mm(c["xb"].T, dA1b).reshape(-1)
    g[lay.ob1:lay.oW2] = dA1.astype(np.float64).sum(0)
    return loss, g


def adam(w, g, m, v, step: int, lr: float, mask, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW update (step counts from 1, decoupled weight decay wd); returns
    new (w, m, v) as float32."""
    g = np.asarray(g, np.float32)
    m2 = np.where(mask, b1 * m + (1 - b1) * g, m).astype(np.float32)
    v2 = np.where(mask, b2 * v + (1 - b2) * g * g, v).astype(np.float32)
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    w2 = np.where(mask, w - lr * ((m2 / c1) / (np.sqrt(v2 / c2) + eps) + wd * w), w).astype(np.float32)
    return w2, m2, v2
