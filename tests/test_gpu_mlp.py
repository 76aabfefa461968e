"""Proposal network on the device (bf16 MFMA, FP32 accumulate) versus the numpy
restatement in oracle/mlp.py, through the C ABI."""
import sys

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

DIN, DH = 12, 128


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def net():
    from messi_amd import lib
    _say("mlp create din=12 dh=128")
    m = lib.Mlp(DIN, DH, seed=7)
    yield m
    m.close()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, DIN)).astype(np.float32)
    t = (0.3 * x + rng.normal(0, 0.7, (n, DIN))).astype(np.float32)
    return x, t


def test_params_layout_and_padding(net):
    lay = omlp.Layout(DIN, DH)
    assert net.num_params == lay.n
    w = net.get_params()
    assert np.all(w[~lay.mask()] == 0.0)
    assert np.count_nonzero(w[lay.oW1:lay.ob1]) > 0.9 * DIN * DH


def test_forward_matches_oracle(net):
    x, _ = _data(37, 1)           # 37 rows: ragged against the 16-row tiles
    _say("mlp forward n=37")
    out = net.forward(x)
    want, _ = omlp.forward(net.get_params(), x, DIN, DH, emulate=True)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(out - want)) <= 1e-3 * scale + 1e-6


def test_loss_grad_matches_oracle(net):
    x, t = _data(53, 2)
    w = np.random.default_rng(3).normal(0, 0.05, net.num_params).astype(np.float32) * omlp.Layout(DIN, DH).mask()
    w = w.astype(np.float32)
    net.set_params(w)
    _say("mlp loss_grad n=53")
    loss, g = net.loss_grad(x, t)
    lo, go = omlp.loss_grad(w, x, t, DIN, DH, emulate=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    lay = omlp.Layout(DIN, DH)
    for a, b in ((lay.oW1, lay.ob1), (lay.ob1, lay.oW2), (lay.oW2, lay.ob2), (lay.ob2, lay.oW3),
                 (lay.oW3, lay.ob3), (lay.ob3, lay.n)):
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 2e-3 * np.linalg.norm(go[a:b]) + 1e-7, (a, b)
    # and the oracle's gradient is the loss's derivative up to bf16 rounding
    _, gx = omlp.loss_grad(w, x, t, DIN, DH, emulate=False)
    assert np.linalg.norm(g - gx) <= 0.02 * np.linalg.norm(gx)


def test_adam_steps_match_oracle(net):
    from messi_amd import lib
    m2 = lib.Mlp(DIN, DH, seed=11)
    try:
        lay = omlp.Layout(DIN, DH)
        mask = lay.mask()
        w = m2.get_params()
        m = np.zeros_like(w)
        v = np.zeros_like(w)
        for step in (1, 2, 3):
            x, t = _data(64, 10 + step)
            _, go = omlp.loss_grad(w, x, t, DIN, DH, emulate=True)
            wdec = 0.05 if step == 3 else 0.0
            w, m, v = omlp.adam(w, go, m, v, step, 1e-3, mask, wd=wdec)
            _say(f"mlp train_step {step}")
            m2.train_step(x, t, lr=1e-3, weight_decay=wdec)
            wd = m2.get_params()
            d = np.abs(wd - w)
            # an update is at most lr per step; a gradient entry near zero can flip sign
            assert np.max(d) <= 2.0 * 1e-3 * step + 1e-6
            assert np.median(d[mask]) <= 1e-5
            assert np.all(wd[~mask] == 0.0)
            w = wd      # continue from the device's weights
    finally:
        m2.close()


def test_training_learns_conditional_gaussian():
    """t = 0.4 x + 0.5 eps, fresh rows every step: after training mu(x) ~ -0.6 x,
    exp(ell) ~ 0.5 and the loss is near its optimum 12 (log 0.5 + 1/2)."""
    from messi_amd import lib
    rng = np.random.default_rng(5)
    m = lib.Mlp(DIN, DH, seed=3)
    try:
        _say("mlp training 800 steps")
        losses = []
        for it in range(800):
            x = rng.normal(0, 1, (256, DIN)).astype(np.float32)
            t = (0.4 * x + 0.5 * rng.normal(0, 1, (256, DIN))).astype(np.float32)
            losses.append(m.train_step(x, t, lr=2e-3))
        x = rng.normal(0, 1, (512, DIN)).astype(np.float32)
        out = m.forward(x)
        mu, ell = out[:, :DIN], out[:, DIN:]
        assert np.sqrt(np.mean((mu + 0.6 * x) ** 2)) < 0.12
        assert abs(np.mean(np.exp(ell)) - 0.5) < 0.05
        optimum = DIN * (np.log(0.5) + 0.5)
        assert abs(np.mean(losses[-50:]) - optimum) < 0.3
        assert np.mean(losses[-50:]) < losses[0] - 5.0
    finally:
        m.close()
