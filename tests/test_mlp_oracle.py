"""The numpy restatement of the proposal network: pinned by torch's bf16
conversion, central finite differences and Adam's closed-form first step."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_matches_torch_rounding():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf]]).astype(np.float32)
    # exact ties: mantissa low half 0x8000, with even and odd kept bit
    u = (np.arange(200, dtype=np.uint32) << 16) | 0x3F808000
    x = np.concatenate([x, u.view(np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = mlp.bf16(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _net(din=3, dh=16, n=7, seed=1, scale=0.5):
    rng = np.random.default_rng(seed)
    lay = mlp.Layout(din, dh)
    w = rng.normal(0, scale, lay.n) * lay.mask()
    x = rng.normal(0, 1, (n, din))
    t = x + rng.normal(0, 1, (n, din))
    return lay, w, x, t


def test_layout_counts():
    lay = mlp.Layout(12, 128)
    assert (lay.D0, lay.D3) == (16, 32)
    assert lay.n == 16 * 128 + 128 + 128 * 128 + 128 + 128 * 32 + 32
    assert lay.mask().sum() == 12 * 128 + 128 + 128 * 128 + 128 + 128 * 24 + 24


def test_gradient_matches_finite_differences():
    lay, w, x, t = _net()
    loss, g = mlp.loss_grad(w, x, t, lay.din, lay.dh, emulate=False)
    rng = np.random.default_rng(2)
    live = np.flatnonzero(lay.mask())
    idx = np.concatenate([rng.choice(live[live < lay.ob1], 10), rng.choice(live[(live >= lay.ob1) & (live < lay.ob2)], 10),
                          rng.choice(live[live >= lay.ob2], 20)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss_grad(wp, x, t, lay.din, lay.dh, emulate=False)[0]
              - mlp.loss_grad(wm, x, t, lay.din, lay.dh, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padding of the input layer and the output layer carries no gradient
    m = lay.mask()
    pad = ~m
    pad[lay.oW2:lay.ob2] = False
    assert np.all(g[pad & (np.arange(lay.n) >= lay.oW3)] == 0.0)
    assert np.all(g[pad & (np.arange(lay.n) < lay.ob1)] == 0.0)


def test_loss_closed_form():
    """With every weight zero the network outputs mu = ell = 0: the loss is
    0.5 * mean_rows sum_d (t - x)^2 and d loss / d b3 is (-(t-x) mean, 1 - (t-x)^2 mean)."""
    lay, w, x, t = _net()
    w[:] = 0.0
    loss, g = mlp.loss_grad(w, x, t, lay.din, lay.dh, emulate=False)
    d = t - x
    assert loss == pytest.approx(0.5 * np.sum(d * d) / len(x), rel=1e-12)
    gb3 = g[lay.ob3:lay.n]
    assert np.allclose(gb3[:lay.din], -d.mean(0), atol=1e-12)
    assert np.allclose(gb3[lay.din:2 * lay.din], (1 - d * d).mean(0), atol=1e-12)


def test_emulated_gradient_close_to_exact():
    lay, w, x, t = _net(din=12, dh=128, n=40, scale=0.1)
    _, ge = mlp.loss_grad(w, x, t, 12, 128, emulate=True)
    _, g = mlp.loss_grad(w, x, t, 12, 128, emulate=False)
    assert np.linalg.norm(ge - g) <= 0.02 * np.linalg.norm(g)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    n = 50
    w = rng.normal(0, 1, n).astype(np.float32)
    g = rng.normal(0, 1, n).astype(np.float32)
    mask = rng.random(n) < 0.8
    z = np.zeros(n, np.float32)
    w2, m2, v2 = mlp.adam(w, g, z, z, 1, 1e-2, mask)
    want = np.where(mask, w - 1e-2 * g / (np.abs(g) + 1e-8), w)
    assert np.allclose(w2, want, atol=1e-6)
    assert np.all(m2[~mask] == 0) and np.all(v2[~mask] == 0)


def test_adamw_decay_closed_form():
    """With a zero gradient AdamW only shrinks: w <- w (1 - lr wd)."""
    w = np.linspace(-1, 1, 11).astype(np.float32)
    z = np.zeros_like(w)
    mask = np.ones(11, bool)
    w2, _, _ = mlp.adam(w, z, z, z, 1, 0.1, mask, wd=0.5)
    assert np.allclose(w2, w * (1 - 0.05), atol=1e-7)
