"""messi_inside_noalloc (every buffer in a static device array) against
messi_inside, through the C ABI: the same results bit for bit, on two calls in a
row with the second one smaller, in both precisions and both tile edges, plus
the length limit."""
import sys

import numpy as np
import pytest

from messi_amd import lib

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _problem(L, seed):
    rng = np.random.default_rng(seed)
    u = -rng.uniform(2.0, 9.0, size=L)
    pm = np.full((L, L), -np.inf)
    iu = np.triu_indices(L, 1)
    v = u[iu[0]] + u[iu[1]] + rng.normal(0.0, 4.0, size=iu[0].size)
    keep = rng.random(iu[0].size) < 0.6
    pm[iu[0][keep], iu[1][keep]] = v[keep]
    pm[iu[1][keep], iu[0][keep]] = v[keep]
    return u, np.asfortranarray(pm)


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_inside_twice_second_smaller(prec, tile):
    ctx = lib.context()
    ctx.set_tile(tile)
    try:
        for L, seed in ((512, 1), (137, 2)):
            _say(f"noalloc inside L={L} {prec} tile={tile}")
            u, pm = _problem(L, seed)
            z0, t0 = lib.inside(u, pm, 1.0, False, prec)
            z1, t1 = lib.inside_noalloc(u, pm, 1.0, False, prec)
            assert z1 == z0
            assert np.array_equal(t1, t0)
            z2 = lib.inside_noalloc(u, pm, 1.0, True, prec)
            assert z2 == z0
            zb = lib.inside(u, pm, 1.7, True, prec)
            assert lib.inside_noalloc(u, pm, 1.7, True, prec) == zb
    finally:
        ctx.set_tile(16)


def test_noalloc_band_and_limit():
    _say("noalloc inside: banded matrix, length limit")
    u, pm = _problem(300, 3)
    d = np.abs(np.subtract.outer(np.arange(300), np.arange(300)))
    pm = np.asfortranarray(np.where(d <= 40, pm, -np.inf))
    for prec in ("fp64", "fp32"):
        z0, t0 = lib.inside(u, pm, 1.0, False, prec)
        z1, t1 = lib.inside_noalloc(u, pm, 1.0, False, prec)
        assert z1 == z0 and np.array_equal(t1, t0)
    u2, pm2 = _problem(lib.NOALLOC_MAX_L + 1, 4)
    with pytest.raises(lib.MessiError, match="512"):
        lib.inside_noalloc(u2, pm2)
    with pytest.raises(lib.MessiError):
        lib.inside_noalloc(u[:0], pm[:0, :0])
    # the context keeps working after the rejected calls
    assert lib.inside_noalloc(u, pm, 1.0, True, "fp64") == lib.inside(u, pm, 1.0, True, "fp64")
