"""Host-side argument checks, content-keyed device-dataset caches and the gap
mask rule (no GPU: nothing here reaches a device call)."""
import numpy as np
import pytest

from messi_amd import hotpath, lib


def test_gapmask_follows_maskgapped_rule():
    """maskgapped! (src/RNATools.jl:441-451): (i, j) is masked when either gap
    frequency is >= the cutoff; exactly at the cutoff counts as masked."""
    g = np.array([0.0, 0.5, 0.49, 0.75, 0.1])
    m = hotpath.gapmask(g, 0.5)
    for i in range(5):
        for j in range(5):
            assert m[i, j] == (g[i] < 0.5 and g[j] < 0.5)
    M = np.arange(25.0).reshape(5, 5)
    hotpath.maskgapped(M, g, 0.5, -np.inf)
    assert np.array_equal(np.isfinite(M), m.astype(bool))
    M2 = np.ones((5, 5))
    hotpath.maskgapped(M2, g, 0.5)           # default v = 0.0, as the reference
    assert np.array_equal(M2, m.astype(float))
    with pytest.raises(ValueError):
        hotpath.maskgapped(np.zeros((4, 4)), g, 0.5)


def test_maskgapped_structure_drops_gapped_pairs():
    g = np.array([0.0, 0.6, 0.0, 0.0, 0.2, 0.0])
    paired = np.array([6, 4, 0, 2, 0, 1])      # 1-6 and 2-4 (1-based)
    s = hotpath.maskgapped_structure(paired, g, 0.5)
    assert list(s) == [6, 0, 0, 0, 0, 1]
    assert list(paired) == [6, 4, 0, 2, 0, 1]   # a copy, as the reference


def test_mask_key_is_content_not_identity():
    a = np.zeros((6, 6), dtype=np.uint8)
    a[1, 3] = 1
    b = a.copy()
    assert hotpath.mask_key(a) == hotpath.mask_key(b)
    b[2, 4] = 7
    assert hotpath.mask_key(a) != hotpath.mask_key(b)
    assert hotpath.mask_key(a) == hotpath.mask_key(np.asfortranarray(a))
    assert hotpath.mask_key(np.zeros((6, 6))) != hotpath.mask_key(np.zeros((5, 5)))


class _Fake:
    def __init__(self):
        self.closed = False

    def close(self):
        self.closed = True


def test_cache_is_bounded_and_closes_evicted_entries():
    class D:
        pass
    d = D()
    made = []

    def mk():
        made.append(_Fake())
        return made[-1]
    x1 = hotpath._cached(d, "_c", 1, mk, 2)
    x2 = hotpath._cached(d, "_c", 2, mk, 2)
    assert hotpath._cached(d, "_c", 1, mk, 2) is x1      # hit: 1 becomes most recent
    x3 = hotpath._cached(d, "_c", 3, mk, 2)              # evicts 2, the least recent
    assert x2.closed and not x1.closed and not x3.closed
    assert len(made) == 3 and set(d._c) == {1, 3}


def test_component_shapes_are_checked():
    with pytest.raises(ValueError, match="transition matrices"):
        lib._components(np.zeros((1, 5, 4, 4)), np.zeros((1, 4)), np.zeros(1), 7, 4, "site")
    with pytest.raises(ValueError, match="frequencies"):
        lib._components(np.zeros((2, 7, 16, 16)), np.zeros((1, 16)), np.zeros(2), 7, 16, "pair")
    K, P, pi, lw = lib._components(np.zeros((7, 4, 4)), np.zeros(4), np.zeros(1), 7, 4, "site")
    assert K == 1 and P.shape == (1, 7, 4, 4) and pi.shape == (1, 4) and lw.shape == (1,)


def test_standalone_table_shapes_are_checked():
    with pytest.raises(ValueError):
        lib._square(np.zeros((4, 5)), 4, "paired")
    with pytest.raises(ValueError):
        lib._table3(np.zeros((3, 4, 5)), 4, "inside table")
    assert lib._table3(np.zeros((3, 4, 4)), 4, "t").flags.f_contiguous
