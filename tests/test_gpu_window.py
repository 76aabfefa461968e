"""The inside schedule's bulk window (messi_set_bulk_window) changes how often a
bulk accumulator tile is read and written, never what is added to it or in
which order: log Z and the inside tables must be bit-identical for windows 1, 2
and 4, in both precisions, both tile edges, with and without a pairing band, for
single instances and for uneven batches (whose posteriors are compared too)."""
import sys

import numpy as np
import pytest

import oracle
from messi_amd import hotpath, lib, model, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    _say("context create")
    c = lib.context()
    yield c
    c.set_bulk_window(4)
    c.set_tile(16)


def _inputs(L, maxbp, seed):
    rng = np.random.default_rng(seed)
    u = rng.normal(-1.2, 0.4, L)
    M = np.full((L, L), -np.inf)
    for i in range(L):
        for j in range(i + 4, min(L, i + maxbp + 1)):
            M[i, j] = M[j, i] = rng.normal(-2.0, 1.5)
    return u, M


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("L,maxbp", [(1, 10), (2, 10), (40, 100), (130, 100000), (300, 100000), (300, 37),
                                     (700, 100000), (700, 90)])
def test_window_is_bit_identical(ctx, prec, tile, L, maxbp):
    _say("bulk window", prec, "tile", tile, "L", L, "maxbp", maxbp)
    u, M = _inputs(L, maxbp, L + maxbp)
    ctx.set_tile(tile)
    try:
        ctx.set_bulk_window(1)
        z1, t1 = lib.inside(u, M, zonly=False, prec=prec)
        for w in (2, 4, 2):
            ctx.set_bulk_window(w)
            zw, tw = lib.inside(u, M, zonly=False, prec=prec)
            assert zw == z1, (w, zw, z1)
            assert np.array_equal(tw, t1), w
    finally:
        ctx.set_bulk_window(4)
        ctx.set_tile(16)
    if L in (130, 300) and prec == "fp64":
        ins = oracle.inside(u, M)
        assert abs(z1 - ins[0, 0, -1]) <= 1e-9 * abs(ins[0, 0, -1])


def _stack(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


def test_window_batch_and_posterior(ctx):
    als = [synth.generate(6 + 3 * i, 90 + 71 * i, 300 + i) for i in range(6)]
    dss = [synth.dataset_of(a) for a in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    stacked = [_stack(model.getparams(a.params, d.tree.branchlengths)) for a, d in zip(als, dss)]
    sites = [s for s, _ in stacked]
    pairs = [q for _, q in stacked]
    b = lib.Batch(dds)
    try:
        for prec in ("fp64", "fp32"):
            ctx.set_bulk_window(1)
            ref = b.eval(sites, pairs, 1.0, prec)
            for w in (4, 2):
                _say("batch window", prec, w)
                ctx.set_bulk_window(w)
                assert np.array_equal(b.eval(sites, pairs, 1.0, prec), ref), w
        p = model.getparams(als[5].params, dss[5].tree.branchlengths)
        for prec in ("fp64", "fp32"):
            ctx.set_bulk_window(1)
            z1, r1, q1 = hotpath.posterior(dss[5], p, prec=prec)
            ctx.set_bulk_window(4)
            _say("posterior window 4", prec)
            z4, r4, q4 = hotpath.posterior(dss[5], p, prec=prec)
            assert z4 == z1 and np.array_equal(r4, r1) and np.array_equal(q4, q1)
    finally:
        ctx.set_bulk_window(4)
        b.close()


def test_window_rejects_bad_values(ctx):
    for w in (0, 3, 8, -1):
        with pytest.raises(lib.MessiError, match="window"):
            ctx.set_bulk_window(w)


def test_step_timings(ctx):
    """messi_batch_step_timings: the inside step kernels' launches per tile edge
    and their device time, from a profiled evaluation only."""
    als = [synth.generate(6, 200, 400), synth.generate(7, 300, 401)]
    dss = [synth.dataset_of(a) for a in als]
    stacked = [_stack(model.getparams(a.params, d.tree.branchlengths)) for a, d in zip(als, dss)]
    b = lib.Batch([hotpath.device_dataset(d) for d in dss])
    try:
        with pytest.raises(lib.MessiError, match="profiled"):
            b.step_timings()
        b.set_tiles([16, 32])
        ctx.set_profiling(True)
        try:
            _say("step timings")
            b.eval([s for s, _ in stacked], [q for _, q in stacked], 1.0, "fp32")
            ms, n = b.step_timings()
        finally:
            ctx.set_profiling(False)
        assert list(n) == [-(-200 // 16), -(-300 // 32)]
        assert ms[0] > 0 and ms[1] > 0
        assert ms.sum() <= b.timings()[2]   # within the inside phase
    finally:
        b.close()
