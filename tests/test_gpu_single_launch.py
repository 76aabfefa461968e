"""The single-launch inside (messi_set_single_launch): the same work items as the
launch sequence, taken from a ticket counter and ordered by per-tile dependency
flags instead of kernel boundaries.  Every addition happens in the same order,
so log Z, the inside tables and the posteriors must be bit-identical to the
launch sequence's: both precisions, both tile edges, every bulk window, with and
without a pairing band, single instances, uneven batches (the 64-instance batch
included) and the 50 x 1500 batch of 16, over repeated evaluations."""
import sys

import numpy as np
import pytest

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
    c.set_single_launch(False)
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


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("tile,window", [(16, 4), (16, 2), (16, 1), (32, 4), (32, 1)])
@pytest.mark.parametrize("L,maxbp", [(1, 10), (2, 10), (40, 100), (130, 100000), (300, 37), (700, 100000),
                                     (700, 90)])
def test_single_launch_is_bit_identical(ctx, prec, tile, window, L, maxbp):
    _say("single launch", prec, "tile", tile, "window", window, "L", L, "maxbp", maxbp)
    u, M = _inputs(L, maxbp, 7 * L + maxbp)
    ctx.set_tile(tile)
    ctx.set_bulk_window(window)
    try:
        ctx.set_single_launch(False)
        z0, t0 = lib.inside(u, M, zonly=False, prec=prec)
        ctx.set_single_launch(True)
        for _ in range(2):
            z1, t1 = lib.inside(u, M, zonly=False, prec=prec)
            assert z1 == z0, (z1, z0)
            assert np.array_equal(t1, t0)
    finally:
        ctx.set_single_launch(False)
        ctx.set_bulk_window(4)
        ctx.set_tile(16)


def _stack(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


def _batch(als):
    dss = [synth.dataset_of(a) for a in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    stacked = [_stack(model.getparams(a.params, d.tree.branchlengths)) for a, d in zip(als, dss)]
    return dss, dds, [s for s, _ in stacked], [q for _, q in stacked]


def test_single_launch_uneven_batch_and_posterior(ctx):
    als = [synth.generate(6 + 3 * i, 90 + 71 * i, 500 + i) for i in range(6)]
    dss, dds, sites, pairs = _batch(als)
    b = lib.Batch(dds)
    try:
        for tiles in ([16] * 6, [16, 32, 16, 32, 32, 16]):
            b.set_tiles(tiles)
            for prec in ("fp64", "fp32"):
                ctx.set_single_launch(False)
                ref = b.eval(sites, pairs, 1.0, prec)
                ctx.set_single_launch(True)
                for rep in range(3):
                    _say("uneven batch", tiles, prec, rep)
                    assert np.array_equal(b.eval(sites, pairs, 1.0, prec), ref)
        p = model.getparams(als[5].params, dss[5].tree.branchlengths)
        for prec in ("fp64", "fp32"):
            ctx.set_single_launch(False)
            z0, r0, q0 = hotpath.posterior(dss[5], p, prec=prec)
            ctx.set_single_launch(True)
            _say("posterior single launch", prec)
            z1, r1, q1 = hotpath.posterior(dss[5], p, prec=prec)
            assert z1 == z0 and np.array_equal(r1, r0) and np.array_equal(q1, q0)
    finally:
        ctx.set_single_launch(False)
        b.close()


def test_single_launch_64_instances(ctx):
    """The 64 uneven instances (lengths 80..600, 10..60 sequences) in one launch."""
    rng = np.random.default_rng(64)
    als = [synth.generate(int(rng.integers(10, 61)), int(rng.integers(80, 601)), 4000 + i) for i in range(64)]
    dss, dds, sites, pairs = _batch(als)
    b = lib.Batch(dds)
    try:
        for prec in ("fp32", "fp64"):
            ctx.set_single_launch(False)
            ref = b.eval(sites, pairs, 1.0, prec)
            ctx.set_single_launch(True)
            for rep in range(2):
                _say("64 instances", prec, rep)
                assert np.array_equal(b.eval(sites, pairs, 1.0, prec), ref)
    finally:
        ctx.set_single_launch(False)
        b.close()


def test_single_launch_config2_batch16(ctx):
    """50 x 1500, 16 parameter vectors: the evaluation the throughput is quoted on."""
    a = synth.config2()
    ds = synth.dataset_of(a)
    dd = hotpath.device_dataset(ds)
    rng = np.random.default_rng(16)
    sites, pairs = [], []
    for k in range(16):
        v = a.params.copy()
        v[:3] *= np.exp(rng.normal(0, 0.05, 3))
        s, q = _stack(model.getparams(v, ds.tree.branchlengths))
        sites.append(s)
        pairs.append(q)
    b = lib.Batch([dd] * 16)
    try:
        for prec in ("fp32", "fp64"):
            ctx.set_single_launch(False)
            ref = b.eval(sites, pairs, 1.0, prec)
            ctx.set_single_launch(True)
            for rep in range(3):
                _say("50 x 1500 batch 16", prec, rep)
                assert np.array_equal(b.eval(sites, pairs, 1.0, prec), ref)
    finally:
        ctx.set_single_launch(False)
        b.close()
