"""Pair nodes computed in their parent's launch (messi_set_pair_fusion).  A node
whose entries mostly serve one entry of its parent is not stored: the parent's
launch computes it from the node's own two children with the same products and
the same scaling, so every result must be bit-identical to storing every node:
the pair log-likelihood matrix (one and several components), log Z of batches
(one dataset, uneven instances, a pairing band with a mask) and posteriors, in
both precisions."""
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
    c.set_pair_fusion(True)


def _pair(ctx, ds, **kw):
    """The same dataset stored node by node and with fused nodes."""
    ctx.set_pair_fusion(False)
    d0 = lib.DeviceDataset(ds, **kw)
    ctx.set_pair_fusion(True)
    d1 = lib.DeviceDataset(ds, **kw)
    return d0, d1


def _stack(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


def _perturbed(a, ds, k, seed=3, **kw):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        v = a.params.copy()
        if i:
            v[0:3] = np.maximum(1.0, v[0:3] * np.exp(rng.normal(0, 0.1, 3)))
            v[3:9] = v[3:9] * np.exp(rng.normal(0, 0.1, 6))
        out.append(model.getparams(v, ds.tree.branchlengths, **kw))
    return out


@pytest.fixture(scope="module")
def c2(ctx):
    _say("50 x 1500 datasets, stored and fused")
    a = synth.config2()
    ds = synth.dataset_of(a)
    d0, d1 = _pair(ctx, ds)
    yield a, ds, d0, d1
    d0.close()
    d1.close()


def test_fusion_applies_at_full_size(c2):
    a, ds, d0, d1 = c2
    s0, s1 = d0.stats(), d1.stats()
    _say("device bytes stored", s0["device_bytes"], "fused", s1["device_bytes"])
    # the same entries; a fused node's own two maps are gone and its parent holds
    # two composed maps in place of one, so the index arena changes size
    assert s0["sum_U"] == s1["sum_U"] and s0["root_entries"] == s1["root_entries"]
    assert s1["device_bytes"] != s0["device_bytes"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pair_matrix_and_batch_full_size(c2, prec):
    a, ds, d0, d1 = c2
    _say("50 x 1500 pair matrix and batch of 4", prec)
    ps = _perturbed(a, ds, 4)
    sp = [_stack(p) for p in ps]
    M0 = d0.pair_loglik(*sp[0][1], prec=prec)
    M1 = d1.pair_loglik(*sp[0][1], prec=prec)
    assert np.array_equal(M0, M1)
    z = []
    for d in (d0, d1):
        b = lib.Batch([d] * 4)
        try:
            z.append(b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, prec))
        finally:
            b.close()
    assert np.array_equal(z[0], z[1]), (z[0], z[1])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mixture_components(ctx, prec):
    _say("3 site categories (9 pair components), 24 x 300", prec)
    a = synth.generate(24, 300, 31)
    ds = synth.dataset_of(a)
    d0, d1 = _pair(ctx, ds)
    try:
        p = model.getparams(a.params, ds.tree.branchlengths, siteCats=3)
        pc = hotpath._stack(model.pair_components(p))
        M0, per0 = d0.pair_loglik(*pc, prec=prec, percomp=True)
        M1, per1 = d1.pair_loglik(*pc, prec=prec, percomp=True)
        assert np.array_equal(M0, M1)
        for x, y in zip(per0, per1):
            assert np.array_equal(x, y)
        assert d0.eval(*_stack(p), 1.0, prec) == d1.eval(*_stack(p), 1.0, prec)
    finally:
        d0.close()
        d1.close()


def test_uneven_instances(ctx):
    _say("batch of 6 uneven instances, stored and fused")
    als = synth.config4(count=6)
    dss = [synth.dataset_of(a) for a in als]
    pairs = [_pair(ctx, d) for d in dss]
    try:
        sp = [_stack(model.getparams(a.params, d.tree.branchlengths)) for a, d in zip(als, dss)]
        for prec in ("fp64", "fp32"):
            z = []
            for k in (0, 1):
                b = lib.Batch([p[k] for p in pairs])
                try:
                    z.append(b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, prec))
                finally:
                    b.close()
            assert np.array_equal(z[0], z[1]), prec
    finally:
        for p in pairs:
            p[0].close()
            p[1].close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_band_mask_and_posterior(ctx, prec):
    _say("30 x 500, band 80 with a pair mask, posteriors", prec)
    a = synth.generate(30, 500, 17)
    ds = synth.dataset_of(a)
    mask = hotpath.gapmask(np.linspace(0.0, 0.6, ds.numcols), 0.5)
    p = model.getparams(a.params, ds.tree.branchlengths)
    for kw in ({"maxbasepairdistance": 80, "mask": mask}, {}):
        d0, d1 = _pair(ctx, ds, **kw)
        try:
            M0 = d0.pair_loglik(*_stack(p)[1], prec=prec)
            M1 = d1.pair_loglik(*_stack(p)[1], prec=prec)
            assert np.array_equal(M0, M1)
            z0, u0, q0 = d0.posterior(*_stack(p), prec=prec)
            z1, u1, q1 = d1.posterior(*_stack(p), prec=prec)
            assert z0 == z1
            assert np.array_equal(u0, u1) and np.array_equal(q0, q1)
        finally:
            d0.close()
            d1.close()
