"""Transition matrices built on the device from eigendecompositions
(messi_batch_load_eigen) against the host-built matrices (messi_batch_eval of
the same parameter vectors), through the C ABI."""
import sys

import numpy as np
import pytest

import oracle
from messi_amd import hotpath, lib, mcmc, model, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("zero_branch", [False, True])
def test_load_eigen_equals_host_matrices(zero_branch):
    a = synth.config1()
    ds = synth.dataset_of(a)
    if zero_branch:
        ds.tree.branchlengths[5] = 0.0      # P(0) = I on that branch
    _say(f"load_eigen vs host matrices, zero branch {zero_branch}")
    pm = mcmc.ParamMap(a.params, ds.tree.branchlengths)
    th = pm.theta_of(a.params) + np.random.default_rng(5).normal(0, 0.3, (6, mcmc.DIM))
    dd = hotpath.device_dataset(ds)
    b = lib.Batch([dd] * 6)
    try:
        for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
            ref = b.eval_stacked(*pm.components(th), 1.0, prec)
            b.load_eigen(*pm.eigen_components(th), prec=prec)
            z = b.eval_loaded(1.0, prec)
            assert np.max(np.abs(z - ref) / np.abs(ref)) < tol, prec
            assert np.array_equal(b.eval_loaded(1.0, prec), z)
        # anchor one vector on the oracle (host matrices from the model functions)
        p15, bl = pm.params_of(th[2])
        mp = model.getparams(p15, bl)
        u = oracle.site_loglik(ds, model.site_components(mp))
        M = oracle.pair_loglik(ds, model.pair_components(mp))
        zr = oracle.inside(u, M)[0, 0, -1]
        b.load_eigen(*pm.eigen_components(th), prec="fp64")
        assert abs(b.eval_loaded(1.0, "fp64")[2] - zr) <= 1e-9 * abs(zr)
    finally:
        b.close()


def test_load_eigen_uneven_members():
    """Members of different trees in one batch: each member's matrices come from
    its own decomposition and branch lengths."""
    _say("load_eigen with uneven members")
    als = [synth.generate(6 + 3 * i, 50 + 20 * i, 60 + i) for i in range(4)]
    dss = [synth.dataset_of(x) for x in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    pms = [mcmc.ParamMap(x.params, d.tree.branchlengths) for x, d in zip(als, dss)]
    ths = [pm.theta_of(x.params) for pm, x in zip(pms, als)]
    b = lib.Batch(dds)
    try:
        ec = [pm.eigen_components(th[None]) for pm, th in zip(pms, ths)]
        stacked = [np.concatenate([e[i] for e in ec]) for i in range(10)]
        b.load_eigen(*stacked, [e[10][0] for e in ec], prec="fp64")
        z = b.eval_loaded(1.0, "fp64")
        for m, (pm, th) in enumerate(zip(pms, ths)):
            P4, pi4, lw4, P16, pi16, lw16 = pm.components(th[None])
            zs = dds[m].eval((P4[0], pi4[0], lw4[0]), (P16[0], pi16[0], lw16[0]), 1.0, "fp64")
            assert abs(z[m] - zs) <= 1e-12 * abs(zs), m
    finally:
        b.close()
