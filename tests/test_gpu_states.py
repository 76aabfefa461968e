"""State sampling given a structure (messi_sample_states) on the device versus
the oracle.

The per-component conditionals the device computes match the oracle's
single-component site and pair log-likelihoods (FP64 1e-12 relative, FP32 1e-5);
given the device's own conditionals the oracle draws the same states (the
uniforms are keyed by seed and column); the returned log-likelihoods equal the
device's mixture values from messi_site_loglik / messi_pair_loglik.
"""
import sys

import numpy as np
import pytest

import oracle
from messi_amd import hotpath, lib, model, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _structure(L, rng, minloop=3):
    partner = np.zeros(L, dtype=np.int64)
    stack = []
    for x in range(L):
        r = rng.random()
        if stack and r < 0.3 and x - stack[-1] > minloop:
            y = stack.pop()
            partner[x], partner[y] = y + 1, x + 1
        elif r < 0.6:
            stack.append(x)
    return partner


def _comps(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


def _oracle_conditionals(ds, p, partner):
    L = ds.numcols
    sitec = np.stack([oracle.site_loglik(ds, [c]) for c in model.site_components(p)])
    mask = np.zeros((L, L), dtype=np.uint8, order="F")
    op = np.flatnonzero(partner > np.arange(1, L + 1))
    mask[op, partner[op] - 1] = 1
    pc = model.pair_components(p)
    pairc = np.full((len(pc), L), -np.inf)
    for k, (P, pi, _) in enumerate(pc):
        M = oracle.pair_loglik(ds, [(P, pi, 0.0)], mask=mask)
        pairc[k, op] = M[op, partner[op] - 1]
    return sitec, pairc


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    return float(np.max(np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-300))) if fin.any() else 0.0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("case", [("config1", 1), ("small", 3), ("config2", 2)])
def test_states_match_oracle(prec, tol, case):
    name, cats = case
    _say("sample states", name, "cats", cats, prec)
    if name == "config1":
        a = synth.config1()
    elif name == "config2":
        a = synth.config2()
    else:
        a = synth.generate(8, 90, 23)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths, siteCats=cats)
    partner = _structure(ds.numcols, np.random.default_rng(cats + ds.numcols))
    dd = hotpath.structure_dataset(ds, partner)
    site, pair = _comps(p)
    st, unp, pll, sc, pc = dd.sample_states(site, pair, partner, cats, seed=11, prec=prec, conditionals=True)
    # conditionals against the oracle's single-component log-likelihoods
    osc, opc = _oracle_conditionals(ds, p, partner)
    assert _rel(sc, osc) < tol
    assert _rel(pc, opc) < tol
    # identical draws from the device's own conditionals
    lw = np.array([c[2] for c in model.pair_components(p)])
    ost, ounp, opll = oracle.sample_states(sc, pc, lw, partner, cats, 1.0, seed=11)
    assert np.array_equal(st, ost)
    assert _rel(unp, ounp) < 1e-13 and _rel(pll, opll) < 1e-13
    # the mixture values of the other entry points
    u = dd.site_loglik(*site, prec=prec)
    free = partner == 0
    assert np.allclose(unp[free], u[free], rtol=1e-12 if prec == "fp64" else 1e-6, atol=0)
    assert np.all((st >= 1) & (st <= cats))
    for x in np.flatnonzero(partner > np.arange(1, ds.numcols + 1)):
        assert np.isfinite(pll[x])


def test_seeds_B_and_repeats():
    _say("sample states seeds / B / repeats")
    a = synth.generate(8, 90, 23)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths, siteCats=4)
    partner = _structure(ds.numcols, np.random.default_rng(2))
    dd = hotpath.structure_dataset(ds, partner)
    site, pair = _comps(p)
    s1 = dd.sample_states(site, pair, partner, 4, seed=1)[0]
    assert np.array_equal(s1, dd.sample_states(site, pair, partner, 4, seed=1)[0])
    assert not np.array_equal(s1, dd.sample_states(site, pair, partner, 4, seed=2)[0])
    # B large: the most probable category (pair) of every column
    st, _, _, sc, pc = dd.sample_states(site, pair, partner, 4, seed=3, B=500.0, conditionals=True)
    lw = np.array([c[2] for c in model.pair_components(p)])
    checked = 0
    for x in range(ds.numcols):
        if partner[x] == 0:
            v = sc[:, x]
        elif partner[x] > x + 1:
            v = np.logaddexp.reduce((lw + pc[:, x]).reshape(16, -1), axis=1)
        else:
            continue
        pr = np.exp(500.0 * (v - v.max()))
        pr /= pr.sum()
        if pr.max() < 0.999:   # (near-)tied categories: no single mode to check
            continue
        k = int(np.argmax(v))
        checked += 1
        if partner[x] == 0:
            assert st[x] == k + 1
        else:
            assert (st[x], st[partner[x] - 1]) == (k // 4 + 1, k % 4 + 1)
    assert checked > ds.numcols // 2
    # the reference-named entry point sets params.states
    unp, pll, states = hotpath.samplestates(7, ds, p, partner)
    assert np.array_equal(p.states, states) and len(states) == ds.numcols


def test_rejects_bad_arguments():
    _say("sample states errors")
    a = synth.generate(6, 60, 5)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths, siteCats=2)
    site, pair = _comps(p)
    partner = _structure(ds.numcols, np.random.default_rng(4))
    far = np.zeros(ds.numcols, dtype=np.int64)
    far[0], far[59] = 60, 1
    dd = hotpath.device_dataset(ds, 20)   # pairs beyond distance 20 are not in the list
    with pytest.raises(lib.MessiError, match="lacks a pair"):
        dd.sample_states(site, pair, far, 2)
    dd2 = hotpath.structure_dataset(ds, partner)
    with pytest.raises(lib.MessiError, match="K4"):
        dd2.sample_states(site, pair, partner, 3)
    with pytest.raises(lib.MessiError, match="B"):
        dd2.sample_states(site, pair, partner, 2, B=0.0)
    bad = partner.copy()
    x = int(np.flatnonzero(partner)[0])
    bad[x] = 0
    with pytest.raises(lib.MessiError, match="symmetric"):
        dd2.sample_states(site, pair, bad, 2)
