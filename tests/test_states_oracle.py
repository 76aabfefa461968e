"""State sampling given a structure (samplestates, src/FARCE.jl:1398-1486) in the
oracle, pinned without retyping its sums:

* the unpaired and paired log-likelihoods it returns equal the oracle's own
  mixture log-likelihoods (site_loglik / pair_loglik over all components, a
  separate mixing path);
* over many seeds the drawn categories follow the normalised conditionals raised
  to B (chi-square), a large B gives the most probable category, and the closing
  column of a pair gets the second category of the opener's draw.
"""
import numpy as np
import pytest
from scipy import stats

import oracle
from messi_amd import model, synth


def _structure(L, rng):
    partner = np.zeros(L, dtype=np.int64)
    stack = []
    for x in range(L):
        r = rng.random()
        if stack and r < 0.35 and x - stack[-1] >= 3:
            y = stack.pop()
            partner[x], partner[y] = y + 1, x + 1
        elif r < 0.6:
            stack.append(x)
    return partner


def _conditionals(ds, p, partner):
    """sitec (cats x L) and pairc (K16 x L) from single-component oracle calls."""
    L = ds.numcols
    sc = model.site_components(p)
    pc = model.pair_components(p)
    sitec = np.stack([oracle.site_loglik(ds, [c]) for c in sc])
    mask = np.zeros((L, L), dtype=np.uint8, order="F")
    op = np.flatnonzero(partner > np.arange(1, L + 1))
    mask[op, partner[op] - 1] = 1
    pairc = np.full((len(pc), L), -np.inf)
    for k, (P, pi, _) in enumerate(pc):
        M = oracle.pair_loglik(ds, [(P, pi, 0.0)], mask=mask)
        pairc[k, op] = M[op, partner[op] - 1]
    return sitec, pairc, np.array([c[2] for c in pc]), mask


@pytest.mark.parametrize("cats", [1, 3])
def test_loglikelihoods_equal_the_mixtures(cats):
    a = synth.generate(6, 40, 17)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths, siteCats=cats)
    partner = _structure(ds.numcols, np.random.default_rng(cats))
    assert np.count_nonzero(partner) > 0
    sitec, pairc, lw, mask = _conditionals(ds, p, partner)
    states, unp, pll = oracle.sample_states(sitec, pairc, lw, partner, cats, 1.0, seed=5)
    x = np.arange(ds.numcols)
    free, op = partner == 0, partner > x + 1
    mix_u = oracle.site_loglik(ds, model.site_components(p))
    mix_p = oracle.pair_loglik(ds, model.pair_components(p), mask=mask)
    assert np.allclose(unp[free], mix_u[free], rtol=1e-12, atol=0)
    assert np.all(unp[~free] == 0.0)
    assert np.allclose(pll[op], mix_p[op, partner[op] - 1], rtol=1e-12, atol=0)
    assert np.all(np.isneginf(pll[~op]))
    assert np.all((states >= 1) & (states <= cats))


def _probs(sitec, pairc, lw, partner, cats, B):
    """Category (pair) probabilities per column from the conditionals, by softmax."""
    L = sitec.shape[1]
    M = pairc.shape[0] // (cats * cats)
    out = {}
    for x in range(L):
        if partner[x] == 0:
            v = sitec[:, x] * B
        elif partner[x] > x + 1:
            comp = (lw + pairc[:, x]).reshape(cats * cats, M)
            v = np.logaddexp.reduce(comp, axis=1) * B
        else:
            continue
        out[x] = np.exp(v - np.logaddexp.reduce(v))
    return out


@pytest.mark.parametrize("B", [1.0, 2.5])
def test_draws_follow_the_conditionals(B):
    rng = np.random.default_rng(3)
    cats, M, L = 3, 2, 9
    partner = np.array([6, 0, 5, 0, 3, 1, 0, 9, 8])
    sitec = rng.normal(-20, 1.0, (cats, L))
    pairc = rng.normal(-40, 1.0, (cats * cats * M, L))
    lw = np.log(rng.dirichlet(np.ones(cats * cats * M)))
    probs = _probs(sitec, pairc, lw, partner, cats, B)
    n = 6000
    draws = np.array([oracle.sample_states(sitec, pairc, lw, partner, cats, B, seed=s)[0] for s in range(n)])
    for x, pr in probs.items():
        if partner[x] == 0:
            idx = draws[:, x] - 1
        else:
            y = partner[x] - 1
            idx = (draws[:, x] - 1) * cats + (draws[:, y] - 1)
        obs = np.bincount(idx, minlength=len(pr))
        assert stats.chisquare(obs, pr * n).pvalue > 1e-4, (x, obs, pr * n)


def test_large_B_gives_the_mode_and_partners_get_the_second_category():
    rng = np.random.default_rng(8)
    cats, L = 2, 6
    partner = np.array([4, 0, 0, 1, 6, 5])
    sitec = rng.normal(-10, 2.0, (cats, L))
    pairc = rng.normal(-30, 2.0, (cats * cats, L))
    lw = np.log(np.full(cats * cats, 0.25))
    probs = _probs(sitec, pairc, lw, partner, cats, 1.0)
    for s in range(20):
        st, _, _ = oracle.sample_states(sitec, pairc, lw, partner, cats, 400.0, seed=s)
        for x, pr in probs.items():
            k = int(np.argmax(pr))
            if partner[x] == 0:
                assert st[x] == k + 1
            else:
                assert (st[x], st[partner[x] - 1]) == (k // cats + 1, k % cats + 1)
