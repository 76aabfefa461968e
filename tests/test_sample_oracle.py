"""Stochastic traceback (structure sampling) and the KH99 structure prior in the
oracle, pinned without their recursions:

* the prior of every nested structure on up to 9 positions equals the log of its
  unique KH99 derivation probability computed by an independent recursion;
* the sampler's structure frequencies match the exact posterior over all
  structures (enumeration, chi-square) on small inputs, and its pair
  frequencies match the outside-based base-pair posteriors on a larger one;
* the returned prior is the prior of the returned structure.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle
from bruteforce import kh99_derivation, structures


def _one_based(partner):
    return [k + 1 if k >= 0 else 0 for k in partner]


@pytest.mark.parametrize("L", [1, 2, 4, 6, 9])
def test_prior_equals_derivation_probability(L):
    n = 0
    for s in structures(L):
        g = kh99_derivation(s)
        if g == 0.0:
            continue
        assert oracle.kh99_prior(_one_based(s)) == pytest.approx(math.log(g), rel=1e-12, abs=1e-12)
        n += 1
    assert n > 0


def _inputs(L, seed):
    rng = np.random.default_rng(seed)
    u = rng.normal(-1.0, 0.5, L)
    M = np.full((L, L), -np.inf)
    for i in range(L):
        for j in range(i + 1, L):
            M[i, j] = M[j, i] = rng.normal(-1.5, 1.5)
    return u, M


def _exact(u, M):
    L = len(u)
    probs = {}
    for s in structures(L):
        g = kh99_derivation(s)
        if g == 0.0:
            continue
        le = sum(u[i] for i, k in enumerate(s) if k < 0) + sum(M[i, k] for i, k in enumerate(s) if k > i)
        probs[tuple(_one_based(s))] = g * math.exp(le)
    Z = sum(probs.values())
    return {k: v / Z for k, v in probs.items()}, math.log(Z)


@pytest.mark.parametrize("L,seed", [(5, 1), (7, 2)])
def test_sampler_matches_exact_structure_distribution(L, seed):
    u, M = _inputs(L, seed)
    ins = oracle.inside(u, M)
    exact, logZ = _exact(u, M)
    assert ins[0, 0, L - 1] == pytest.approx(logZ, rel=1e-12)
    n = 20000
    S, pri = oracle.sample_structures(ins, M, n, seed=11)
    counts = {}
    for row in S:
        counts[tuple(int(v) for v in row)] = counts.get(tuple(int(v) for v in row), 0) + 1
    assert set(counts) <= set(exact)   # only structures of positive probability
    keys = sorted(exact, key=lambda k: -exact[k])
    exp_c = np.array([exact[k] * n for k in keys])
    obs_c = np.array([counts.get(k, 0) for k in keys], dtype=float)
    big = exp_c >= 5
    obs = np.append(obs_c[big], obs_c[~big].sum())
    expd = np.append(exp_c[big], exp_c[~big].sum())
    if expd[-1] == 0:
        obs, expd = obs[:-1], expd[:-1]
    p = stats.chisquare(obs, expd).pvalue
    assert p > 1e-4, p
    for row, pr in zip(S[:50], pri[:50]):
        assert pr == oracle.kh99_prior(row)


def test_pair_frequencies_match_posteriors():
    rng = np.random.default_rng(5)
    L = 40
    u = rng.normal(-1.0, 0.3, L)
    M = np.full((L, L), -np.inf)
    # a planted helix makes some pairs likely
    for i in range(L):
        for j in range(i + 4, L):
            M[i, j] = M[j, i] = rng.normal(-3.0, 1.0)
    for k in range(6):
        M[5 + k, 30 - k] = M[30 - k, 5 + k] = 3.0
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    n = 6000
    S, _ = oracle.sample_structures(ins, M, n, seed=3)
    freq = np.zeros((L, L))
    unp = np.zeros(L)
    for row in S:
        for i, k in enumerate(row):
            if k > 0:
                freq[i, k - 1] += 1
            else:
                unp[i] += 1
    freq /= n
    unp /= n
    se = np.sqrt(np.maximum(pair * (1 - pair), 1e-4) / n)
    assert np.max(np.abs(freq - pair) / se) < 5.5
    assert np.max(np.abs(unp - res) / np.sqrt(np.maximum(res * (1 - res), 1e-4) / n)) < 5.5
    assert pair[5, 30] > 0.5


def test_uniforms_are_keyed_and_in_range():
    vals = [oracle.sample_uniform(9, s, sym, x, y) for s in range(3) for sym in range(3)
            for x in range(4) for y in range(x, 6)]
    assert all(0.0 < v <= 1.0 for v in vals)
    assert len(set(vals)) == len(vals)
    assert oracle.sample_uniform(9, 1, 0, 2, 5) == oracle.sample_uniform(9, 1, 0, 2, 5)
    # roughly uniform
    many = np.array([oracle.sample_uniform(1, s, 0, 0, 1) for s in range(20000)])
    assert stats.kstest(many, "uniform").pvalue > 1e-4
