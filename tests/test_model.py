"""Host model layer: rate matrices and P(t) pinned by independent mathematics."""
import numpy as np
import pytest
from scipy.linalg import expm

from messi_amd import model


def _rng_params(seed):
    rng = np.random.default_rng(seed)
    q = rng.dirichlet(np.ones(6)) * rng.uniform(0.5, 3.0)
    pi = rng.dirichlet(np.full(4, 5.0))
    lam = rng.uniform(1.0, 3.0, size=3)
    return q, pi, lam


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gtr_rows_and_detailed_balance(seed):
    q, pi, _ = _rng_params(seed)
    Q = model.gtr(*q, pi)
    assert np.allclose(Q.sum(axis=1), 0.0, atol=1e-14)
    F = pi[:, None] * Q
    assert np.allclose(F, F.T, atol=1e-15)
    assert np.allclose(pi @ Q, 0.0, atol=1e-14)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_muse_stationary_and_detailed_balance(seed):
    q, pi, lam = _rng_params(seed)
    Q = model.muse(*lam, *q, pi, 1.3, 0.7)
    pi16 = model.calculatemusefreqs(pi, *lam)
    assert abs(pi16.sum() - 1.0) < 1e-14
    assert np.allclose(Q.sum(axis=1), 0.0, atol=1e-13)
    F = pi16[:, None] * Q
    assert np.allclose(F, F.T, atol=1e-14)
    # Only single-nucleotide changes are allowed.
    for h in range(16):
        for v in range(16):
            if h != v and (h // 4 != v // 4) and (h % 4 != v % 4):
                assert Q[h, v] == 0.0


def test_muse_lambda_one_is_kronecker_sum():
    q, pi, _ = _rng_params(5)
    r1, r2 = 0.6, 1.7
    Q16 = model.muse(1.0, 1.0, 1.0, *q, pi, r1, r2)
    Q4 = model.gtr(*q, pi)
    K = np.kron(Q4 * r1, np.eye(4)) + np.kron(np.eye(4), Q4 * r2)
    assert np.allclose(Q16, K, atol=1e-14)
    assert np.allclose(model.calculatemusefreqs(pi, 1, 1, 1), np.kron(pi, pi), atol=1e-16)


@pytest.mark.parametrize("dim", [4, 16])
def test_transition_matrices_match_expm(dim):
    q, pi, lam = _rng_params(7)
    Q = model.gtr(*q, pi) if dim == 4 else model.muse(*lam, *q, pi)
    bl = np.array([0.0, 0.0, 1e-6, 0.05, 0.3, 2.5])
    P = model.gettransitionmatriceslist(bl, Q)
    assert np.all(P[0] == 0.0)                        # root entry unused
    assert np.array_equal(P[1], np.eye(dim))          # zero branch -> identity
    for n in range(2, len(bl)):
        E = expm(Q * bl[n])
        E = np.where(E <= 0.0, 1e-50, E)
        assert np.allclose(P[n], E, rtol=1e-9, atol=1e-13)
        assert np.allclose(P[n].sum(axis=1), 1.0, atol=1e-12)
    # Chapman-Kolmogorov
    Pa = model.gettransitionmatriceslist([0.0, 0.2], Q)[1]
    Pb = model.gettransitionmatriceslist([0.0, 0.3], Q)[1]
    Pab = model.gettransitionmatriceslist([0.0, 0.5], Q)[1]
    assert np.allclose(Pa @ Pb, Pab, atol=1e-12)


def test_getparams_single_category():
    v = np.array([2.0, 1.5, 1.2, .1, .2, .3, .4, .5, .6, 1.0, 1.0, 0.3, 0.2, 0.25, 1e-4])
    p = model.getparams(v, np.zeros(5))
    assert np.allclose(p.freqs, [0.3, 0.2, 0.25, 0.25])
    assert p.lambdaratesGC[0] == 1.0 and p.lambdaratesAT[0] == 0.5
    comps = model.pair_components(p)
    assert len(comps) == 1 and comps[0][2] == 0.0
    sites = model.site_components(p)
    assert len(sites) == 1 and sites[0][2] == 0.0


def test_discretizegamma3_closed_form():
    """With the gamma parameter at 1 (getparams' lambdaGammaScaleGC) the rates are
    Exp(1) quantiles at the category midpoints, -log(1 - (i - 1/2)/n), scaled by
    lambda - 1, after a rate-0 category of weight lambdazeroweight."""
    n = 4
    w, gc, w2, at, w3, gt = model.discretizegamma3(0.3, 2.0, 1.5, 1.0, 1.0, n)
    q = -np.log(1.0 - (np.arange(1, n + 1) - 0.5) / n)
    assert np.allclose(gc, np.concatenate([[0.0], q]) * 1.0, rtol=1e-12, atol=0)
    assert np.allclose(at, np.concatenate([[0.0], q]) * 0.5, rtol=1e-12, atol=0)
    assert np.all(gt == 0.0)
    assert np.allclose(w, [0.3] + [0.7 / n] * n) and abs(w.sum() - 1.0) < 1e-15
    assert np.array_equal(w, w2) and np.array_equal(w, w3)


def test_getparams_lambda_categories_and_fix_flags():
    v = np.array([2.0, 1.5, 1.2, 0.3, 1.1, 0.4, 0.7, 2.0, 1.0, 1.7, 0.8, 0.2, 0.3, 0.25, 0.2])
    bl = np.array([0.0, 0.1, 0.2])
    p = model.getparams(v, bl, siteCats=3, lambdaCats=5)
    q = -np.log(1.0 - (np.arange(1, 5) - 0.5) / 4)
    assert np.allclose(p.lambdaweightsGC, [0.2] + [0.2] * 4)
    assert np.allclose(p.lambdaratesGC[1:], q * 1.0) and np.allclose(p.lambdaratesAT[1:], q * 0.5)
    assert np.allclose(p.lambdaratesGT[1:], q * 0.2)
    assert len(model.site_components(p)) == 3 and len(model.pair_components(p)) == 45
    w = np.exp([c[2] for c in model.pair_components(p)])
    assert abs(w.sum() - 1.0) < 1e-12
    pg = model.getparams(v, bl, siteCats=1, lambdaCats=5, fixGU=True, fixGCAU=True)
    assert np.all(pg.lambdaratesGT == 0.0) and np.array_equal(pg.lambdaratesGC, pg.lambdaratesAT)
    # one category: lambda - 1 with weight 1, whatever the flags asked for (src/FARCE.jl:645-652)
    p1 = model.getparams(v, bl, lambdaCats=1, fixGU=True)
    assert np.allclose(p1.lambdaratesGT, [0.2]) and np.allclose(p1.lambdaweightsGC, [1.0])


def test_lambda_mixture_collapses_at_lambda_one():
    """With every lambda equal to 1 all lambda categories are the uncoupled model
    and their weights sum to one, so the 45-component mixture gives the pair
    log-likelihoods of the 9-component one (closed form of the mixture)."""
    import oracle
    from messi_amd import synth
    a = synth.generate(6, 14, 8)
    ds = synth.dataset_of(a)
    v = a.params.copy()
    v[0:3] = 1.0
    v[14] = 0.35
    p5 = model.getparams(v, ds.tree.branchlengths, siteCats=3, lambdaCats=5)
    p1 = model.getparams(v, ds.tree.branchlengths, siteCats=3, lambdaCats=1)
    M5 = oracle.pair_loglik(ds, model.pair_components(p5))
    M1 = oracle.pair_loglik(ds, model.pair_components(p1))
    fin = np.isfinite(M1)
    assert np.array_equal(np.isfinite(M5), fin)
    assert np.max(np.abs(M5[fin] - M1[fin])) < 1e-12
