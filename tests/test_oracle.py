"""The FP64 CPU oracle, pinned by brute force, closed forms and invariants.

These checks do not retype the oracle's recursions: pruning is checked against
explicit enumeration of internal-node states, the inside/outside/bpp recursions
against enumeration of every nested structure scored by its unique KH99
derivation, plus closed forms (all-unpaired Z, lambda == 1 factorisation,
all-gap columns) and the pulley principle.
"""
import math

import numpy as np
import pytest

import oracle
from messi_amd import model
from messi_amd.dataset import Dataset
from tests.bruteforce import kh99_enumerate, prune_enumerate

LP = {k: math.log(v) for k, v in dict(S_LS=0.868534, S_s=0.117609877998, L_s=0.894603).items()}


def _tiny_dataset(newick, L, seed, gapfrac=0.15, amb=True):
    rng = np.random.default_rng(seed)
    names = sorted(set(t.split(":")[0].strip("(),; ") for t in newick.replace("(", ",").split(",")
                       if t.split(":")[0].strip("(),; ")))
    alphabet = list("ACGU") + (list("RYN") if amb else [])
    seqs = []
    for _ in names:
        s = "".join(rng.choice(alphabet) if rng.random() > gapfrac else "-" for _ in range(L))
        seqs.append(s)
    return Dataset(names, seqs, newick)


def _params(seed, lam=None):
    rng = np.random.default_rng(seed)
    q = rng.dirichlet(np.ones(6)) * 2.0
    pi = rng.dirichlet(np.full(4, 5.0))
    lam = lam if lam is not None else rng.uniform(1.0, 3.0, size=3)
    v = np.array([*lam, *q, 1.0, 1.0, pi[0], pi[1], pi[2], 1e-4])
    return v


TREES = ["(a:0.1,b:0.2,c:0.3);", "((a:0.1,b:0.2):0.05,c:0.3,d:0.15);", "(a:0.4,(b:0.1,c:0.2):0.3,d:0.05);"]


def _leafvecs(ds, cols):
    out = {}
    for v in ds.tree.leaves():
        codes = [ds.codes[ds.tree.seqindex[v], c] for c in cols]
        vecs = [np.array([(c >> a) & 1 for a in range(4)], dtype=float) for c in codes]
        out[v] = vecs[0] if len(vecs) == 1 else np.kron(vecs[0], vecs[1])
    return out


@pytest.mark.parametrize("tree", TREES)
def test_site_loglik_bruteforce(tree):
    ds = _tiny_dataset(tree, 12, 3)
    p = model.getparams(_params(1), ds.tree.branchlengths)
    comps = model.site_components(p)
    u = oracle.site_loglik(ds, comps)
    P = comps[0][0]
    for c in range(ds.numcols):
        bf = prune_enumerate(ds.tree.children, _leafvecs(ds, [c]), P, p.freqs)
        assert abs(u[c] - bf) < 1e-12 * max(1.0, abs(bf))


@pytest.mark.parametrize("tree", TREES[:2])
def test_pair_loglik_bruteforce(tree):
    ds = _tiny_dataset(tree, 6, 4)
    p = model.getparams(_params(2), ds.tree.branchlengths)
    comps = model.pair_components(p)
    M = oracle.pair_loglik(ds, comps)
    P, pi16 = comps[0][0], comps[0][1]
    for x in range(ds.numcols):
        for y in range(x + 1, ds.numcols):
            bf = prune_enumerate(ds.tree.children, _leafvecs(ds, [x, y]), P, pi16)
            assert abs(M[x, y] - bf) < 1e-11 * max(1.0, abs(bf))
            assert M[y, x] == M[x, y]
    assert np.all(np.isneginf(np.diag(M)))


def test_pair_factorises_at_lambda_one():
    from messi_amd import synth
    a = synth.config1()
    ds = synth.dataset_of(a)
    p = model.getparams(_params(3, lam=[1.0, 1.0, 1.0]), ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p), maxbp=40)
    L = ds.numcols
    for x in range(0, L, 7):
        for y in range(x + 1, min(L, x + 41)):
            assert abs(M[x, y] - (u[x] + u[y])) < 1e-10 * abs(M[x, y])
    assert np.isneginf(M[0, 60])        # beyond the distance bound


def test_all_gap_columns_give_zero():
    ds = Dataset(["a", "b", "c", "d"], ["A-G", "C-U", "G-A", "N-C"], TREES[1])
    p = model.getparams(_params(4), ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    assert abs(u[1]) < 1e-13
    ds2 = Dataset(["a", "b", "c", "d"], ["--", "--", "--", "--"], TREES[1])
    M = oracle.pair_loglik(ds2, model.pair_components(p))
    assert abs(M[0, 1]) < 1e-13


def test_pulley_rerooting_invariance():
    t1 = "((a:0.1,b:0.2):0.05,c:0.3,d:0.15);"
    t2 = "(c:0.3,d:0.15,(a:0.1,b:0.2):0.05);"
    t3 = "(a:0.1,b:0.2,(c:0.3,d:0.15):0.05);"
    ds = [_tiny_dataset(t, 10, 9) for t in (t1, t2, t3)]
    # same alignment for all three rootings
    for d in ds[1:]:
        assert d.names == ds[0].names and d.sequences == ds[0].sequences
    vals = []
    for d in ds:
        p = model.getparams(_params(5), d.tree.branchlengths)
        vals.append((oracle.site_loglik(d, model.site_components(p)),
                     oracle.pair_loglik(d, model.pair_components(p))))
    for u, M in vals[1:]:
        assert np.allclose(u, vals[0][0], rtol=1e-12)
        iu = np.triu_indices(10, 1)
        assert np.allclose(M[iu], vals[0][1][iu], rtol=1e-11)


def _random_inputs(L, seed, pfrac=0.5):
    rng = np.random.default_rng(seed)
    u = -rng.uniform(2.0, 6.0, size=L)
    pm = np.full((L, L), -np.inf)
    for i in range(L):
        for j in range(i + 1, L):
            if rng.random() < pfrac:
                pm[i, j] = pm[j, i] = u[i] + u[j] + rng.normal(0.0, 3.0)
    return u, pm


@pytest.mark.parametrize("L,seed", [(1, 0), (2, 1), (5, 2), (8, 3), (10, 4)])
def test_inside_outside_bpp_bruteforce(L, seed):
    u, pm = _random_inputs(L, seed)
    logZ, unp, pp = kh99_enumerate(u, pm)
    ins = oracle.inside(u, pm)
    assert abs(ins[0, 0, L - 1] - logZ) < 1e-12 * max(1, abs(logZ))
    if L < 2:
        return
    out = oracle.outside(ins, u, pm)
    res, pair = oracle.bpp(ins, out, u, pm)
    assert np.allclose(res, unp, atol=1e-12)
    assert np.allclose(pair, pp, atol=1e-12)


def test_inside_B_scaling_bruteforce():
    u, pm = _random_inputs(7, 11)
    B = 0.5
    # exp(B * log weight) for every derivation: scale emissions and rule logs by B.
    import tests.bruteforce as bf
    saved = {k: getattr(bf, k) for k in ("P_S_LS", "P_S_s", "P_S_dFd", "P_F_dFd", "P_F_LS", "P_L_s", "P_L_dFd")}
    try:
        for k, v in saved.items():
            setattr(bf, k, v ** B)
        logZ, _, _ = bf.kh99_enumerate(u * B, pm * B)
    finally:
        for k, v in saved.items():
            setattr(bf, k, v)
    ins = oracle.inside(u, pm, B)
    assert abs(ins[0, 0, 6] - logZ) < 1e-12 * abs(logZ)


@pytest.mark.parametrize("L", [3, 30, 200])
def test_inside_all_unpaired_closed_form(L):
    rng = np.random.default_rng(L)
    u = -rng.uniform(1, 30, size=L)
    pm = np.full((L, L), -np.inf)
    z = oracle.inside(u, pm)[0, 0, L - 1]
    closed = LP["S_s"] + (L - 1) * (LP["S_LS"] + LP["L_s"]) + u.sum()
    assert abs(z - closed) < 1e-12 * abs(closed)


def test_bpp_rows_sum_to_one_on_config1():
    from messi_amd import synth
    ds = synth.dataset_of(synth.config1())
    p = model.getparams(synth.config1().params, ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p))
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    assert np.allclose(res + pair.sum(axis=1), 1.0, atol=1e-12)
    # Before row normalisation the matrix is symmetric and rows already sum to one,
    # so the normalised matrix is symmetric to rounding.
    assert np.allclose(pair, pair.T, atol=1e-12)


@pytest.mark.parametrize("K", [2, 3, 7])
def test_mixture_of_identical_components_is_the_component(K):
    """Closed form of the mixture (src/FARCE.jl:1207-1218, 1780-1792): K copies of
    one component with weights 1/K give that component's site and pair
    log-likelihoods; weights w and 1-w on two copies as well."""
    ds = _tiny_dataset(TREES[1], 9, 17)
    p = model.getparams(_params(5), ds.tree.branchlengths)
    (P4, pi4, _), (P16, pi16, _) = model.site_components(p)[0], model.pair_components(p)[0]
    u1 = oracle.site_loglik(ds, [(P4, pi4, 0.0)])
    M1 = oracle.pair_loglik(ds, [(P16, pi16, 0.0)])
    lw = math.log(1.0 / K)
    uK = oracle.site_loglik(ds, [(P4, pi4, lw)] * K)
    MK = oracle.pair_loglik(ds, [(P16, pi16, lw)] * K)
    assert np.allclose(uK, u1, rtol=0, atol=1e-12)
    fin = np.isfinite(M1)
    assert np.array_equal(np.isfinite(MK), fin)
    assert np.allclose(MK[fin], M1[fin], rtol=0, atol=1e-12)
    w = 0.3
    u2 = oracle.site_loglik(ds, [(P4, pi4, math.log(w)), (P4, pi4, math.log(1 - w))])
    assert np.allclose(u2, u1, rtol=0, atol=1e-12)


def test_mixture_of_two_components_brute_force():
    """A two-component mixture against the mixture of the enumerated per-component
    likelihoods: log(w1 L1 + w2 L2), with L_k by brute force over internal states."""
    ds = _tiny_dataset(TREES[0], 5, 23)
    p1 = model.getparams(_params(6), ds.tree.branchlengths)
    p2 = model.getparams(_params(7), ds.tree.branchlengths)
    c1, c2 = model.site_components(p1)[0], model.site_components(p2)[0]
    w1 = 0.35
    got = oracle.site_loglik(ds, [(c1[0], c1[1], math.log(w1)), (c2[0], c2[1], math.log(1 - w1))])
    for c in range(ds.numcols):
        lv = _leafvecs(ds, [c])
        L1 = math.exp(prune_enumerate(ds.tree.children, lv, c1[0], c1[1]))
        L2 = math.exp(prune_enumerate(ds.tree.children, lv, c2[0], c2[1]))
        assert got[c] == pytest.approx(math.log(w1 * L1 + (1 - w1) * L2), rel=1e-12, abs=1e-12)


def test_oracle_results_do_not_depend_on_threads():
    """Each cell is computed by one thread with the serial arithmetic, so one
    thread and several give the same bits."""
    from messi_amd import synth
    a = synth.generate(12, 160, 5)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    n0 = oracle.get_threads()
    res = []
    try:
        for n in (1, 4):
            oracle.set_threads(n)
            u = oracle.site_loglik(ds, model.site_components(p))
            M = oracle.pair_loglik(ds, model.pair_components(p))
            ins = oracle.inside(u, M)
            out = oracle.outside(ins, u, M)
            res.append((u, M, ins, out) + oracle.bpp(ins, out, u, M))
    finally:
        oracle.set_threads(n0)
    for a1, a2 in zip(*res):
        assert np.array_equal(a1, a2)


def _posterior_inputs(ds, p):
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p))
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    return oracle.bpp(ins, out, u, M), M


def test_site_rate_posteriors_identical_categories_get_equal_weight():
    """Two site-rate categories with the same rate and weight are the same model,
    so every column gives both the same posterior weight."""
    from messi_amd import synth
    a = synth.generate(7, 30, 12)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths, siteCats=2)
    p.siteRates = np.array([0.8, 0.8])
    (res, pair), _ = _posterior_inputs(ds, p)
    w = oracle.site_rate_posteriors(ds, model.site_components(p), model.pair_components(p), 2, res, pair)
    assert np.all(np.isfinite(w))
    assert np.max(np.abs(w[0] - w[1])) < 1e-12


def test_site_rate_posteriors_one_category_is_the_mixture_formula():
    """With one site category the weight of column x is
    log(single[x] L(x) + sum_y pair[min, max] L_pair(x, y)), the pair likelihood
    mixed over the lambda categories (the mixture the pair matrix holds)."""
    from messi_amd import synth
    a = synth.generate(7, 30, 13)
    ds = synth.dataset_of(a)
    v = a.params.copy()
    v[14] = 0.3
    p = model.getparams(v, ds.tree.branchlengths, siteCats=1, lambdaCats=3)
    (res, pair), M = _posterior_inputs(ds, p)
    w = oracle.site_rate_posteriors(ds, model.site_components(p), model.pair_components(p), 1, res, pair)
    u = oracle.site_loglik(ds, model.site_components(p))
    L = ds.numcols
    lpair = np.log(np.where(pair > 0, pair, 1.0)) + np.where(pair > 0, 0.0, -np.inf)
    for x in range(L):
        terms = [np.log(res[x]) + u[x]]
        for y in range(L):
            if y != x:
                i, j = min(x, y), max(x, y)
                terms.append(lpair[i, j] + M[i, j])
        want = np.logaddexp.reduce(np.array(terms))
        assert abs(w[0, x] - want) <= 1e-12 * max(1.0, abs(want))
