"""HIP path versus the FP64 oracle at the full workload sizes, through the C ABI.

* 50 x 1500 (98-branch tree, 1,124,250 pairs): site log-likelihoods in full; the
  pair matrix on 20,000 pairs drawn with every distance equally likely; then the
  oracle's inside, outside and posteriors run on the HIP pair matrix, against
  which the FP64 path is held to 1e-9 relative on log Z and 1e-8 absolute on the
  posteriors, and the FP32 path to 1e-6 relative and 1e-5 absolute.
* 30 x 9000 at full span and at maximum base-pair distance 500, where the
  oracle's O(L^3) recursions would take hours: the pair matrix on a sample;
  every inside cell (j, y) of the 9000-long table inside a window [s, s + W)
  depends only on the positions of the window, so three windows of the GPU's
  table are checked cell by cell against the oracle's inside of the window;
  posterior rows sum to one; FP32 against FP64; the banded run equals the
  full-span recursion on the banded pair matrix (pairing-band skip off).
* All 64 instances of the uneven batch: batched against single evaluations,
  and 8 of them against the oracle.
* maskgapped (src/RNATools.jl:441-451): the gap cutoff rule applied as a device
  pair mask equals the oracle's matrix masked by the rule.
* A mixture of identical components equals the component (closed form).
"""
import sys

import numpy as np
import pytest

import oracle
from messi_amd import autotune, hotpath, lib, model, synth

pytestmark = pytest.mark.gpu

REL64 = 1e-9
BPP64 = 1e-8
BPP32 = 1e-5


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    """Largest |a - b| / max(1, |b|) over the finite entries; -inf patterns must agree."""
    a, b = np.asarray(a), np.asarray(b)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    assert np.array_equal(a[~fin], b[~fin])
    if not fin.any():
        return 0.0
    return float(np.max(np.abs(a[fin] - b[fin]) / np.maximum(1.0, np.abs(b[fin]))))


def _pairs_all_distances(L, n, rng, maxd=None):
    """n column pairs x < y with y - x uniform over 1 .. maxd, then x uniform."""
    maxd = L - 1 if maxd is None else min(maxd, L - 1)
    d = rng.integers(1, maxd + 1, n)
    x = (rng.random(n) * (L - d)).astype(np.int64)
    return x, x + d


def _mask_of(L, x, y):
    m = np.zeros((L, L), dtype=np.uint8, order="F")
    m[x, y] = 1
    return m


def _upper(t, lo, hi):
    """(3, W, W) window [lo, hi) of a (3, L, L) table with the cells j > y set to -inf."""
    w = np.array(t[:, lo:hi, lo:hi])
    tri = np.tril(np.ones((hi - lo, hi - lo), dtype=bool), -1)
    w[:, tri] = -np.inf
    return w


@pytest.fixture(scope="module")
def ctx():
    _say("context create; oracle threads", oracle.get_threads())
    return lib.context()


# ---------------------------------------------------------------------------
# 50 x 1500
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(ctx):
    _say("config2 50 x 1500: generate, site and pair log-likelihoods (fp64)")
    a = synth.config2()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = hotpath.computeunpairedlikelihoods(ds, p)
    M, _ = hotpath.felsensteincuda(ds, p)
    return dict(a=a, ds=ds, p=p, u=u, M=M)


@pytest.fixture(scope="module")
def c2_oracle(c2):
    _say("config2: oracle inside / outside / posteriors on the HIP pair matrix")
    u, M = c2["u"], c2["M"]
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    return dict(ins=ins, out=out, res=res, pair=pair, z=ins[0, 0, -1])


def test_config2_site_and_pairs_over_all_distances(c2):
    ds, p, L = c2["ds"], c2["p"], c2["ds"].numcols
    _say("config2: site log-likelihoods and 20,000 pairs over all distances vs oracle")
    assert _rel(c2["u"], oracle.site_loglik(ds, model.site_components(p))) < 1e-12
    M = c2["M"]
    assert np.array_equal(M, M.T)
    assert np.all(np.diag(M) == -np.inf)
    x, y = _pairs_all_distances(L, 20000, np.random.default_rng(7))
    assert np.unique(y - x).size > 1300      # nearly every distance is present
    mask = _mask_of(L, x, y)
    ref = oracle.pair_loglik(ds, model.pair_components(p), mask=mask)
    got, want = M[x, y], ref[x, y]
    assert np.all(np.isfinite(want))
    err = float(np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))))
    _say(f"config2 sampled pairs: max rel err {err:.2e}")
    assert err < 1e-12
    # the same pairs through a masked device dataset: those entries, -inf elsewhere
    Mm, _ = hotpath.felsensteincuda(ds, p, mask=mask)
    allowed = (mask > 0) | (mask.T > 0)
    assert np.array_equal(Mm[allowed], M[allowed])
    assert np.all(Mm[~allowed] == -np.inf)


def test_config2_fp64_posterior_matches_oracle(ctx, c2, c2_oracle):
    ds, p, u, M, o = c2["ds"], c2["p"], c2["u"], c2["M"], c2_oracle
    zr = o["z"]
    _say("config2: fused fp64 evaluation and posterior vs oracle")
    z = hotpath.loglikelihood(ds, p)
    assert abs(z - zr) <= REL64 * abs(zr), (z, zr)
    z2, r, pp = hotpath.posterior(ds, p)
    er = max(np.max(np.abs(r - o["res"])), np.max(np.abs(pp - o["pair"])))
    _say(f"config2 fp64: logZ {z2:.9f} oracle {zr:.9f} rel {abs(z2 - zr) / abs(zr):.2e}; posterior err {er:.2e}")
    assert abs(z2 - zr) <= REL64 * abs(zr)
    assert er < BPP64
    assert np.max(np.abs(r + pp.sum(axis=1) - 1.0)) < 1e-10


def test_config2_fp64_tables_match_oracle(ctx, c2, c2_oracle):
    u, M, o = c2["u"], c2["M"], c2_oracle
    _say("config2: standalone fp64 inside / outside / bpp tables vs oracle")
    t = hotpath.computeinsideKH99(u, M, 1.0, False)
    ei = _rel(t, o["ins"])
    out = hotpath.computeoutsideKH99(o["ins"], u, M)
    eo = _rel(out, o["out"])
    res, pair = hotpath.computebasepairprobs(o["ins"], o["out"], u, M)
    eb = max(np.max(np.abs(res - o["res"])), np.max(np.abs(pair - o["pair"])))
    _say(f"config2 tables: inside {ei:.2e} outside {eo:.2e} bpp {eb:.2e}")
    assert ei < 1e-11 and eo < 1e-10 and eb < BPP64


def test_config2_fp32_error_against_oracle(ctx, c2, c2_oracle):
    ds, p, o = c2["ds"], c2["p"], c2_oracle
    _say("config2: fp32 posterior vs the fp64 oracle")
    z32, r32, p32 = hotpath.posterior(ds, p, prec="fp32")
    ez = abs(z32 - o["z"]) / abs(o["z"])
    eb = max(np.max(np.abs(r32 - o["res"])), np.max(np.abs(p32 - o["pair"])))
    _say(f"config2 fp32: logZ rel err {ez:.2e}, posterior abs err {eb:.2e}")
    assert ez < 1e-6
    assert eb < BPP32


# ---------------------------------------------------------------------------
# 30 x 9000
# ---------------------------------------------------------------------------
WINDOWS = ((0, 1200), (4000, 5200), (7800, 9000))


@pytest.fixture(scope="module")
def c3(ctx):
    _say("config3 30 x 9000: generate, dataset, site and full pair matrix (fp64)")
    a = synth.config3()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = hotpath.computeunpairedlikelihoods(ds, p)
    M, _ = hotpath.felsensteincuda(ds, p)
    return dict(a=a, ds=ds, p=p, u=u, M=M)


def test_config3_site_and_sampled_pairs(c3):
    ds, p, u, M, L = c3["ds"], c3["p"], c3["u"], c3["M"], c3["ds"].numcols
    _say("config3: site log-likelihoods and sampled pairs vs oracle")
    assert L == 9000
    assert _rel(u, oracle.site_loglik(ds, model.site_components(p))) < 1e-12
    rng = np.random.default_rng(9)
    x, y = _pairs_all_distances(L, 20000, rng)
    xb, yb = _pairs_all_distances(L, 5000, rng, maxd=500)
    x, y = np.concatenate([x, xb]), np.concatenate([y, yb])
    ref = oracle.pair_loglik(ds, model.pair_components(p), mask=_mask_of(L, x, y))
    err = float(np.max(np.abs(M[x, y] - ref[x, y]) / np.maximum(1.0, np.abs(ref[x, y]))))
    _say(f"config3 sampled pairs: max rel err {err:.2e}")
    assert err < 1e-12
    assert np.array_equal(M[x, y], M[y, x])
    # the banded dataset returns the same scores inside the band and -inf beyond it
    Mb, _ = hotpath.felsensteincuda(ds, p, maxbasepairdistance=500)
    d = np.abs(np.subtract.outer(np.arange(L, dtype=np.int32), np.arange(L, dtype=np.int32)))
    band = d <= 500
    del d
    assert np.array_equal(Mb[band], M[band])
    assert np.all(Mb[~band] == -np.inf)
    c3["Mb"] = Mb


def test_config3_full_span_posterior(ctx, c3):
    ds, p = c3["ds"], c3["p"]
    _say("config3: full-span posteriors fp64 and fp32")
    z64, r64, p64 = hotpath.posterior(ds, p)
    rs = float(np.max(np.abs(r64 + p64.sum(axis=1) - 1.0)))
    z32, r32, p32 = hotpath.posterior(ds, p, prec="fp32")
    ez = abs(z32 - z64) / abs(z64)
    eb = max(float(np.max(np.abs(r32 - r64))), float(np.max(np.abs(p32 - p64))))
    _say(f"config3 full span: logZ64 {z64:.6f} logZ32 {z32:.6f} rel {ez:.2e}; posterior err {eb:.2e}; "
         f"row sums {rs:.2e}")
    assert rs < 1e-10
    assert ez < 1e-6
    assert eb < BPP32
    assert np.max(np.abs(r32 + p32.sum(axis=1) - 1.0)) < 1e-5
    c3["z64"] = z64


def test_config3_inside_table_windows_match_oracle(ctx, c3):
    u, M = c3["u"], c3["M"]
    _say("config3: full-span inside table, three 1200-wide windows vs the oracle")
    t = hotpath.computeinsideKH99(u, M, 1.0, False)
    if "z64" in c3:
        assert abs(t[0, 0, -1] - c3["z64"]) <= 1e-12 * abs(c3["z64"])
    for lo, hi in WINDOWS:
        ref = oracle.inside(u[lo:hi], M[lo:hi, lo:hi])
        e = _rel(_upper(t, lo, hi), ref)
        _say(f"config3 inside window [{lo}, {hi}): max rel err {e:.2e}")
        assert e < 1e-11
    del t


def test_config3_band500_equals_full_recursion(ctx, c3):
    ds, p, u = c3["ds"], c3["p"], c3["u"]
    Mb = c3.get("Mb")
    if Mb is None:
        Mb, _ = hotpath.felsensteincuda(ds, p, maxbasepairdistance=500)
    _say("config3 D=500: band skip on against off (full-span recursion on the banded matrix)")
    zb, rb, pb = hotpath.posterior(ds, p, maxbasepairdistance=500)
    ctx.set_band_skip(False)
    try:
        zf, rf, pf = hotpath.posterior(ds, p, maxbasepairdistance=500)
        zf32, rf32, pf32 = hotpath.posterior(ds, p, maxbasepairdistance=500, prec="fp32")
    finally:
        ctx.set_band_skip(True)
    ef = max(float(np.max(np.abs(rb - rf))), float(np.max(np.abs(pb - pf))))
    _say(f"config3 D=500 fp64: logZ {zb:.9f} vs full recursion {zf:.9f}; posterior diff {ef:.2e}")
    assert abs(zb - zf) <= 1e-12 * abs(zf)
    assert ef < 1e-12
    assert np.max(np.abs(rb + pb.sum(axis=1) - 1.0)) < 1e-10
    zb32, rb32, pb32 = hotpath.posterior(ds, p, maxbasepairdistance=500, prec="fp32")
    assert abs(zb32 - zf32) <= 1e-6 * abs(zf32)
    ez = abs(zb32 - zb) / abs(zb)
    eb = max(float(np.max(np.abs(rb32 - rb))), float(np.max(np.abs(pb32 - pb))))
    _say(f"config3 D=500 fp32 vs fp64: logZ rel {ez:.2e}, posterior err {eb:.2e}")
    assert ez < 1e-6 and eb < BPP32
    del rb, pb, rf, pf, rb32, pb32, rf32, pf32
    # the banded standalone inside table (band detected from the matrix) on windows
    t = hotpath.computeinsideKH99(u, Mb, 1.0, False)
    assert abs(t[0, 0, -1] - zb) <= 1e-12 * abs(zb)
    for lo, hi in WINDOWS[:2]:
        ref = oracle.inside(u[lo:hi], Mb[lo:hi, lo:hi])
        e = _rel(_upper(t, lo, hi), ref)
        _say(f"config3 D=500 inside window [{lo}, {hi}): max rel err {e:.2e}")
        assert e < 1e-11


# ---------------------------------------------------------------------------
# 64 uneven instances
# ---------------------------------------------------------------------------
def _stack_params(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


def test_config4_all_64_instances(ctx):
    _say("config4: 64 uneven instances, batch vs single vs oracle")
    als = synth.config4()
    assert len(als) == 64
    dss = [synth.dataset_of(a) for a in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    ps = [model.getparams(a.params, d.tree.branchlengths) for a, d in zip(als, dss)]
    sp = [_stack_params(p) for p in ps]
    Ls = np.array([d.numcols for d in dss])
    b = lib.Batch(dds)
    try:
        zs = {}
        for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
            z = b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, prec)
            single = np.array([dd.eval(*x, 1.0, prec) for dd, x in zip(dds, sp)])
            assert np.max(np.abs(z - single) / np.abs(single)) <= tol, prec
            b.load([x[0] for x in sp], [x[1] for x in sp], prec)
            ch = autotune.tune_batch(b, prec, reps=1)
            z2 = b.eval_loaded(1.0, prec)
            assert np.max(np.abs(z2 - z) / np.abs(z)) <= tol, (prec, ch["threshold"])
            zs[prec] = z
        # the oracle on the four shortest and the four longest members
        order = np.argsort(Ls)
        for m in list(order[:4]) + list(order[-4:]):
            u = oracle.site_loglik(dss[m], model.site_components(ps[m]))
            M = oracle.pair_loglik(dss[m], model.pair_components(ps[m]))
            zr = oracle.inside(u, M)[0, 0, -1]
            assert abs(zs["fp64"][m] - zr) <= REL64 * abs(zr), m
            assert abs(zs["fp32"][m] - zr) <= 1e-6 * abs(zr), m
    finally:
        b.close()


# ---------------------------------------------------------------------------
# gap mask and mixtures
# ---------------------------------------------------------------------------
def test_maskgapped_cutoff_rule_on_device(ctx):
    _say("maskgapped cutoff rule as a device pair mask")
    a = synth.config1()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    g = ds.gapfrequency
    vals = np.unique(g)
    assert vals.size >= 3
    cutoff = float(vals[1])              # an attained value: columns at exactly the cutoff are masked
    assert np.any(g == cutoff) and np.any(g < cutoff) and np.any(g > cutoff)
    ref = oracle.pair_loglik(ds, model.pair_components(p))
    hotpath.maskgapped(ref, g, cutoff, -np.inf)
    M = hotpath.felsensteincuda_gapmasked(ds, p, cutoff)
    assert _rel(M, ref) < 1e-12
    bad = g >= cutoff
    assert np.all(M[bad, :] == -np.inf) and np.all(M[:, bad] == -np.inf)
    ok = np.flatnonzero(~bad)
    assert np.all(np.isfinite(M[np.ix_(ok, ok)][~np.eye(ok.size, dtype=bool)]))
    # the ML objective's log Z on the masked matrix (src/MESSI.jl:305-311)
    u = oracle.site_loglik(ds, model.site_components(p))
    zr = oracle.inside(u, ref)[0, 0, -1]
    z = hotpath.computeinsideKH99(hotpath.computeunpairedlikelihoods(ds, p), M)
    assert abs(z - zr) <= REL64 * abs(zr)
    # a structure keeps only the pairs whose columns are both below the cutoff
    st = hotpath.maskgapped_structure(a.pairedsites + 1, g, cutoff)   # 1-based partners, 0 unpaired
    i = np.flatnonzero(st > np.arange(1, len(st) + 1))
    assert np.all(g[i] < cutoff) and np.all(g[st[i] - 1] < cutoff)


def test_mixture_of_identical_components_is_the_component(ctx):
    """K identical components with weights summing to one give the single
    component's site and pair log-likelihoods and log Z (closed form)."""
    _say("mixture of identical components")
    a = synth.generate(10, 60, 41)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    dd = hotpath.device_dataset(ds)
    (P4, pi4, lw4), (P16, pi16, lw16) = _stack_params(p)
    assert P4.shape[0] == 1 and P16.shape[0] == 1
    for K in (2, 3, 5):
        w = np.log(np.full(K, 1.0 / K))
        for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
            u1 = dd.site_loglik(P4, pi4, lw4, prec)
            uK = dd.site_loglik(np.repeat(P4, K, 0), np.repeat(pi4, K, 0), w, prec)
            assert np.max(np.abs(uK - u1) / np.maximum(1, np.abs(u1))) < tol
            M1 = dd.pair_loglik(P16, pi16, lw16, prec)
            MK = dd.pair_loglik(np.repeat(P16, K, 0), np.repeat(pi16, K, 0), w, prec)
            assert _rel(MK, M1) < tol
            z1 = dd.eval((P4, pi4, lw4), (P16, pi16, lw16), 1.0, prec)
            zK = dd.eval((np.repeat(P4, K, 0), np.repeat(pi4, K, 0), w),
                         (np.repeat(P16, K, 0), np.repeat(pi16, K, 0), w), 1.0, prec)
            assert abs(zK - z1) <= tol * abs(z1)


@pytest.mark.parametrize("siteCats,lambdaCats", [(3, 5), (2, 3)])
def test_site_and_lambda_category_mixture_matches_oracle(ctx, siteCats, lambdaCats):
    """The reference CLI's default mixture (3 site-rate categories, 5 lambda
    categories: 3 site and 45 pair components) on the 20 x 120 workload: site and
    pair log-likelihoods, log Z and posteriors against the oracle."""
    _say(f"mixture siteCats={siteCats} lambdaCats={lambdaCats}")
    a = synth.config1()
    ds = synth.dataset_of(a)
    v = a.params.copy()
    v[14] = 0.3                      # lambda-zero weight
    p = model.getparams(v, ds.tree.branchlengths, siteCats=siteCats, lambdaCats=lambdaCats)
    sc, pc = model.site_components(p), model.pair_components(p)
    assert len(sc) == siteCats and len(pc) == siteCats ** 2 * lambdaCats
    u = oracle.site_loglik(ds, sc)
    M = oracle.pair_loglik(ds, pc)
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    zr = ins[0, 0, -1]
    assert _rel(hotpath.computeunpairedlikelihoods(ds, p), u) < 1e-12
    Mg, per = hotpath.felsensteincuda(ds, p, keepmuseconditionals=True)
    assert _rel(Mg, M) < 1e-12
    assert len(per) == len(pc)
    z, r, pp = hotpath.posterior(ds, p)
    assert abs(z - zr) <= REL64 * abs(zr)
    assert max(np.max(np.abs(r - res)), np.max(np.abs(pp - pair))) < BPP64
    z32, r32, p32 = hotpath.posterior(ds, p, prec="fp32")
    assert abs(z32 - zr) <= 1e-6 * abs(zr)
    assert max(np.max(np.abs(r32 - res)), np.max(np.abs(p32 - pair))) < BPP32


@pytest.mark.parametrize("siteCats,lambdaCats", [(3, 1), (2, 3), (1, 3)])
def test_site_rate_posteriors_match_oracle(ctx, siteCats, lambdaCats):
    """felsensteinposteriorsiterateconditionalscuda on the device against the
    oracle's restatement, from the oracle's own posteriors (20 x 120)."""
    _say(f"site-rate posteriors siteCats={siteCats} lambdaCats={lambdaCats}")
    a = synth.config1()
    ds = synth.dataset_of(a)
    v = a.params.copy()
    v[14] = 0.3
    p = model.getparams(v, ds.tree.branchlengths, siteCats=siteCats, lambdaCats=lambdaCats)
    sc, pc = model.site_components(p), model.pair_components(p)
    u = oracle.site_loglik(ds, sc)
    M = oracle.pair_loglik(ds, pc)
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    want = oracle.site_rate_posteriors(ds, sc, pc, siteCats, res, pair)
    got = hotpath.felsensteinposteriorsiterateconditionalscuda(res, pair, ds, p)
    assert got.shape == (siteCats, ds.numcols)
    assert _rel(got, want) < 1e-12
    got32 = hotpath.felsensteinposteriorsiterateconditionalscuda(res, pair, ds, p, prec="fp32")
    assert _rel(got32, want) < 1e-6
    # normalised over the categories: a distribution per column
    w = np.exp(got - np.logaddexp.reduce(got, axis=0))
    assert np.allclose(w.sum(axis=0), 1.0, atol=1e-12)
