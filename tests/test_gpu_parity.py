"""HIP path versus the FP64 oracle, through the C ABI.

FP64 bar (from the task): log-likelihoods within 1e-9 relative, pairing
probabilities within 1e-8 absolute.  The FP32 path reports its error against
the oracle and is held to the bounds written in each test.
"""
import sys

import numpy as np
import pytest

import oracle
from messi_amd import hotpath, lib, model, synth
from messi_amd.dataset import Dataset

pytestmark = pytest.mark.gpu

REL64 = 1e-9
BPP64 = 1e-8


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    _say("context create")
    return lib.context()


@pytest.fixture(scope="module")
def c1():
    a = synth.config1()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p))
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    return dict(ds=ds, p=p, u=u, M=M, ins=ins, out=out, res=res, pair=pair)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    assert np.array_equal(a[~fin], b[~fin])
    if not fin.any():
        return 0.0
    return float(np.max(np.abs(a[fin] - b[fin]) / np.maximum(1.0, np.abs(b[fin]))))


def test_selftest(ctx):
    ctx.selftest()


def test_site_loglik_fp64(ctx, c1):
    _say("site loglik fp64 config1")
    u = hotpath.computeunpairedlikelihoods(c1["ds"], c1["p"])
    assert _rel(u, c1["u"]) < 1e-12


def test_pair_loglik_fp64(ctx, c1):
    _say("pair loglik fp64 config1")
    M, _ = hotpath.felsensteincuda(c1["ds"], c1["p"])
    assert _rel(M, c1["M"]) < 1e-12


def test_inside_fp64(ctx, c1):
    _say("inside fp64 config1")
    z = hotpath.computeinsideKH99(c1["u"], c1["M"])
    zr = c1["ins"][0, 0, -1]
    assert abs(z - zr) <= REL64 * abs(zr)
    t = hotpath.computeinsideKH99(c1["u"], c1["M"], 1.0, False)
    assert _rel(t, c1["ins"]) < 1e-11


def test_outside_bpp_fp64(ctx, c1):
    _say("outside + bpp fp64 config1")
    out = hotpath.computeoutsideKH99(c1["ins"], c1["u"], c1["M"])
    assert _rel(out, c1["out"]) < 1e-10
    res, pair = hotpath.computebasepairprobs(c1["ins"], c1["out"], c1["u"], c1["M"])
    assert np.max(np.abs(res - c1["res"])) < BPP64
    assert np.max(np.abs(pair - c1["pair"])) < BPP64


def test_eval_and_posterior_fp64(ctx, c1):
    _say("eval + posterior fp64 config1")
    z = hotpath.loglikelihood(c1["ds"], c1["p"])
    zr = c1["ins"][0, 0, -1]
    assert abs(z - zr) <= REL64 * abs(zr)
    z2, res, pair = hotpath.posterior(c1["ds"], c1["p"])
    assert abs(z2 - zr) <= REL64 * abs(zr)
    assert np.max(np.abs(res - c1["res"])) < BPP64
    assert np.max(np.abs(pair - c1["pair"])) < BPP64


def test_fp32_path_error_report(ctx, c1):
    _say("fp32 path config1")
    u = hotpath.computeunpairedlikelihoods(c1["ds"], c1["p"], prec="fp32")
    M, _ = hotpath.felsensteincuda(c1["ds"], c1["p"], prec="fp32")
    zr = c1["ins"][0, 0, -1]
    z, res, pair = hotpath.posterior(c1["ds"], c1["p"], prec="fp32")
    eu, eM = _rel(u, c1["u"]), _rel(M, c1["M"])
    ez = abs(z - zr) / abs(zr)
    eb = max(np.max(np.abs(res - c1["res"])), np.max(np.abs(pair - c1["pair"])))
    _say(f"fp32 errors: site {eu:.2e} pair {eM:.2e} logZ(rel) {ez:.2e} bpp(abs) {eb:.2e}")
    assert eu < 1e-6 and eM < 1e-6
    assert ez < 1e-6
    assert eb < 1e-5


def test_mixture_components_match_oracle(ctx):
    _say("K>1 mixture fp64")
    a = synth.generate(8, 40, 11)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths, siteCats=3)
    sc, pc = model.site_components(p), model.pair_components(p)
    assert len(sc) == 3 and len(pc) == 9
    u = hotpath.computeunpairedlikelihoods(ds, p)
    assert _rel(u, oracle.site_loglik(ds, sc)) < 1e-12
    M, per = hotpath.felsensteincuda(ds, p, keepmuseconditionals=True)
    assert _rel(M, oracle.pair_loglik(ds, pc)) < 1e-12
    for k in (0, 4, 8):
        assert _rel(per[k], oracle.pair_loglik(ds, [(pc[k][0], pc[k][1], 0.0)])) < 1e-12


@pytest.mark.parametrize("maxbp", [1, 7, 1000000])
def test_distance_bound_and_mask(ctx, maxbp):
    _say(f"distance bound {maxbp}")
    a = synth.generate(6, 30, 12)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    M, _ = hotpath.felsensteincuda(ds, p, maxbasepairdistance=maxbp)
    assert _rel(M, oracle.pair_loglik(ds, model.pair_components(p), maxbp=maxbp)) < 1e-12
    rng = np.random.default_rng(1)
    mask = (rng.random((30, 30)) < 0.5).astype(np.uint8)
    M2, _ = hotpath.felsensteincuda(ds, p, maxbasepairdistance=maxbp, mask=mask)
    assert _rel(M2, oracle.pair_loglik(ds, model.pair_components(p), maxbp=maxbp, mask=mask)) < 1e-12


@pytest.mark.parametrize("L", [1, 2, 3, 9])
def test_small_and_degenerate_lengths(ctx, L):
    _say(f"inside/outside L={L}")
    rng = np.random.default_rng(L)
    u = -rng.uniform(2, 6, size=L)
    pm = np.full((L, L), -np.inf)
    for i in range(L):
        for j in range(i + 1, L):
            pm[i, j] = pm[j, i] = u[i] + u[j] + rng.normal(0, 3)
    ins = oracle.inside(u, pm)
    z = hotpath.computeinsideKH99(u, pm)
    assert abs(z - ins[0, 0, -1]) <= 1e-12 * abs(ins[0, 0, -1])
    if L >= 2:
        out = oracle.outside(ins, u, pm)
        assert _rel(hotpath.computeoutsideKH99(ins, u, pm), out) < 1e-11
        res, pair = oracle.bpp(ins, out, u, pm)
        r2, p2 = hotpath.computebasepairprobs(ins, out, u, pm)
        assert np.max(np.abs(r2 - res)) < BPP64 and np.max(np.abs(p2 - pair)) < BPP64


def test_all_unpaired_closed_form(ctx):
    _say("all-unpaired closed form L=700")
    L = 700
    rng = np.random.default_rng(3)
    u = -rng.uniform(1, 40, size=L)
    pm = np.full((L, L), -np.inf)
    closed = np.log(0.117609877998) + (L - 1) * (np.log(0.868534) + np.log(0.894603)) + u.sum()
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        z = hotpath.computeinsideKH99(u, pm, prec=prec)
        assert abs(z - closed) <= tol * abs(closed), prec


@pytest.mark.parametrize("score", [300.0, 900.0, -900.0])
def test_fp32_extreme_pair_scores(ctx, score):
    """FP32 inside table against the oracle when helices carry pair scores of
    +-300 and +-900 nats: the in-tile wavefront of the scaled path runs in double
    precision at one scale per tile and hands a tile to its per-cell-exponent form
    when a value would leave [2^-1000, 2^1000] (a single +-900 nat pair term is
    2^+-1298), so both forms are exercised; both must match the oracle."""
    _say(f"fp32 extreme pair scores {score}")
    L = 300
    rng = np.random.default_rng(11)
    u = -rng.uniform(2, 6, size=L)
    pm = np.full((L, L), -np.inf)
    for i in range(L):
        for j in range(i + 1, L):
            pm[i, j] = pm[j, i] = u[i] + u[j] + rng.normal(-2, 2)
    for (i0, j0, n) in ((20, 170, 10), (181, 260, 6)):  # two helices with extreme pair terms
        for k in range(n):
            pm[i0 + k, j0 - k] = pm[j0 - k, i0 + k] = u[i0 + k] + u[j0 - k] + score
    ins = oracle.inside(u, pm)
    for tile in (16, 32):
        ctx.set_tile(tile)
        try:
            z, t = lib.inside(u, pm, zonly=False, prec="fp32")
        finally:
            ctx.set_tile(16)
        zr = ins[0, 0, -1]
        assert abs(z - zr) <= 1e-6 * abs(zr), (tile, z, zr)
        _rel(t, ins)  # the same cells are -inf
        fin = np.isfinite(ins)
        err = float(np.max(np.abs(t[fin] - ins[fin])))
        _say(f"tile {tile}: log Z rel err {abs(z - zr) / abs(zr):.2e}, max table log err {err:.2e}")
        assert err < 1e-4, tile
    # the outside pass from the oracle's inside table (its in-tile wavefront has the
    # same double / per-cell-exponent split) and the posteriors
    out = oracle.outside(ins, u, pm)
    res, pair = oracle.bpp(ins, out, u, pm)
    o32 = lib.outside(ins, u, pm, prec="fp32")
    _rel(o32, out)
    fin = np.isfinite(out)
    oerr = float(np.max(np.abs(o32[fin] - out[fin])))
    r32, p32 = lib.bpp(ins, o32, u, pm, prec="fp32")
    berr = max(float(np.max(np.abs(r32 - res))), float(np.max(np.abs(p32 - pair))))
    _say(f"outside max log err {oerr:.2e}, posterior max abs err {berr:.2e}")
    assert oerr < 1e-4 and berr < 1e-5


def test_config2_scale_properties(ctx):
    """Full 1,500 nt x 50 sequence workload: dedup consistency against a banded
    run, symmetry, posterior row sums and FP32 versus FP64 agreement."""
    _say("config2 properties")
    a = synth.config2()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = hotpath.computeunpairedlikelihoods(ds, p)
    assert _rel(u, oracle.site_loglik(ds, model.site_components(p))) < 1e-12
    M, _ = hotpath.felsensteincuda(ds, p)
    assert np.array_equal(M, M.T)
    Mb, _ = hotpath.felsensteincuda(ds, p, maxbasepairdistance=12)
    ref = oracle.pair_loglik(ds, model.pair_components(p), maxbp=12)
    assert _rel(Mb, ref) < 1e-12
    band = np.abs(np.subtract.outer(np.arange(1500), np.arange(1500))) <= 12
    assert _rel(np.where(band, M, -np.inf), ref) < 1e-12
    z64, r64, p64 = hotpath.posterior(ds, p)
    assert np.max(np.abs(r64 + p64.sum(axis=1) - 1.0)) < 1e-10
    z32, r32, p32 = hotpath.posterior(ds, p, prec="fp32")
    err = max(np.max(np.abs(r32 - r64)), np.max(np.abs(p32 - p64)))
    _say(f"config2 logZ64 {z64:.6f} logZ32 {z32:.6f} rel {abs(z32 - z64) / abs(z64):.2e} bpp err {err:.2e}")
    assert abs(z32 - z64) <= 1e-6 * abs(z64)
    assert err < 1e-5


def _perturbed(a, ds, k, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        v = a.params.copy()
        v[0:3] = np.maximum(1.0, v[0:3] * np.exp(rng.normal(0, 0.2, 3)))
        v[3:9] = v[3:9] * np.exp(rng.normal(0, 0.2, 6))
        out.append(model.getparams(v, ds.tree.branchlengths))
    return out


def _stack_params(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_same_dataset_matches_single(ctx, prec):
    _say(f"batch of 5 parameter vectors, 20 x 120, {prec}")
    a = synth.config1()
    ds = synth.dataset_of(a)
    dd = hotpath.device_dataset(ds)
    ps = _perturbed(a, ds, 5)
    single = [dd.eval(*_stack_params(p), 1.0, prec) for p in ps]
    b = lib.Batch([dd] * 5)
    sp = [_stack_params(p) for p in ps]
    z = b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, prec)
    assert np.allclose(z, single, rtol=1e-12, atol=0)
    if prec == "fp64":
        for p, zz in zip(ps[:2], z[:2]):
            u = oracle.site_loglik(ds, model.site_components(p))
            M = oracle.pair_loglik(ds, model.pair_components(p))
            zr = oracle.inside(u, M)[0, 0, -1]
            assert abs(zz - zr) <= REL64 * abs(zr)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_runs_of_shared_datasets(ctx, prec):
    """Members on one dataset that follow each other share their pair-term lookups
    (one pass per run); runs broken by other datasets or another distance bound
    must give each member exactly its single evaluation's log Z."""
    _say(f"batch with runs of shared datasets, {prec}")
    a1, a2 = synth.config1(), synth.generate(16, 90, 23)
    d1, d2 = synth.dataset_of(a1), synth.dataset_of(a2)
    dd1, dd2, dd1b = hotpath.device_dataset(d1), hotpath.device_dataset(d2), hotpath.device_dataset(d1, 25)
    p1, p2 = _perturbed(a1, d1, 4, seed=5), _perturbed(a2, d2, 2, seed=6)
    members = [(dd1, p1[0]), (dd1, p1[1]), (dd2, p2[0]), (dd1, p1[2]), (dd1b, p1[3]), (dd1b, p1[0]),
               (dd2, p2[1]), (dd2, p2[0])]
    sp = [_stack_params(p) for _, p in members]
    b = lib.Batch([d for d, _ in members])
    try:
        z = b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, prec)
    finally:
        b.close()
    single = [d.eval(*x, 1.0, prec) for (d, _), x in zip(members, sp)]
    assert np.array_equal(z, np.array(single)), (z, single)


def test_batch_uneven_instances(ctx):
    _say("batch of 6 uneven instances (config4 generator)")
    als = synth.config4(count=6)
    dss = [synth.dataset_of(a) for a in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    ps = [model.getparams(a.params, d.tree.branchlengths) for a, d in zip(als, dss)]
    sp = [_stack_params(p) for p in ps]
    b = lib.Batch(dds)
    for prec in ("fp64", "fp32"):
        z = b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, prec)
        single = [dd.eval(*x, 1.0, prec) for dd, x in zip(dds, sp)]
        assert np.allclose(z, single, rtol=1e-12, atol=0), prec
    # anchor one instance on the oracle
    d0, p0 = dss[0], ps[0]
    u = oracle.site_loglik(d0, model.site_components(p0))
    M = oracle.pair_loglik(d0, model.pair_components(p0))
    zr = oracle.inside(u, M)[0, 0, -1]
    z = b.eval([x[0] for x in sp], [x[1] for x in sp], 1.0, "fp64")
    assert abs(z[0] - zr) <= REL64 * abs(zr)


def test_tile32_matches_tile16(ctx):
    _say("inside tile 32 vs 16")
    a = synth.generate(10, 300, 21)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    dd = hotpath.device_dataset(ds)
    z16 = [dd.eval(*_stack_params(p), 1.0, prec) for prec in ("fp64", "fp32")]
    ctx.set_tile(32)
    try:
        z32 = [dd.eval(*_stack_params(p), 1.0, prec) for prec in ("fp64", "fp32")]
    finally:
        ctx.set_tile(16)
    assert abs(z32[0] - z16[0]) <= 1e-12 * abs(z16[0])
    assert abs(z32[1] - z16[1]) <= 1e-6 * abs(z16[1])


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("maxbp", [0, 5, 17, 40])
def test_pairing_band_posterior(ctx, tile, maxbp):
    """With a distance bound the inside pass skips out-of-band split-point tiles
    and the outside pass the out-of-band sums; results equal the oracle's full
    recursions on the same bounded pair matrix."""
    _say(f"band posterior maxbp={maxbp} tile={tile}")
    a = synth.generate(12, 150, 31)
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p), maxbp=maxbp)
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    zr = ins[0, 0, -1]
    ctx.set_tile(tile)
    try:
        z, r2, p2 = hotpath.posterior(ds, p, maxbasepairdistance=maxbp)
        assert abs(z - zr) <= REL64 * abs(zr)
        assert np.max(np.abs(r2 - res)) < BPP64 and np.max(np.abs(p2 - pair)) < BPP64
        z32, r32, p32 = hotpath.posterior(ds, p, prec="fp32", maxbasepairdistance=maxbp)
        assert abs(z32 - zr) <= 1e-6 * abs(zr)
        assert max(np.max(np.abs(r32 - res)), np.max(np.abs(p32 - pair))) < 1e-5
        # standalone entry points detect the band from the pair matrix
        t = hotpath.computeinsideKH99(u, M, 1.0, False)
        assert _rel(t, ins) < 1e-11
        assert _rel(hotpath.computeoutsideKH99(ins, u, M), out) < 1e-10
    finally:
        ctx.set_tile(16)


def test_batch_load_then_eval_loaded(ctx):
    """messi_batch_load + messi_batch_eval_loaded equal messi_batch_eval; a
    precision other than the loaded one is an error."""
    _say("batch load / eval_loaded")
    a = synth.config1()
    ds = synth.dataset_of(a)
    dd = hotpath.device_dataset(ds)
    ps = [model.getparams(a.params * (1 + 0.02 * i), ds.tree.branchlengths) for i in range(3)]
    comps = [_stack_params(p) for p in ps]
    b = lib.Batch([dd] * 3)
    try:
        for prec in ("fp64", "fp32"):
            want = b.eval([c[0] for c in comps], [c[1] for c in comps], 1.0, prec)
            b.load([c[0] for c in comps], [c[1] for c in comps], prec)
            z1 = b.eval_loaded(1.0, prec)
            z2 = b.eval_loaded(1.0, prec)
            assert np.array_equal(z1, want) and np.array_equal(z2, want)
        with pytest.raises(lib.MessiError):
            b.eval_loaded(1.0, "fp64")      # fp32 components are loaded
    finally:
        b.close()


def test_batch_mixed_tiles_and_autotuner(ctx):
    """Members split over tile-16 and tile-32 launch sequences give the same log Z
    as one sequence; the autotuner leaves a valid assignment behind."""
    from messi_amd import autotune
    _say("batch with mixed tiles + autotuner")
    als = [synth.generate(8 + 3 * i, 90 + 70 * i, 50 + i) for i in range(5)]
    dss = [synth.dataset_of(a) for a in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    comps = [_stack_params(model.getparams(a.params, d.tree.branchlengths)) for a, d in zip(als, dss)]
    b = lib.Batch(dds)
    try:
        for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
            b.load([c[0] for c in comps], [c[1] for c in comps], prec)
            b.set_tiles(None)
            ref = b.eval_loaded(1.0, prec)
            b.set_tiles([16, 32, 16, 32, 32])
            for _ in range(3):   # each tile-32 evaluation, not only the first
                z = b.eval_loaded(1.0, prec)
                assert np.max(np.abs(z - ref) / np.abs(ref)) < tol
            ch = autotune.tune_batch(b, prec, reps=1)
            assert len(ch["candidates"]) >= 2
            z2 = b.eval_loaded(1.0, prec)
            assert np.max(np.abs(z2 - ref) / np.abs(ref)) < tol
    finally:
        b.close()


def test_batch_member_groups_are_bit_identical(ctx):
    """A pipelined evaluation (member groups on their own streams) returns the
    same bits as one group, for uneven members, mixed tile edges and two pair
    components; repeated calls agree."""
    _say("batch member groups")
    als = [synth.generate(6 + 2 * i, 70 + 37 * i, 80 + i) for i in range(7)]
    dss = [synth.dataset_of(a) for a in als]
    dds = [hotpath.device_dataset(d) for d in dss]
    sites, pairs = [], []
    for a, d in zip(als, dss):
        s1, p1 = _stack_params(model.getparams(a.params, d.tree.branchlengths))
        _, p2 = _stack_params(model.getparams(a.params * 1.05, d.tree.branchlengths))
        sites.append(s1)
        pairs.append((np.concatenate([p1[0], p2[0]]), np.concatenate([p1[1], p2[1]]),
                      np.concatenate([p1[2], p2[2]]) + np.log(0.5)))
    b = lib.Batch(dds)
    try:
        for prec in ("fp64", "fp32"):
            _say("groups", prec)
            b.load(sites, pairs, prec)
            b.set_tiles([16, 32, 16, 16, 32, 16, 16])
            b.set_groups(1)
            ref = b.eval_loaded(1.0, prec)
            for g in (2, 3, 4, 0):
                b.set_groups(g)
                for _ in range(2):
                    assert np.array_equal(b.eval_loaded(1.0, prec), ref), g
            ctx.set_profiling(True)
            try:
                assert np.array_equal(b.eval_loaded(1.0, prec), ref)
            finally:
                ctx.set_profiling(False)
            single = lib.Batch([dds[3]])
            try:
                single.load([sites[3]], [pairs[3]], prec)
                single.set_tiles([16])
                assert np.array_equal(single.eval_loaded(1.0, prec), ref[3:4])
            finally:
                single.close()
        with pytest.raises(lib.MessiError, match="groups"):
            b.set_groups(5)
    finally:
        b.close()


def test_posterior_into_pinned_buffer(ctx, c1):
    """messi_posterior writes into a caller's page-locked buffer (messi_host_alloc)
    exactly what it returns in a fresh array; the buffer can be reused."""
    _say("posterior into a pinned buffer")
    dd = hotpath.device_dataset(c1["ds"])
    site, pair = _stack_params(c1["p"])
    L = c1["ds"].numcols
    buf = lib.pinned_empty((L, L), order="F")
    assert buf.shape == (L, L) and buf.flags.f_contiguous and buf.flags.writeable
    for prec in ("fp64", "fp32", "fp64"):
        z, unp, pp = dd.posterior(site, pair, prec)
        z2, unp2, pp2 = dd.posterior(site, pair, prec, out=buf)
        assert pp2 is buf
        assert z == z2 and np.array_equal(unp, unp2) and np.array_equal(pp, buf)
    with pytest.raises(ValueError):
        dd.posterior(site, pair, "fp64", out=np.zeros((L, L)))   # C order
    del buf
