"""Structure sampling on the device versus the oracle, through the C ABI.

With the same inside table and the same keyed uniforms, the device's draws are
the oracle's draws: identical structures (FP64 table) and priors equal to
rounding.  Drawn from the device's own evaluation, the pair frequencies of many
samples match the base-pair posteriors.
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
    return dict(ds=ds, p=p, u=u, M=M, ins=ins)


def _random_problem(L, seed):
    rng = np.random.default_rng(seed)
    u = rng.normal(-1.0, 0.5, L)
    M = np.full((L, L), -np.inf)
    for i in range(L):
        for j in range(i + 1, L):
            M[i, j] = M[j, i] = rng.normal(-1.5, 1.5)
    return u, M, oracle.inside(u, M)


@pytest.mark.parametrize("L,n", [(1, 4), (2, 8), (7, 2000), (33, 300)])
def test_draws_equal_oracle_small(ctx, L, n):
    _say("sample small", L, n)
    u, M, ins = _random_problem(L, L)
    S, pri = lib.sample_structures(ins, M, u, n, seed=123)
    So, prio = oracle.sample_structures(ins, M, n, seed=123)
    assert np.array_equal(S, So)
    assert np.allclose(pri, prio, rtol=0, atol=1e-10)


def test_draws_equal_oracle_config1(ctx, c1):
    _say("sample config1 fp64")
    S, pri = lib.sample_structures(c1["ins"], c1["M"], c1["u"], 256, seed=2024)
    So, prio = oracle.sample_structures(c1["ins"], c1["M"], 256, seed=2024)
    assert np.array_equal(S, So)
    assert np.allclose(pri, prio, rtol=0, atol=1e-10)
    assert (S > 0).sum() > 0


def test_fp32_table_draws(ctx, c1):
    """The scaled FP32 table differs from the FP64 one by rounding only: nearly
    every draw is the same, and each prior is its structure's."""
    _say("sample config1 fp32")
    S, pri = lib.sample_structures(c1["ins"], c1["M"], c1["u"], 256, seed=7, prec="fp32")
    So, _ = oracle.sample_structures(c1["ins"], c1["M"], 256, seed=7)
    same = np.mean(np.all(S == So, axis=1))
    assert same >= 0.97, same
    for row, pr in zip(S[:32], pri[:32]):
        assert abs(pr - oracle.kh99_prior(row)) < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_posterior_sampling_matches_posteriors(ctx, c1, prec):
    """Draws from the device's own evaluation (no inside table on the host)."""
    _say("sample posterior", prec)
    dd = hotpath.device_dataset(c1["ds"])
    site, pair = hotpath._stack(model.site_components(c1["p"])), hotpath._stack(model.pair_components(c1["p"]))
    z, unp, pp = dd.posterior(site, pair, prec)
    n = 4000
    z2, S, pri = dd.sample_posterior(site, pair, n, seed=99, prec=prec)
    assert abs(z2 - z) <= 1e-12 * abs(z)
    L = S.shape[1]
    freq = np.zeros((L, L))
    rows = np.repeat(np.arange(n), L)
    cols = np.tile(np.arange(L), n)
    part = S.reshape(-1)
    m = part > 0
    np.add.at(freq, (cols[m], part[m] - 1), 1.0)
    freq /= n
    se = np.sqrt(np.maximum(pp * (1 - pp), 1e-4) / n)
    assert np.max(np.abs(freq - pp) / se) < 5.5
    unf = np.mean(S == 0, axis=0)
    assert np.max(np.abs(unf - unp) / np.sqrt(np.maximum(unp * (1 - unp), 1e-4) / n)) < 5.5
    # structures are nested and symmetric
    for row in S[:64]:
        for i, k in enumerate(row):
            if k:
                assert row[k - 1] == i + 1
    for row, p in zip(S[:16], pri[:16]):
        assert abs(p - hotpath.calculateKH99prior(row)) < 1e-9


def test_rejects_bad_arguments(ctx):
    with pytest.raises(ValueError):
        lib.sample_structures(np.zeros((3, 4, 4)), np.zeros((3, 3)), np.zeros(4), 1)
    small = np.zeros(1)
    out = np.zeros(1, dtype=np.intc)
    rc = ctx.lib.messi_sample_structures(ctx.handle, 16384, lib.ptr(small), lib.ptr(small), lib.ptr(small),
                                         lib.FP64, 1, 0, 1.0, out.ctypes.data_as(lib._ip), None)
    assert rc != 0 and b"16383" in ctx.lib.messi_last_error(ctx.handle)
