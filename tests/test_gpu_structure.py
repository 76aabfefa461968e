"""Fixed-structure log-likelihood on the device versus the oracle.

The oracle's value is the sum of its site log-likelihoods over the unpaired
columns and of its pair log-likelihoods over the structure's pairs (both pinned
by enumeration elsewhere); FP64 within 1e-9 relative, FP32 within 1e-5.
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
def c1():
    _say("config1 oracle tables")
    a = synth.config1()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p))
    true = np.asarray(a.pairedsites)
    paired = np.where(true >= 0, true + 1, 0)
    return dict(ds=ds, p=p, u=u, M=M, paired=paired)


def _oracle_ll(u, M, paired):
    un = sum(u[x] for x in range(len(u)) if paired[x] == 0)
    pa = sum(M[x, paired[x] - 1] for x in range(len(u)) if paired[x] - 1 > x)
    return un + pa, un, pa


def _comps(p):
    return hotpath._stack(model.site_components(p)), hotpath._stack(model.pair_components(p))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
@pytest.mark.parametrize("which", ["true", "sampled", "empty"])
def test_structure_loglik_matches_oracle(c1, prec, tol, which):
    _say("structure loglik", which, prec)
    L = len(c1["u"])
    if which == "true":
        paired = c1["paired"]
    elif which == "empty":
        paired = np.zeros(L, dtype=np.int64)
    else:
        ins = oracle.inside(c1["u"], c1["M"])
        S, _ = oracle.sample_structures(ins, c1["M"], 1, seed=4)
        paired = S[0]
    want, wu, wp = _oracle_ll(c1["u"], c1["M"], paired)
    site, pair = _comps(c1["p"])
    for dd in (hotpath.device_dataset(c1["ds"]), hotpath.structure_dataset(c1["ds"], paired)):
        ll, (un, pa) = dd.structure_loglik(site, pair, paired, prec)
        assert abs(ll - want) <= tol * abs(want)
        assert abs(un - wu) <= tol * abs(wu) + 1e-12
        assert abs(pa - wp) <= tol * max(1.0, abs(wp))


def test_reference_named_entry_points(c1):
    _say("felsensteincudapaired / computemodellikelihood")
    want, wu, wp = _oracle_ll(c1["u"], c1["M"], c1["paired"])
    got = hotpath.felsensteincudapaired(c1["ds"], c1["paired"], c1["p"])
    assert abs(got - wp) <= 1e-9 * abs(wp)
    post = hotpath.computemodellikelihood(c1["ds"], c1["p"], c1["paired"])
    assert abs(post - (model.computeprior(c1["p"]) + want)) <= 1e-9 * abs(want)


def test_missing_and_malformed_pairs_are_rejected(c1):
    L = len(c1["u"])
    site, pair = _comps(c1["p"])
    near = lib.DeviceDataset(c1["ds"], 5)
    paired = np.zeros(L, dtype=np.int64)
    paired[0], paired[40] = 41, 1   # distance 40 > 5
    with pytest.raises(lib.MessiError, match="pair list"):
        near.structure_loglik(site, pair, paired)
    bad = np.zeros(L, dtype=np.int64)
    bad[0] = 5                       # partner 5 does not point back
    with pytest.raises(lib.MessiError, match="symmetric"):
        hotpath.device_dataset(c1["ds"]).structure_loglik(site, pair, bad)
    near.close()


def test_dataset_from_files_evaluates_like_the_python_dataset(c1, tmp_path):
    """FASTA + Newick parsed by the library give the same device dataset."""
    _say("dataset from files")
    ds = c1["ds"]
    fa = tmp_path / "c1.fa"
    fa.write_text(ds.to_fasta())
    nw = tmp_path / "c1.nwk"
    nw.write_text(ds.newick + "\n")
    al = lib.Alignment(str(fa), str(nw))
    p = model.getparams(synth.config1().params, al.branchlengths)
    site, pair = _comps(p)
    d1 = al.device_dataset()
    d2 = hotpath.device_dataset(ds)
    assert d1.stats() == d2.stats()
    for prec in ("fp64", "fp32"):
        assert d1.eval(site, pair, 1.0, prec) == d2.eval(site, pair, 1.0, prec)
    d1.close()
    al.close()
