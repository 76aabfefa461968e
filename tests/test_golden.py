"""Golden vectors of the 20 x 120 nt workload: generator, host model and oracle
reproduce the committed fixture."""
import os

import numpy as np

import oracle
from messi_amd import model, synth
from messi_amd.dataset import Dataset

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load():
    return np.load(os.path.join(HERE, "config1.npz"), allow_pickle=False)


def test_generator_reproduces_fixture_inputs():
    g = _load()
    a = synth.config1()
    assert list(g["sequences"]) == a.sequences
    assert str(g["newick"]) == a.newick
    assert np.array_equal(g["params"], a.params)


def test_oracle_reproduces_fixture_outputs():
    g = _load()
    ds = Dataset(list(g["names"]), list(g["sequences"]), str(g["newick"]))
    p = model.getparams(g["params"], ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    assert np.allclose(u, g["site_loglik"], rtol=1e-13, atol=0)
    M = oracle.pair_loglik(ds, model.pair_components(p))
    iu = np.triu_indices(ds.numcols, 1)
    assert np.allclose(M[iu], g["pair_loglik_upper"], rtol=1e-12, atol=0)
    ins = oracle.inside(u, M)
    assert abs(ins[0, 0, -1] - float(g["logZ"])) < 1e-11 * abs(float(g["logZ"]))
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    assert np.max(np.abs(res - g["unpaired"])) < 1e-12
    assert np.max(np.abs(pair - g["pairprob"])) < 1e-12
