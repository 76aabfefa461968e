"""Write the committed golden vectors for the 20 x 120 nt workload (seed 1).

Inputs (alignment, Newick, parameter vector) come from the seeded generator;
expected outputs come from the FP64 oracle, which is itself pinned by the
brute-force and closed-form tests.  The fixture freezes both so that a change
to the generator, the host model or the oracle shows up as a diff.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from messi_amd import model, synth  # noqa: E402


def main():
    a = synth.config1()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    u = oracle.site_loglik(ds, model.site_components(p))
    M = oracle.pair_loglik(ds, model.pair_components(p))
    ins = oracle.inside(u, M)
    out = oracle.outside(ins, u, M)
    res, pair = oracle.bpp(ins, out, u, M)
    iu = np.triu_indices(ds.numcols, 1)
    np.savez_compressed(
        os.path.join(HERE, "config1.npz"),
        names=np.array(a.names), sequences=np.array(a.sequences), newick=np.array(a.newick),
        params=a.params, site_loglik=u, pair_loglik_upper=M[iu], logZ=np.array(ins[0, 0, -1]),
        unpaired=res, pairprob=pair.astype(np.float64),
        outside_S_diag=np.array([out[0, i, i] for i in range(ds.numcols)]))
    print("wrote", os.path.join(HERE, "config1.npz"))


if __name__ == "__main__":
    main()
