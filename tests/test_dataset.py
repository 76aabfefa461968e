"""Host data model: encoding, Newick/rooting conventions, pattern dedup."""
import numpy as np

from messi_amd import synth
from messi_amd.dataset import Dataset, codes_to_onehot, encode, pair_keys_root
from messi_amd.tree import gettreefromnewick, roottree, tree_arrays


def test_iupac_encoding():
    c = encode(["ACGTUNRY-BDHKMSVW"])[0]
    expect = [1, 2, 4, 8, 8, 15, 5, 10, 15, 14, 13, 11, 12, 3, 6, 7, 9]
    assert list(c) == expect
    oh = codes_to_onehot(c[None, :])
    assert oh.shape == (1, 17, 4) and oh[0, 0].tolist() == [1, 0, 0, 0]


def test_rooting_and_preorder():
    root = roottree(gettreefromnewick("(a:0.2,(b:0.1,c:0.3):0.4,d:0.6);"), 0)
    t = tree_arrays(root, {"a": 0, "b": 1, "c": 2, "d": 3})
    # new root -> [a (0.1), old root (0.1) -> [(b,c) 0.4, d 0.6]]
    assert t.names[0] == "" and t.names[1] == "a"
    assert abs(t.branchlengths[1] - 0.1) < 1e-15 and abs(t.branchlengths[2] - 0.1) < 1e-15
    assert t.branchlengths[0] == 0.0
    assert all(t.children[v, 0] > v for v in t.internal())
    assert t.nnodes == 7


def _subcolumnrefs_naive(ds):
    t = ds.tree
    out = np.zeros((t.nnodes, ds.numcols), dtype=np.int64)
    for v in range(t.nnodes):
        # leaves of the subtree in pre-order
        stack, leaves = [v], []
        while stack:
            w = stack.pop()
            if t.children[w, 0] < 0:
                leaves.append(t.seqindex[w])
            else:
                stack.extend([t.children[w, 1], t.children[w, 0]])
        seen = {}
        for c in range(ds.numcols):
            key = tuple(ds.codes[leaves, c])
            out[v, c] = seen.setdefault(key, c)
    return out


def test_pattern_dedup_matches_naive():
    a = synth.generate(12, 80, 7)
    ds = synth.dataset_of(a)
    assert np.array_equal(ds.subcolumnrefs, _subcolumnrefs_naive(ds))


def test_root_pair_dedup():
    ds = Dataset(["a", "b", "c"], ["AAGA", "CCUC", "GGAG"], "(a:0.1,b:0.2,c:0.3);")
    xs, ys, idx = pair_keys_root(ds, 10)
    # columns 0,1,3 share a pattern: (0,1),(0,3),(1,3) -> one key; (0,2),(1,2) -> one; (2,3) -> one
    assert len(xs) == 3
    assert len(idx) == 6


def test_generator_shapes():
    a = synth.config1()
    assert len(a.sequences) == 20 and len(a.sequences[0]) == 120
    frac = np.mean(a.pairedsites >= 0)
    assert 0.2 < frac <= 0.36
    p = a.pairedsites
    for i in range(120):
        if p[i] >= 0:
            assert p[p[i]] == i and abs(p[i] - i) >= 4
