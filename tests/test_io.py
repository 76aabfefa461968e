"""Input formats parsed by the library (host code, no device): FASTA + Newick
into an alignment handle, .ct files and dot-bracket strings.  Expected values
are worked out by hand from the reference's rules, and the parsed alignments
agree with the Python Dataset on the synthetic workloads."""
import numpy as np
import pytest

from messi_amd import lib, synth
from messi_amd.dataset import Dataset


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_hand_made_alignment(tmp_path):
    fasta = _write(tmp_path, "a.fa", ">s1 first\r\nACGU\nRYN-\n>s2\nacgt\n  ..KM\n>s3\nA-G-TTac\n")
    newick = _write(tmp_path, "t.nwk", "( s1 first:0.4, (s2:0.1,s3:0.2)in:0.3, x:0.5);\nignored line\n")
    # the tree has a leaf 'x' without a sequence
    with pytest.raises(lib.MessiError, match="no sequence"):
        lib.Alignment(fasta, newick)
    newick = _write(tmp_path, "t.nwk", "( s1 first:0.4, (s2:0.1,s3:0.2)in:0.3);\n")
    # rooted at its first child the root would have one child left: not binary
    with pytest.raises(lib.MessiError, match="binary"):
        lib.Alignment(fasta, newick)
    newick = _write(tmp_path, "t.nwk", "((s1 first:0.4,s2:0.6)ab:0.3, s3:0.2, (s2:0.1,s3:0.2):0.7);")
    a = lib.Alignment(fasta, newick)
    assert (a.numseqs, a.numcols) == (3, 8)
    assert a.names == ["s1 first", "s2", "s3"]
    want_codes = np.array([[1, 2, 4, 8, 5, 10, 15, 15],
                           [1, 2, 4, 8, 15, 15, 12, 3],
                           [1, 15, 4, 15, 8, 8, 1, 2]], dtype=np.uint8)
    assert np.array_equal(a.codes, want_codes)
    # gap frequency: anything but upper-case A C G T U (lower case included)
    gap = np.array([1, 2, 1, 2, 2, 2, 3, 3]) / 3
    assert np.allclose(a.gapfrequency, gap)
    obs = np.array([2, 1, 2, 3], dtype=float)   # A: s1,s3; C: s1; G: s1,s3; T/U: s1,s3 T,T
    assert np.allclose(a.obsfreqs, obs / obs.sum())
    # roottree: the first child (ab, 0.3) goes under a new root with 0.15; the old
    # root (children s3, (s2,s3)) gets 0.15; pre-order numbering
    assert a.numnodes == 9
    ch = [[1, 4], [2, 3], [-1, -1], [-1, -1], [5, 6], [-1, -1], [7, 8], [-1, -1], [-1, -1]]
    assert np.array_equal(a.children, ch)
    assert np.array_equal(a.seqindex, [-1, -1, 0, 1, -1, 2, -1, 1, 2])
    assert np.allclose(a.branchlengths, [0.0, 0.15, 0.4, 0.6, 0.15, 0.2, 0.7, 0.1, 0.2])
    d = a.data()
    assert d.shape == (3, 8, 4)
    assert np.array_equal(d[..., 0] + 2 * d[..., 1] + 4 * d[..., 2] + 8 * d[..., 3], want_codes)
    a.close()


def test_errors(tmp_path):
    nwk = _write(tmp_path, "t.nwk", "(a:1,b:1,c:1);")
    with pytest.raises(lib.MessiError, match="cannot open"):
        lib.Alignment(str(tmp_path / "missing.fa"), nwk)
    fa = _write(tmp_path, "a.fa", ">a\nACG\n>b\nAC\n>c\nACG\n")
    with pytest.raises(lib.MessiError, match="equal length"):
        lib.Alignment(fa, nwk)
    fa = _write(tmp_path, "b.fa", ">a\nACG\n>b\nACC\n>c\nACG\n")
    bad = _write(tmp_path, "b.nwk", "(a:x,b:1,c:1);")
    with pytest.raises(lib.MessiError, match="branch length"):
        lib.Alignment(fa, bad)


@pytest.mark.parametrize("cfg", ["config1", "config4"])
def test_agrees_with_python_dataset(tmp_path, cfg):
    als = synth.config1() if cfg == "config1" else synth.config4()[:3]
    for k, al in enumerate(als if isinstance(als, list) else [als]):
        ds = synth.dataset_of(al)
        fa = _write(tmp_path, f"{k}.fa", ds.to_fasta())
        nw = _write(tmp_path, f"{k}.nwk", ds.newick + "\n")
        a = lib.Alignment(fa, nw)
        assert a.names == ds.names
        assert np.array_equal(a.codes, ds.codes)
        assert np.allclose(a.gapfrequency, ds.gapfrequency, rtol=0, atol=0)
        assert np.allclose(a.obsfreqs, ds.obsfreqs, rtol=1e-15, atol=0)
        t = ds.tree
        assert np.array_equal(a.children, t.children)
        assert np.array_equal(a.seqindex, t.seqindex)
        assert np.array_equal(a.branchlengths, t.branchlengths)
        # lower case counts as a gap, as in the reference
        low = Dataset(ds.names, [s.lower() for s in ds.sequences], ds.newick)
        assert np.all(low.gapfrequency == 1.0) and np.array_equal(low.codes, ds.codes)


def _write_ct(path, seq, paired):
    # writectfile layout (src/RNATools.jl:395-402)
    with open(path, "w") as f:
        f.write(f"{len(paired)}\n")
        for i, p in enumerate(paired, 1):
            f.write(f"{i}\t{seq[i - 1]}\t{i - 1}\t{i + 1}\t{p}\t{i}\n")


def test_ct_round_trip(tmp_path):
    seq = "GGGAAACCCUU"
    paired = [9, 8, 7, 0, 0, 0, 3, 2, 1, 0, 0]
    p = str(tmp_path / "s.ct")
    _write_ct(p, seq, paired)
    s2, p2 = lib.read_ct(p)
    assert s2 == seq and list(p2) == paired
    # a position already paired keeps its first partner
    lines = open(p).read().splitlines()
    lines[4] = "4\tA\t3\t5\t11\t4"       # 4 -> 11 (both free): pairs
    lines.append("12\tX\t11\t13\t1\t12")  # 12 -> 1: y < x, ignored
    lines[0] = "11 title words"
    open(p, "w").write("\n".join(lines[:12]) + "\n")
    s3, p3 = lib.read_ct(p)
    assert p3[3] == 11 and p3[10] == 4 and p3[0] == 9


def test_dbn():
    assert list(lib.parse_dbn("((..))")) == [6, 5, 0, 0, 2, 1]
    # ')' does not close past the open '<' (ignored); '>' closes '<'
    assert list(lib.parse_dbn("(<)>")) == [0, 4, 0, 2]
    assert list(lib.parse_dbn("(<>)")) == [4, 3, 2, 1]
    # a closing bracket of another kind than the innermost open one is ignored
    assert list(lib.parse_dbn("(<)")) == [0, 0, 0]
    assert list(lib.parse_dbn("AaBb..")) == [2, 1, 4, 3, 0, 0]
    assert list(lib.parse_dbn("  ..((..))\n")) == [0, 0, 8, 7, 0, 0, 4, 3]
    assert list(lib.parse_dbn("")) == []
