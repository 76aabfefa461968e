"""Alignment + tree container (host side), following src/Dataset.jl:99-199.

Nucleotides are stored as 4-bit state masks (bit a set <=> state a is allowed),
which is the reference's one-hot ``data[s, c, 1:4]`` 0/1 encoding packed into a
byte (src/Dataset.jl:13-96): A=1, C=2, G=4, T/U=8, IUPAC ambiguity codes set
several bits, and a gap or any other character sets all four (0xF).

Per-node column patterns (the reference's ``subcolumnrefs``, src/Dataset.jl:180-196)
are kept as dense pattern ids per node, numbered in order of first occurrence,
together with the reference-style ``subcolumnrefs`` (first column carrying the
same pattern).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .tree import TreeArrays, gettreefromnewick, roottree, tree_arrays

_IUPAC = {
    "A": 0b0001, "C": 0b0010, "G": 0b0100, "T": 0b1000, "U": 0b1000,
    "B": 0b1110, "D": 0b1101, "H": 0b1011, "K": 0b1100, "M": 0b0011,
    "R": 0b0101, "S": 0b0110, "V": 0b0111, "W": 0b1001, "Y": 0b1010,
}
_NUCMAP = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3}

_LUT = np.full(256, 0xF, dtype=np.uint8)
for _ch, _m in _IUPAC.items():
    _LUT[ord(_ch)] = _m
    _LUT[ord(_ch.lower())] = _m


def encode(sequences: Sequence[str]) -> np.ndarray:
    """(nseq, ncol) uint8 state masks, src/Dataset.jl:13-96."""
    arr = np.array([np.frombuffer(s.encode("ascii"), dtype=np.uint8) for s in sequences])
    return _LUT[arr]


def codes_to_onehot(codes: np.ndarray) -> np.ndarray:
    """(nseq, ncol, 4) float64 0/1 array, the reference's ``data`` layout."""
    bits = np.arange(4, dtype=np.uint8)
    return ((codes[..., None] >> bits) & 1).astype(np.float64)


def onehot_to_codes(data: np.ndarray) -> np.ndarray:
    d = np.asarray(data)
    if not np.all((d == 0.0) | (d == 1.0)):
        raise ValueError("data must hold 0/1 state indicators")
    return (d[..., 0] + 2 * d[..., 1] + 4 * d[..., 2] + 8 * d[..., 3]).astype(np.uint8)


def _dense_ids(keys: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Dense ids in order of first occurrence and the first column of each id."""
    _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return rank[inv].astype(np.int32), first[order].astype(np.int32)


class Dataset:
    """The reference's ``Dataset`` (src/Dataset.jl:99-199), reduced to the fields the
    likelihood path reads."""

    def __init__(self, names: Sequence[str], sequences: Sequence[str], newick: str,
                 doroottree: bool = True):
        if len(set(len(s) for s in sequences)) != 1:
            raise ValueError("aligned sequences must have equal length")
        self.names = list(names)
        self.sequences = list(sequences)
        self.numseqs = len(sequences)
        self.numcols = len(sequences[0])
        self.codes = encode(self.sequences)
        self.seqnametoindex: Dict[str, int] = {n: i for i, n in enumerate(self.names)}
        obs = np.zeros(4)
        gap = np.zeros(self.numcols)
        for s in self.sequences:
            for j, ch in enumerate(s):
                k = _NUCMAP.get(ch, -1)
                if k >= 0:
                    obs[k] += 1
                else:
                    gap[j] += 1
        self.obsfreqs = obs / max(obs.sum(), 1.0)           # src/Dataset.jl:142-149
        self.gapfrequency = gap / self.numseqs              # src/Dataset.jl:150-152
        root = gettreefromnewick(newick)
        if doroottree:
            root = roottree(root, 0)
        self.newick = newick
        self.tree: TreeArrays = tree_arrays(root, self.seqnametoindex)
        self.numnodes = self.tree.nnodes
        self._patterns()

    def _patterns(self) -> None:
        t = self.tree
        n, L = t.nnodes, self.numcols
        pid = np.zeros((n, L), dtype=np.int32)
        first = [None] * n
        for v in reversed(range(n)):          # children before parents (pre-order reversed)
            if t.children[v, 0] < 0:
                pid[v], first[v] = _dense_ids(self.codes[t.seqindex[v]])
            else:
                l, r = t.children[v]
                key = pid[l].astype(np.int64) * (L + 1) + pid[r]
                pid[v], first[v] = _dense_ids(key)
        self.patternid = pid                   # (nnodes, L) dense pattern ids
        self.patternfirst = first              # per node: first column of each pattern id
        self.numpatterns = np.array([len(f) for f in first], dtype=np.int64)
        # subcolumnrefs (0-based): first column with the same subtree pattern.
        self.subcolumnrefs = np.stack([first[v][pid[v]] for v in range(n)])

    @property
    def data(self) -> np.ndarray:
        return codes_to_onehot(self.codes)

    @classmethod
    def from_fasta_newick(cls, fasta_text: str, newick: str) -> "Dataset":
        names, seqs, cur = [], [], None
        for line in fasta_text.splitlines():
            line = line.strip()
            if not line:
                continue
            if line.startswith(">"):
                names.append(line[1:].strip())
                seqs.append("")
            else:
                seqs[-1] += line
        return cls(names, seqs, newick)

    def to_fasta(self) -> str:
        return "".join(f">{n}\n{s}\n" for n, s in zip(self.names, self.sequences))


def pair_keys_root(ds: Dataset, maxbasepairdistance: int,
                   mask: Optional[np.ndarray] = None) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Root-level pair dedup, following ``getindexdict`` at the root (src/FelsensteinCUDA.jl:86-100).

    Returns (xs, ys, index): representative column pairs (x < y) of each distinct
    ordered root-pattern pair, and for every in-range column pair of the upper
    triangle (row-major order over x then y) the index of its representative.
    """
    L = ds.numcols
    pid = ds.patternid[0].astype(np.int64)
    D = min(maxbasepairdistance, L)
    xs_all, ys_all = [], []
    for x in range(L - 1):
        y = np.arange(x + 1, min(x + D, L - 1) + 1)
        if mask is not None:
            y = y[mask[x, y] > 0]
        xs_all.append(np.full(len(y), x, dtype=np.int64))
        ys_all.append(y)
    xs_all = np.concatenate(xs_all) if xs_all else np.zeros(0, np.int64)
    ys_all = np.concatenate(ys_all) if ys_all else np.zeros(0, np.int64)
    key = pid[xs_all] * L + pid[ys_all]
    if len(key) == 0:
        return xs_all, ys_all, np.zeros(0, np.int64)
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return xs_all[first[order]], ys_all[first[order]], rank[inv.reshape(-1)]
