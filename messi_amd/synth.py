"""Seeded synthetic alignments for the benchmark configurations.

The generator follows the idea of the reference's simulator
(src/Simulations.jl:72-183): a nested secondary structure is drawn, paired
columns evolve jointly under the 16-state MuSE model and unpaired columns under
4-state GTR, from a stationary root state down a tree.  Sizes:

* ``config1``: 20 sequences x 120 nt, ~35 % paired, helices 4-8 bp.
* ``config2``: 50 sequences x 1,500 nt, ~40 % paired.
* ``config3``: 30 sequences x 9,000 nt, ~30 % paired, helices up to 12 bp.
* ``config4``: 64 instances, L ~ U[80, 600], N ~ U[10, 60].

Tree: Yule topology, branch lengths ~ Exp(mean 0.1), written as Newick with a
trifurcating root (the reference's rooting rule needs one).  Rates:
GTR q ~ Dirichlet(1^6) (unnormalised, as the reference uses them), base
frequencies ~ Dirichlet(5^4), lambdaGC = 2, lambdaAU = 1.5, lambdaGU = 1.
A fraction ``gapfrac`` of cells is replaced by gaps.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .model import calculatemusefreqs, gettransitionmatriceslist, gtr, muse

NUC = np.array(list("ACGU"))


@dataclass
class SynthAlignment:
    names: List[str]
    sequences: List[str]
    newick: str
    pairedsites: np.ndarray         # 0-based partner or -1
    params: np.ndarray              # 15-vector of the generating parameters
    seed: int


def _yule(n: int, rng) -> Tuple[List[List[int]], np.ndarray]:
    """Rooted binary Yule topology: children lists (node 0 = root), iid Exp(0.1) branches."""
    children: List[List[int]] = [[]]
    tips = [0]
    while len(tips) < n:
        k = int(rng.integers(len(tips)))
        v = tips.pop(k)
        a, b = len(children), len(children) + 1
        children.append([])
        children.append([])
        children[v] = [a, b]
        tips += [a, b]
    bl = rng.exponential(0.1, size=len(children))
    bl[0] = 0.0
    return children, bl


def _structure(L: int, frac: float, hmin: int, hmax: int, rng, minloop: int = 3) -> np.ndarray:
    """Random nested structure with ~frac of sites paired, helices of hmin..hmax bp."""
    partner = -np.ones(L, dtype=np.int64)
    target = int(frac * L)
    paired = 0
    attempts = 0
    while paired + 2 * hmin <= target and attempts < 20000:
        attempts += 1
        h = int(rng.integers(hmin, hmax + 1))
        h = min(h, (target - paired) // 2)
        if h < hmin:
            break
        span = 2 * h + minloop + int(rng.integers(0, max(1, L // 8)))
        if span > L:
            continue
        i = int(rng.integers(0, L - span + 1))
        j = i + span - 1
        ok = True
        for k in range(h):
            a, b = i + k, j - k
            if partner[a] >= 0 or partner[b] >= 0:
                ok = False
                break
        if not ok:
            continue
        # Nestedness: no existing pair (c, d) may cross (i, j) or sit between the helix arms.
        ps = np.nonzero(partner >= 0)[0]
        for c in ps:
            d = partner[c]
            if c > d:
                continue
            if (c < i < d < j) or (i < c < j < d):
                ok = False
                break
            if (i < c < i + h) or (j - h < c < j) or (i < d < i + h) or (j - h < d < j):
                ok = False
                break
        if not ok:
            continue
        for k in range(h):
            partner[i + k] = j - k
            partner[j - k] = i + k
        paired += 2 * h
    return partner


def _evolve(children, bl, root_states: np.ndarray, P: np.ndarray, rng) -> List[np.ndarray]:
    """States at every node; ``P[node]`` is the matrix over the branch above ``node``."""
    st = [None] * len(children)
    st[0] = root_states
    stack = [0]
    while stack:
        v = stack.pop()
        for c in children[v]:
            cum = np.cumsum(P[c][st[v]], axis=1)
            u = rng.random(len(st[v]))[:, None] * cum[:, -1:]
            st[c] = np.minimum((cum < u).sum(axis=1), P[c].shape[0] - 1)
            stack.append(c)
    return st


def _newick_trifurcating(children, bl, names) -> str:
    def h(v: int) -> str:
        if not children[v]:
            return f"{names[v]}:{bl[v]:.10g}"
        return "(" + ",".join(h(c) for c in children[v]) + f"):{bl[v]:.10g}"
    a, b = children[0]
    if children[a]:
        a1, a2 = children[a]
        parts = [h(a1), h(a2), f"{h(b).rsplit(':', 1)[0]}:{bl[a] + bl[b]:.10g}"]
    else:
        b1, b2 = children[b]
        parts = [h(b1), h(b2), f"{h(a).rsplit(':', 1)[0]}:{bl[a] + bl[b]:.10g}"]
    return "(" + ",".join(parts) + ");"


def generate(nseq: int, L: int, seed: int, pairfrac: float = 0.35, hmin: int = 4, hmax: int = 8,
             gapfrac: float = 0.05, lambdas=(2.0, 1.5, 1.0)) -> SynthAlignment:
    rng = np.random.default_rng(seed)
    children, bl = _yule(nseq, rng)
    q = rng.dirichlet(np.ones(6))
    pi = rng.dirichlet(np.full(4, 5.0))
    lgc, lau, lgu = lambdas
    partner = _structure(L, pairfrac, hmin, hmax, rng)
    leaves = [v for v in range(len(children)) if not children[v]]
    names = [""] * len(children)
    for k, v in enumerate(leaves):
        names[v] = f"seq{k}"
    cols = np.zeros((nseq, L), dtype=np.int64)
    unp = np.nonzero(partner < 0)[0]
    if len(unp):
        P4 = gettransitionmatriceslist(bl, gtr(*q, pi))
        st = _evolve(children, bl, rng.choice(4, size=len(unp), p=pi), P4, rng)
        for k, v in enumerate(leaves):
            cols[k, unp] = st[v]
    left = np.nonzero(partner > np.arange(L))[0]
    if len(left):
        pi16 = calculatemusefreqs(pi, lgc, lau, lgu)
        P16 = gettransitionmatriceslist(bl, muse(lgc, lau, lgu, *q, pi))
        st = _evolve(children, bl, rng.choice(16, size=len(left), p=pi16 / pi16.sum()), P16, rng)
        for k, v in enumerate(leaves):
            cols[k, left] = st[v] // 4
            cols[k, partner[left]] = st[v] % 4
    chars = NUC[cols]
    gaps = rng.random((nseq, L)) < gapfrac
    chars[gaps] = "-"
    seqs = ["".join(r) for r in chars]
    params = np.array([lgc, lau, lgu, *q, 1.0, 1.0, pi[0], pi[1], pi[2], 1e-4])
    return SynthAlignment([f"seq{k}" for k in range(nseq)], seqs,
                          _newick_trifurcating(children, bl, names), partner, params, seed)


def config1(seed: int = 1) -> SynthAlignment:
    return generate(20, 120, seed, pairfrac=0.35)


def config2(seed: int = 2) -> SynthAlignment:
    return generate(50, 1500, seed, pairfrac=0.40)


def config3(seed: int = 3) -> SynthAlignment:
    return generate(30, 9000, seed, pairfrac=0.30, hmax=12)


def config4(seed: int = 4, count: int = 64) -> List[SynthAlignment]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        L = int(rng.integers(80, 601))
        N = int(rng.integers(10, 61))
        out.append(generate(N, L, seed * 1000 + i, pairfrac=0.35))
    return out


def dataset_of(a: SynthAlignment):
    from .dataset import Dataset
    return Dataset(a.names, a.sequences, a.newick)
