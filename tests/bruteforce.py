"""Brute-force references that pin the oracle independently of its recursions.

* ``prune_enumerate``: the likelihood of one column (pair) by summing over every
  assignment of states to the internal nodes (|S|^k terms) -- no pruning.
* ``kh99_enumerate``: Z and posteriors by enumerating every nested secondary
  structure and scoring it with the unique KH99 derivation of that structure
  (grammar: S -> LS | s | dFd, F -> dFd | LS, L -> s | dFd; rule probabilities of
  src/InsideOutsideAlgorithm.jl:40-46), times the emissions exp(u_i) of unpaired
  and exp(paired[i, j]) of paired positions.
"""
from __future__ import annotations

import itertools
import math
from functools import lru_cache

import numpy as np

P_S_LS, P_S_s, P_S_dFd = 0.868534, 0.117609877998, 0.013856122002
P_F_dFd, P_F_LS = 0.787640, 0.21236
P_L_s, P_L_dFd = 0.894603, 0.105397


def prune_enumerate(children, leafvec, P, pi):
    """children: (n, 2) with -1 for leaves (root 0); leafvec[v]: state vector of leaf v;
    P[v][a, d]: transition over the branch above v; pi: root distribution."""
    n = len(children)
    dim = len(pi)
    internal = [v for v in range(n) if children[v][0] >= 0]
    total = 0.0
    for states in itertools.product(range(dim), repeat=len(internal)):
        st = dict(zip(internal, states))
        w = pi[st[0]]
        for v in internal:
            a = st[v]
            for c in children[v]:
                if children[c][0] >= 0:
                    w *= P[c][a, st[c]]
                else:
                    w *= float(np.dot(P[c][a], leafvec[c]))
            if w == 0.0:
                break
        total += w
    return math.log(total)


def structures(L):
    """All nested secondary structures on L positions as partner lists (-1 unpaired)."""
    @lru_cache(maxsize=None)
    def rec(i, j):
        if i > j:
            return [()]
        out = []
        for rest in rec(i + 1, j):
            out.append(((i, -1),) + rest)
        for k in range(i + 1, j + 1):
            for inner in rec(i + 1, k - 1):
                for rest in rec(k + 1, j):
                    out.append(((i, k), (k, i)) + inner + rest)
        return out

    res = []
    for s in rec(0, L - 1):
        p = [-1] * L
        for a, b in s:
            p[a] = b
        res.append(tuple(p))
    return res


def kh99_derivation(partner):
    """Probability of the unique KH99 derivation of a structure (0 if none)."""
    L = len(partner)

    def S(i, j):
        if i > j:
            return 0.0
        k = partner[i]
        if i == j:
            return P_S_s if k < 0 else 0.0
        if k < 0:
            return P_S_LS * P_L_s * S(i + 1, j)
        if k == j:
            return P_S_dFd * F(i + 1, j - 1)
        if i < k < j:
            return P_S_LS * P_L_dFd * F(i + 1, k - 1) * S(k + 1, j)
        return 0.0

    def F(i, j):
        if i >= j:
            return 0.0
        k = partner[i]
        if k == j:
            return P_F_dFd * F(i + 1, j - 1)
        if k < 0:
            return P_F_LS * P_L_s * S(i + 1, j)
        if i < k < j:
            return P_F_LS * P_L_dFd * F(i + 1, k - 1) * S(k + 1, j)
        return 0.0

    return S(0, L - 1)


def kh99_enumerate(u, paired):
    """(log Z, unpaired posteriors (L), pair posteriors (L x L)) by enumeration."""
    L = len(u)
    Z = 0.0
    unp = np.zeros(L)
    pp = np.zeros((L, L))
    for s in structures(L):
        g = kh99_derivation(s)
        if g == 0.0:
            continue
        le = 0.0
        for i, k in enumerate(s):
            if k < 0:
                le += u[i]
            elif k > i:
                le += paired[i, k]
        if le == -np.inf:
            continue
        w = g * math.exp(le)
        Z += w
        for i, k in enumerate(s):
            if k < 0:
                unp[i] += w
            else:
                pp[i, k] += w
    return math.log(Z), unp / Z, pp / Z
