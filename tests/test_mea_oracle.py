"""MEA posterior decoding in the oracle, pinned without its recursion:

* enumeration of every nested structure of every sub-interval, scored in closed
  form (alpha times the pair posteriors of its pairs, plus the single posterior
  of each unpaired position that is not the last one of its loop -- the
  reference leaves E[i,i] at zero, src/RNATools.jl:11-13);
* decoding the posteriors of one structure returns that structure;
* the first maximising split point on ties, and a traceback worked by hand.
"""
import numpy as np
import pytest

import oracle
from bruteforce import structures


def mea_score(partner, single, p, alpha):
    """Score of a structure on positions 0..n-1 of the arrays given."""
    n = len(partner)
    sc = 0.0
    for x, k in enumerate(partner):
        if k > x:
            sc += alpha * p(x, k)
        elif k < 0:
            last = x == n - 1 or (0 <= partner[x + 1] < x)
            if not last:
                sc += single[x]
    return sc


def brute(single, pair, alpha, i, j, upper=True):
    """(max score, argmax structure as 1-based partners) on [i, j]."""
    n = j - i + 1
    p = (lambda a, b: pair[i + a, i + b]) if upper else (lambda a, b: pair[i + b, i + a])
    best, arg = -np.inf, None
    for s in structures(n):
        v = mea_score(s, single[i:j + 1], p, alpha)
        if v > best:
            best, arg = v, s
    return best, arg


def _inputs(L, seed):
    rng = np.random.default_rng(seed)
    single = rng.random(L)
    pair = rng.random((L, L)) * 0.6   # deliberately not symmetric
    return single, pair


@pytest.mark.parametrize("L", [2, 3, 5, 8])
def test_fill_matches_enumeration_on_every_interval(L):
    single, pair = _inputs(L, L)
    alpha = 1.7
    E, S = oracle.mea(single, pair, alpha, "fp64")
    for i in range(L):
        for j in range(i + 1, L):
            want, _ = brute(single, pair, alpha, i, j)
            assert E[i, j] == pytest.approx(want, rel=1e-13, abs=1e-13), (i, j)
    assert np.all(np.tril(E) == 0) and np.all(np.tril(S) == 0)


@pytest.mark.parametrize("L", [4, 7, 9])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_traceback_is_the_enumerated_argmax(L, prec):
    single, pair = _inputs(L, 100 + L)
    E, S = oracle.mea(single, pair, 2.0, prec)
    # FP64 reads pair[i, k] (the Julia loop); FP32 pair[k, i] (the CUDA kernel's i*L + k)
    want_v, want_s = brute(single, pair, 2.0, 0, L - 1, upper=(prec == "fp64"))
    got = oracle.mea_traceback(S)
    assert E[0, L - 1] == pytest.approx(want_v, rel=1e-6 if prec == "fp32" else 1e-13)
    assert list(got) == [k + 1 if k >= 0 else 0 for k in want_s]


def test_transposed_pair_matrix_changes_the_answer():
    """The asymmetric inputs above do tell the two readings apart."""
    single, pair = _inputs(9, 109)
    E64, _ = oracle.mea(single, pair, 2.0, "fp64")
    Et, _ = oracle.mea(single, np.ascontiguousarray(pair.T), 2.0, "fp64")
    assert E64[0, 8] != Et[0, 8]


def random_structure(L, rng, frac=0.6):
    partner = [-1] * L
    stack = []
    for x in range(L):
        r = rng.random()
        if stack and r < 0.3 and x - stack[-1] >= 4:
            y = stack.pop()
            partner[x], partner[y] = y, x
        elif r < 0.3 + frac * 0.5:
            stack.append(x)
    return partner


def pairprobs(structs):
    """getpairprobs (src/RNATools.jl:87-102): posteriors of a list of structures."""
    L = len(structs[0])
    single = np.zeros(L)
    pp = np.zeros((L, L))
    for s in structs:
        for x, k in enumerate(s):
            if k > x:
                pp[x, k] += 1
                pp[k, x] += 1
            elif k < 0:
                single[x] += 1
    return single / len(structs), pp / len(structs)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_decoding_one_structure_returns_it(prec):
    rng = np.random.default_rng(7)
    for L in (30, 151, 400):
        s = random_structure(L, rng)
        single, pp = pairprobs([s])
        _, S = oracle.mea(single, pp, 2.0, prec)
        got = oracle.mea_traceback(S)
        assert list(got) == [k + 1 if k >= 0 else 0 for k in s]


def test_first_split_point_wins_ties():
    # E[0,3]: k = 1 and k = 2 both give alpha * 0.25; the first (1-based 2) is kept
    single = np.zeros(4)
    pair = np.zeros((4, 4))
    pair[0, 1] = pair[0, 2] = 0.25
    E, S = oracle.mea(single, pair, 2.0, "fp64")
    assert S[0, 3] == 2 and E[0, 3] == 0.5
    # e1 is kept over an equal e2 (strict comparison)
    pair = np.zeros((2, 2))
    pair[0, 1] = 0.5
    E, S = oracle.mea(np.array([1.0, 0.0]), pair, 2.0, "fp64")
    assert S[0, 1] == -1 and E[0, 1] == 1.0


def test_traceback_by_hand():
    # (0,7) pairs; inside, position 1 pairs with 3 (split, 1-based 4), then [4,6]
    # has 4 unpaired and (5,6) paired.
    L = 8
    S = np.zeros((L, L), dtype=np.int64)
    S[0, 7] = -2
    S[1, 6] = 4
    S[2, 2] = 0
    S[4, 6] = -1
    S[5, 6] = -2
    got = oracle.mea_traceback(S)
    assert list(got) == [8, 4, 0, 2, 0, 7, 6, 1]
