"""FP64 CPU oracle (test infrastructure only).

Python bindings for ``messi_oracle.c``, the serial CPU restatement of the
reference's likelihood algorithms (see the header of that file for the
reference file:line of each function).  Only the tests, the smoke check and
the CPU-baseline leg of the benchmark may import this package; the product
library never does.

``build()`` compiles the C file with gcc (OpenMP) into ``oracle/_build/``;
``set_threads`` sets the number of threads of its loops, which changes no
result (each cell is computed by one thread with the serial arithmetic).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "messi_oracle.c")
_LIB = os.path.join(_HERE, "_build", "libmessi_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8 = ctypes.POINTER(ctypes.c_uint8)


def build() -> str:
    os.makedirs(os.path.dirname(_LIB), exist_ok=True)
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-fno-fast-math", "-ffp-contract=off", "-o", _LIB, _SRC,
           "-lm"]
    subprocess.run(cmd, check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_get_threads.argtypes = []
        L.oracle_get_threads.restype = ctypes.c_int
        L.oracle_site_loglik.argtypes = [ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int, _u8, _dp, _dp, _dp]
        L.oracle_pair_loglik.argtypes = [ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int, _u8, _dp, _dp,
                                         ctypes.c_int, _u8, _dp]
        L.oracle_mix_accumulate.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_double, ctypes.c_int]
        L.oracle_inside.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, _dp]
        L.oracle_outside.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_bpp.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_mea.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _ip]
        L.oracle_mea_traceback.argtypes = [ctypes.c_int, _ip, _ip]
        L.oracle_sample_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_sample_uniform.restype = ctypes.c_double
        L.oracle_sample_structure.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_uint64,
                                              ctypes.c_int, _ip]
        L.oracle_sample_structure.restype = ctypes.c_double
        L.oracle_kh99_prior.argtypes = [ctypes.c_int, _ip]
        L.oracle_kh99_prior.restype = ctypes.c_double
        L.oracle_sample_states.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip,
                                           ctypes.c_double, ctypes.c_uint64, _ip, _dp, _dp]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops (results do not depend on it)."""
    lib().oracle_set_threads(int(n))


def get_threads() -> int:
    return int(lib().oracle_get_threads())


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _colmajor_P(P: np.ndarray) -> np.ndarray:
    """(nnodes, dim, dim) row-indexed [a, d] -> column-major per matrix."""
    return np.ascontiguousarray(np.transpose(P, (0, 2, 1)))


def _tree(ds):
    ch = np.ascontiguousarray(ds.tree.children.reshape(-1), dtype=np.int32)
    sq = np.ascontiguousarray(ds.tree.seqindex, dtype=np.int32)
    return ch, sq


def site_loglik(ds, components) -> np.ndarray:
    """Site log-likelihoods with the site-rate mixture, src/FARCE.jl:1198-1220.
    ``components``: list of (P4 (nnodes,4,4), pi4, logw)."""
    L = ds.numcols
    ch, sq = _tree(ds)
    codes = np.ascontiguousarray(ds.codes)
    acc = np.zeros(L)
    tmp = np.zeros(L)
    for i, (P, pi, logw) in enumerate(components):
        Pc = _colmajor_P(np.asarray(P, dtype=np.float64))
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        lib().oracle_site_loglik(ds.numnodes, _p(ch, _ip), _p(sq, _ip), ds.numseqs, L, _p(codes, _u8),
                                 _p(Pc), _p(pi), _p(tmp))
        lib().oracle_mix_accumulate(L, _p(acc), _p(tmp), float(logw), int(i == 0))
    return acc


def pair_loglik(ds, components, maxbp: int = 1000000, mask: Optional[np.ndarray] = None) -> np.ndarray:
    """Pair log-likelihood matrix (L x L, -inf outside the computed band), mixed
    over components with log weights, src/FARCE.jl:1736-1792 semantics."""
    L = ds.numcols
    ch, sq = _tree(ds)
    codes = np.ascontiguousarray(ds.codes)
    acc = np.full(L * L, -np.inf)
    tmp = np.full(L * L, -np.inf)
    m = None if mask is None else np.asfortranarray(mask, dtype=np.uint8).reshape(-1, order="F")
    for i, (P, pi, logw) in enumerate(components):
        Pc = _colmajor_P(np.asarray(P, dtype=np.float64))
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        tmp[:] = -np.inf
        lib().oracle_pair_loglik(ds.numnodes, _p(ch, _ip), _p(sq, _ip), ds.numseqs, L, _p(codes, _u8),
                                 _p(Pc), _p(pi), int(min(maxbp, 2**31 - 1)),
                                 None if m is None else _p(m, _u8), _p(tmp))
        lib().oracle_mix_accumulate(L * L, _p(acc), _p(tmp), float(logw), int(i == 0))
    out = acc.reshape(L, L, order="F")
    return out


def inside(u: np.ndarray, paired: np.ndarray, B: float = 1.0) -> np.ndarray:
    """(3, L, L) inside table, index [s, i, j] with S=0, F=1, L=2."""
    L = len(u)
    u = np.ascontiguousarray(u, dtype=np.float64)
    pm = np.asfortranarray(paired, dtype=np.float64).reshape(-1, order="F")
    out = np.empty(3 * L * L)
    lib().oracle_inside(L, _p(u), _p(pm), float(B), _p(out))
    return out.reshape((3, L, L), order="F")


def outside(inside_t: np.ndarray, u: np.ndarray, paired: np.ndarray) -> np.ndarray:
    L = len(u)
    it = np.asfortranarray(inside_t).reshape(-1, order="F")
    u = np.ascontiguousarray(u, dtype=np.float64)
    pm = np.asfortranarray(paired, dtype=np.float64).reshape(-1, order="F")
    out = np.empty(3 * L * L)
    lib().oracle_outside(L, _p(it), _p(u), _p(pm), _p(out))
    return out.reshape((3, L, L), order="F")


def bpp(inside_t, outside_t, u, paired):
    L = len(u)
    it = np.asfortranarray(inside_t).reshape(-1, order="F")
    ot = np.asfortranarray(outside_t).reshape(-1, order="F")
    u = np.ascontiguousarray(u, dtype=np.float64)
    pm = np.asfortranarray(paired, dtype=np.float64).reshape(-1, order="F")
    res = np.empty(L)
    pr = np.empty(L * L)
    lib().oracle_bpp(L, _p(it), _p(ot), _p(u), _p(pm), _p(res), _p(pr))
    return res, pr.reshape(L, L, order="F")


def mea(single, pair, alpha: float = 2.0, prec: str = "fp64"):
    """MEA fill: (E, S) as L x L arrays (E float64, S int: -1, -2 or the 1-based
    split point; 0 on and below the diagonal).  ``pair`` is the L x L posterior
    matrix as computebasepairprobs returns it (Julia indexing pair[i, k])."""
    single = np.ascontiguousarray(single, dtype=np.float64)
    L = len(single)
    pm = np.asfortranarray(pair, dtype=np.float64).reshape(-1, order="F")
    E = np.empty(L * L)
    S = np.empty(L * L, dtype=np.intc)
    lib().oracle_mea(L, _p(single), _p(pm), float(alpha), 0 if prec == "fp64" else 1, _p(E), _p(S, _ip))
    return E.reshape(L, L, order="F"), S.reshape(L, L, order="F").astype(np.int64)


def mea_traceback(S) -> np.ndarray:
    """Pairing from an S table: 1-based partner per position, 0 when unpaired."""
    L = S.shape[0]
    s = np.asfortranarray(S, dtype=np.intc).reshape(-1, order="F")
    out = np.empty(L, dtype=np.intc)
    lib().oracle_mea_traceback(L, _p(s, _ip), _p(out, _ip))
    return out.astype(np.int64)


def sample_structures(inside_t, paired, nsamples: int, seed: int = 0, B: float = 1.0, first: int = 0):
    """Stochastic tracebacks from (S, 1, L): (partners (nsamples x L, 1-based, 0
    unpaired), KH99 log priors (nsamples)).  Sample s draws its uniforms from
    (seed, first + s)."""
    L = inside_t.shape[1]
    it = np.asfortranarray(inside_t, dtype=np.float64).reshape(-1, order="F")
    pm = np.asfortranarray(paired, dtype=np.float64).reshape(-1, order="F")
    out = np.empty((nsamples, L), dtype=np.intc)
    pri = np.empty(nsamples)
    f = lib().oracle_sample_structure
    for k in range(nsamples):
        row = out[k]
        pri[k] = f(L, _p(it), _p(pm), float(B), int(seed), first + k, row.ctypes.data_as(_ip))
    return out.astype(np.int64), pri


def kh99_prior(partner) -> float:
    """calculateKH99prior of a structure given as 1-based partners (0 unpaired)."""
    p = np.ascontiguousarray(partner, dtype=np.intc)
    return float(lib().oracle_kh99_prior(len(p), _p(p, _ip)))


def sample_uniform(seed: int, sample: int, sym: int, x: int, y: int) -> float:
    return float(lib().oracle_sample_uniform(int(seed), sample, sym, x, y))


def sample_states(sitec, pairc, logw16, partner, cats: int, B: float = 1.0, seed: int = 0):
    """State sampling given a structure (oracle_sample_states): sitec (K4 x L),
    pairc (K16 x L), logw16 (K16), partner 1-based.  Returns (states 1-based,
    unpaired log-likelihoods, paired log-likelihoods at the openers)."""
    sitec = np.ascontiguousarray(sitec, dtype=np.float64)
    pairc = np.ascontiguousarray(pairc, dtype=np.float64)
    lw = np.ascontiguousarray(logw16, dtype=np.float64)
    part = np.ascontiguousarray(partner, dtype=np.int32)
    L = len(part)
    K16 = pairc.shape[0]
    assert sitec.shape == (cats, L) and pairc.shape == (K16, L) and K16 % (cats * cats) == 0
    st = np.zeros(L, dtype=np.int32)
    unp = np.zeros(L)
    pll = np.zeros(L)
    lib().oracle_sample_states(L, cats, K16 // (cats * cats), _p(sitec), _p(pairc), _p(lw), _p(part, _ip),
                               float(B), int(seed), _p(st, _ip), _p(unp), _p(pll))
    return st.astype(np.int64), unp, pll


def site_rate_posteriors(ds, site_comps, pair_comps, cats: int, single, pairprob, maxbp: int = 1000000,
                         mask: Optional[np.ndarray] = None) -> np.ndarray:
    """(cats, L) log posterior weights of the site-rate categories of every column,
    felsensteinposteriorsiterateconditionalscuda (src/FelsensteinCUDA.jl:518-717):
    for category c and column x the log-sum-exp of log w_c + log single[x] +
    log L_c(x) and of log pairprob[x, y] + log w_k + log L_k(x, y) over the pairs
    (x, y), x < y, of the pair list and the pair components k = (c1, c2, lambda
    class), counted for c1 at x and for c2 at y.  Components are ordered c1, c2,
    lambda class (K16 = cats^2 * classes)."""
    L = ds.numcols
    single = np.asarray(single, dtype=np.float64)
    pp = np.asarray(pairprob, dtype=np.float64)
    out = np.full((cats, L), -np.inf)
    with np.errstate(divide="ignore"):
        ls = np.log(single)
        lpp = np.log(pp)
    for c, (P, pi, lw) in enumerate(site_comps):
        u = site_loglik(ds, [(P, pi, 0.0)])
        t = np.where(np.isfinite(u), lw + ls + u, -np.inf)
        out[c] = np.logaddexp(out[c], t)
    M = len(pair_comps) // (cats * cats)
    upper = np.triu(np.ones((L, L), dtype=bool), 1)
    for k, (P, pi, lw) in enumerate(pair_comps):
        c1, c2 = k // (cats * M), (k // M) % cats
        Mk = pair_loglik(ds, [(P, pi, 0.0)], maxbp=maxbp, mask=mask)
        T = np.where(upper & np.isfinite(Mk), lpp + lw + Mk, -np.inf)
        out[c1] = np.logaddexp(out[c1], np.logaddexp.reduce(T, axis=1))
        out[c2] = np.logaddexp(out[c2], np.logaddexp.reduce(T, axis=0))
    return out
