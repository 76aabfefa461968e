"""ctypes binding of ``libmessi.so`` (the C ABI declared in include/messi.h).

The shared library is built in-tree by ``messi_amd.build``.  There is no CPU
fallback: loading fails loudly when the library is missing, and every call
raises ``MessiError`` with the library's message when it returns non-zero.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import threading
import weakref
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MESSI_LIBRARY selects another build of the same library (for comparing builds).
LIBPATH = os.environ.get("MESSI_LIBRARY") or os.path.join(HERE, "libmessi.so")

FP64 = 0
FP32 = 1

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_i64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every entry point declared in include/messi.h.
SIGNATURES = {
    "messi_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "messi_ctx_destroy": (None, [_vp]),
    "messi_last_error": (ctypes.c_char_p, [_vp]),
    "messi_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_selftest": (ctypes.c_int, [_vp]),
    "messi_dataset_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip,
                                            ctypes.c_int, _u8, ctypes.POINTER(_vp)]),
    "messi_dataset_destroy": (None, [_vp]),
    "messi_dataset_stats": (ctypes.c_int, [_vp, _i64]),
    "messi_site_loglik": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "messi_pair_loglik": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "messi_inside": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    _dp, _dp]),
    "messi_inside_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, _dp, _dp]),
    "messi_outside": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "messi_bpp": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "messi_eval": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                  ctypes.c_double, ctypes.c_int, _dp]),
    "messi_posterior": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                       ctypes.c_int, _dp, _dp, _dp]),
    "messi_last_timings": (ctypes.c_int, [_vp, _dp]),
    "messi_mea": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp, _ip]),
    "messi_mea_timings": (ctypes.c_int, [_vp, _dp]),
    "messi_io_error": (ctypes.c_char_p, []),
    "messi_alignment_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "messi_alignment_destroy": (None, [_vp]),
    "messi_alignment_info": (ctypes.c_int, [_vp, _ip, _ip, _ip]),
    "messi_alignment_data": (ctypes.c_int, [_vp, _dp, _u8, _dp, _dp]),
    "messi_alignment_tree": (ctypes.c_int, [_vp, _ip, _ip, _dp]),
    "messi_alignment_name": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "messi_dataset_from_alignment": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u8, ctypes.POINTER(_vp)]),
    "messi_read_ct": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _ip, _ip, ctypes.c_char_p]),
    "messi_parse_dbn": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _ip, _ip]),
    "messi_structure_loglik": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _ip,
                                              ctypes.c_int, _dp, _dp]),
    "messi_sample_states": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                           ctypes.c_int, _ip, ctypes.c_double, ctypes.c_uint64, ctypes.c_int,
                                           _ip, _dp, _dp, _dp, _dp]),
    "messi_site_rate_posteriors": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                                  ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]),
    "messi_sample_structures": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_uint64, ctypes.c_double, _ip, _dp]),
    "messi_sample_posterior": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, _dp,
                                              _ip, _dp]),
    "messi_posterior_consensus": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _ip]),
    "messi_set_tile": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_set_bulk_window": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_set_band_skip": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_set_single_launch": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_set_pair_fusion": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_batch_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "messi_batch_destroy": (None, [_vp]),
    "messi_batch_eval": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                        ctypes.c_double, ctypes.c_int, _dp]),
    "messi_batch_timings": (ctypes.c_int, [_vp, _dp]),
    "messi_batch_step_timings": (ctypes.c_int, [_vp, _dp, _ip]),
    "messi_batch_load": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    "messi_batch_eval_loaded": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp]),
    "messi_batch_load_eigen": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp,
                                              _dp, _dp, _dp, _dp, ctypes.c_int]),
    "messi_batch_set_tiles": (ctypes.c_int, [_vp, _ip]),
    "messi_batch_set_groups": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_batch_set_window": (ctypes.c_int, [_vp, ctypes.c_int]),
    "messi_host_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "messi_host_free": (ctypes.c_int, [_vp, _vp]),
    "messi_mlp_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "messi_mlp_destroy": (None, [_vp]),
    "messi_mlp_num_params": (ctypes.c_int, [_vp, _i64]),
    "messi_mlp_get_params": (ctypes.c_int, [_vp, _fp]),
    "messi_mlp_set_params": (ctypes.c_int, [_vp, _fp]),
    "messi_mlp_forward": (ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp]),
    "messi_mlp_loss_grad": (ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, _fp, _fp]),
    "messi_mlp_train_step": (ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, ctypes.c_float, ctypes.c_float, _fp]),
}

_lib = None
_lock = threading.Lock()
_live = weakref.WeakSet()      # device objects to release before the HIP runtime unloads
_shutdown = False


@atexit.register
def _close_all():
    global _shutdown
    for o in list(_live):
        if isinstance(o, Batch):
            o.close()
    for o in list(_live):
        if isinstance(o, (DeviceDataset, Mlp)):
            o.close()
    for o in list(_live):
        o.close()
    _shutdown = True


class MessiError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load libmessi.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBPATH):
                raise MessiError(f"{LIBPATH} is missing: build it with `python -m messi_amd.build`")
            lib = ctypes.CDLL(LIBPATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            _lib = lib
    return _lib


def pinned_empty(shape, dtype=np.float64, order: str = "C", ctx: Optional["Context"] = None) -> np.ndarray:
    """Uninitialised array in page-locked host memory (messi_host_alloc), freed
    when the array and its views are gone.  Device-to-host copies into it run at
    the link's full rate."""
    ctx = ctx or context()
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    p = ctypes.c_void_p()
    ctx.check(ctx.lib.messi_host_alloc(ctx.handle, max(n, 1), ctypes.byref(p)))
    buf = (ctypes.c_char * max(n, 1)).from_address(p.value)
    fin = weakref.finalize(buf, ctx.lib.messi_host_free, None, ctypes.c_void_p(p.value))
    fin.atexit = False  # at exit the process's memory goes with it
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape, order=order)


def ptr(a: Optional[np.ndarray], t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def prec_code(prec) -> int:
    if prec in (FP64, "fp64", "float64", np.float64):
        return FP64
    if prec in (FP32, "fp32", "float32", np.float32):
        return FP32
    raise ValueError(f"unknown precision {prec!r}")


class Context:
    """One HIP device context + stream (one per host thread)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _vp()
        rc = lib.messi_ctx_create(int(device), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = lib.messi_last_error(h) if h else b"messi_ctx_create failed"
            if h:
                lib.messi_ctx_destroy(h)
            self._h = None
            raise MessiError((msg or b"").decode())
        self.lib = lib
        _live.add(self)

    @property
    def handle(self):
        return self._h

    def check(self, rc: int) -> None:
        if rc != 0:
            raise MessiError(self.lib.messi_last_error(self._h).decode())

    def set_profiling(self, on: bool) -> None:
        self.check(self.lib.messi_set_profiling(self._h, int(bool(on))))

    def selftest(self) -> None:
        self.check(self.lib.messi_selftest(self._h))

    def set_tile(self, tile: int) -> None:
        self.check(self.lib.messi_set_tile(self._h, int(tile)))

    def set_bulk_window(self, window: int) -> None:
        """Launches per bulk accumulator visit of the inside schedule (1, 2 or 4; default 4)."""
        self.check(self.lib.messi_set_bulk_window(self._h, int(window)))

    def set_band_skip(self, on: bool) -> None:
        """Whether inside/outside passes skip the terms a pairing band makes -inf
        (default on; results do not depend on it)."""
        self.check(self.lib.messi_set_band_skip(self._h, int(bool(on))))

    def set_single_launch(self, on: bool) -> None:
        """Whether inside passes run as one launch with per-tile dependency flags
        instead of one launch per tile-diagonal (default off; results are
        bit-identical)."""
        self.check(self.lib.messi_set_single_launch(self._h, int(bool(on))))

    def set_pair_fusion(self, on: bool) -> None:
        """Whether datasets created later compute the pair nodes whose entries
        mostly serve one parent entry inside the parent's launch instead of
        storing them (default on; results are bit-identical)."""
        self.check(self.lib.messi_set_pair_fusion(self._h, int(bool(on))))

    def close(self) -> None:
        if self._h:
            self.lib.messi_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if _shutdown:
            return
        try:
            self.close()
        except Exception:
            pass


_ctx: Optional[Context] = None


def context() -> Context:
    """Process-wide default context on device 0."""
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def colmajor_P(P: np.ndarray) -> np.ndarray:
    """(..., nnodes, dim, dim) [a, d] -> contiguous column-major-per-matrix doubles."""
    P = np.asarray(P, dtype=np.float64)
    return np.ascontiguousarray(np.swapaxes(P, -1, -2))


def _components(P, pi, logw, nnodes: int, dim: int, what: str):
    """Column-major transition matrices (K, nnodes, dim, dim) and the matching
    (K, dim) frequencies and (K,) log weights, shape-checked (the library reads
    exactly K * nnodes * dim^2, K * dim and K values from these pointers)."""
    Pc = colmajor_P(P)
    if Pc.ndim == 3:
        Pc = Pc[None]
    if Pc.ndim != 4 or Pc.shape[1:] != (nnodes, dim, dim):
        raise ValueError(f"{what} transition matrices must have shape (K, {nnodes}, {dim}, {dim}), got "
                         f"{np.shape(P)}")
    K = Pc.shape[0]
    pi_ = np.ascontiguousarray(pi, dtype=np.float64)
    lw = np.ascontiguousarray(logw, dtype=np.float64)
    if pi_.size != K * dim or lw.size != K:
        raise ValueError(f"{what}: {K} components need {K * dim} frequencies and {K} log weights, got "
                         f"{pi_.size} and {lw.size}")
    return K, Pc, pi_.reshape(K, dim), lw.reshape(K)


def _square(a, L: int, what: str) -> np.ndarray:
    m = np.asfortranarray(a, dtype=np.float64)
    if m.shape != (L, L):
        raise ValueError(f"{what} must be {L} x {L}, got {m.shape}")
    return m


def _table3(a, L: int, what: str) -> np.ndarray:
    t = np.asfortranarray(a, dtype=np.float64)
    if t.shape != (3, L, L):
        raise ValueError(f"{what} must have shape (3, {L}, {L}), got {t.shape}")
    return t


class DeviceDataset:
    """Device-resident dataset (alignment, tree, dedup lists) for one distance bound."""

    def __init__(self, ds, maxbasepairdistance: int = 1000000, mask: Optional[np.ndarray] = None,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        self.L = ds.numcols
        self.nnodes = ds.numnodes
        data = np.asfortranarray(ds.data, dtype=np.float64)
        ch = np.ascontiguousarray(ds.tree.children.reshape(-1), dtype=np.int32)
        sq = np.ascontiguousarray(ds.tree.seqindex, dtype=np.int32)
        m = None if mask is None else np.asfortranarray(mask > 0, dtype=np.uint8)
        self.maxbp = int(min(maxbasepairdistance, 2**31 - 1))
        h = _vp()
        self.ctx.check(self.ctx.lib.messi_dataset_create(
            self.ctx.handle, ds.numseqs, ds.numcols, ptr(data), ds.numnodes, ptr(ch, _ip), ptr(sq, _ip),
            self.maxbp, ptr(m, _u8), ctypes.byref(h)))
        self._h = h
        _live.add(self)

    @classmethod
    def _adopt(cls, ctx, h, L: int, nnodes: int, maxbp: int) -> "DeviceDataset":
        self = cls.__new__(cls)
        self.ctx = ctx
        self.L = L
        self.nnodes = nnodes
        self.maxbp = maxbp
        self._h = h
        _live.add(self)
        return self

    @property
    def handle(self):
        return self._h

    def stats(self) -> dict:
        a = np.zeros(6, dtype=np.int64)
        self.ctx.check(self.ctx.lib.messi_dataset_stats(self._h, ptr(a, _i64)))
        return {"sum_U": int(a[0]), "root_entries": int(a[1]), "pairs_in_range": int(a[2]),
                "sum_site_patterns": int(a[3]), "levels": int(a[4]), "device_bytes": int(a[5])}

    def timings(self) -> np.ndarray:
        t = np.zeros(5)
        self.ctx.check(self.ctx.lib.messi_last_timings(self._h, ptr(t)))
        return t

    def _mixture(self, site, pair):
        return (_components(*site, self.nnodes, 4, "site") + _components(*pair, self.nnodes, 16, "pair"))

    def site_loglik(self, P4, pi4, logw, prec=FP64) -> np.ndarray:
        K, P, pi, lw = _components(P4, pi4, logw, self.nnodes, 4, "site")
        out = np.empty(self.L)
        self.ctx.check(self.ctx.lib.messi_site_loglik(self._h, K, ptr(P), ptr(pi), ptr(lw), prec_code(prec),
                                                      ptr(out)))
        return out

    def pair_loglik(self, P16, pi16, logw, prec=FP64, percomp: bool = False):
        K, P, pi, lw = _components(P16, pi16, logw, self.nnodes, 16, "pair")
        out = np.empty((self.L, self.L), order="F")
        pc = np.empty((K, self.L, self.L)) if percomp else None
        self.ctx.check(self.ctx.lib.messi_pair_loglik(self._h, K, ptr(P), ptr(pi), ptr(lw), prec_code(prec),
                                                      ptr(out), ptr(pc)))
        if percomp:
            return out, [np.asfortranarray(pc[k].T) for k in range(K)]
        return out

    def eval(self, site, pair, B: float = 1.0, prec=FP64) -> float:
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        z = ctypes.c_double()
        self.ctx.check(self.ctx.lib.messi_eval(self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c),
                                               ptr(pi16), ptr(lw16), float(B), prec_code(prec),
                                               ctypes.byref(z)))
        return z.value

    def posterior(self, site, pair, prec=FP64, want_pairs: bool = True, out: Optional[np.ndarray] = None):
        """log Z, unpaired posteriors (L) and the L x L pair posteriors (Julia
        layout).  out: an L x L float64 Fortran-ordered array to receive the pair
        posteriors (e.g. from pinned_empty, reused across calls)."""
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        z = ctypes.c_double()
        unp = np.empty(self.L)
        if out is not None:
            if out.shape != (self.L, self.L) or out.dtype != np.float64 or not out.flags.f_contiguous:
                raise ValueError(f"out must be a Fortran-ordered float64 array of shape {(self.L, self.L)}")
            pp = out
        else:
            pp = np.empty((self.L, self.L), order="F") if want_pairs else None
        self.ctx.check(self.ctx.lib.messi_posterior(self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c),
                                                    ptr(pi16), ptr(lw16), prec_code(prec), ctypes.byref(z),
                                                    ptr(unp), ptr(pp)))
        return z.value, unp, pp

    def posterior_consensus(self, site, pair, prec=FP64, mea_prec=FP64, alpha: float = 2.0):
        """log Z and the MEA consensus pairing (1-based partners, 0 unpaired), with
        the posteriors kept on the device."""
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        z = ctypes.c_double()
        paired = np.empty(self.L, dtype=np.intc)
        self.ctx.check(self.ctx.lib.messi_posterior_consensus(
            self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c), ptr(pi16), ptr(lw16), prec_code(prec),
            prec_code(mea_prec), float(alpha), ctypes.byref(z), paired.ctypes.data_as(_ip)))
        return z.value, paired.astype(np.int64)

    def structure_loglik(self, site, pair, paired, prec=FP64):
        """(log-likelihood of the fixed structure, (unpaired sum, pair sum)); paired
        holds 1-based partners, 0 unpaired."""
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        pa = np.ascontiguousarray(paired, dtype=np.intc)
        if pa.shape != (self.L,):
            raise ValueError(f"paired must have length {self.L}")
        ll = ctypes.c_double()
        parts = np.zeros(2)
        self.ctx.check(self.ctx.lib.messi_structure_loglik(
            self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c), ptr(pi16), ptr(lw16),
            pa.ctypes.data_as(_ip), prec_code(prec), ctypes.byref(ll), ptr(parts)))
        return ll.value, (float(parts[0]), float(parts[1]))

    def sample_states(self, site, pair, paired, cats: int, seed: int = 0, B: float = 1.0, prec=FP64,
                      conditionals: bool = False):
        """Site-category states given the structure paired (1-based partners):
        (states 1-based, unpaired log-likelihoods, paired log-likelihoods at the
        openers), plus (site conditionals K4 x L, pair components K16 x L) when
        conditionals is set (messi_sample_states)."""
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        pa = np.ascontiguousarray(paired, dtype=np.intc)
        if pa.shape != (self.L,):
            raise ValueError(f"paired must have length {self.L}")
        st = np.zeros(self.L, dtype=np.intc)
        unp = np.zeros(self.L)
        pll = np.zeros(self.L)
        sc = np.zeros((K4, self.L)) if conditionals else None
        pc = np.zeros((K16, self.L)) if conditionals else None
        self.ctx.check(self.ctx.lib.messi_sample_states(
            self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c), ptr(pi16), ptr(lw16), int(cats),
            pa.ctypes.data_as(_ip), float(B), int(seed), prec_code(prec), st.ctypes.data_as(_ip), ptr(unp),
            ptr(pll), ptr(sc), ptr(pc)))
        out = (st.astype(np.int64), unp, pll)
        return out + (sc, pc) if conditionals else out

    def site_rate_posteriors(self, site, pair, cats: int, unpaired, pairprob, prec=FP64) -> np.ndarray:
        """(cats, L) log posterior weights of the site-rate categories of every
        column (messi_site_rate_posteriors) from the base-pair posteriors."""
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        u = np.ascontiguousarray(unpaired, dtype=np.float64)
        if u.shape != (self.L,):
            raise ValueError(f"unpaired must have length {self.L}")
        pp = _square(pairprob, self.L, "pairprob")
        out = np.empty((int(cats), self.L), order="F")
        self.ctx.check(self.ctx.lib.messi_site_rate_posteriors(
            self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c), ptr(pi16), ptr(lw16), int(cats), ptr(u),
            ptr(pp), prec_code(prec), ptr(out)))
        return out

    def sample_posterior(self, site, pair, nsamples: int, seed: int = 0, B: float = 1.0, prec=FP64):
        """log Z, then nsamples structures (nsamples x L, 1-based partners) drawn from
        the evaluation's inside table on the device, and their KH99 log priors."""
        K4, P4c, pi4, lw4, K16, P16c, pi16, lw16 = self._mixture(site, pair)
        z = ctypes.c_double()
        out = np.empty((nsamples, self.L), dtype=np.intc)
        pri = np.empty(nsamples)
        self.ctx.check(self.ctx.lib.messi_sample_posterior(
            self._h, K4, ptr(P4c), ptr(pi4), ptr(lw4), K16, ptr(P16c), ptr(pi16), ptr(lw16), prec_code(prec),
            int(nsamples), int(seed) & 0xFFFFFFFFFFFFFFFF, float(B), ctypes.byref(z), out.ctypes.data_as(_ip),
            ptr(pri)))
        return z.value, out.astype(np.int64), pri

    def close(self) -> None:
        if self._h:
            self.ctx.lib.messi_dataset_destroy(self._h)
            self._h = None

    def __del__(self):
        if _shutdown:
            return
        try:
            self.close()
        except Exception:
            pass


def inside(u, paired, B: float = 1.0, zonly: bool = True, prec=FP64, ctx: Optional[Context] = None):
    ctx = ctx or context()
    u = np.ascontiguousarray(u, dtype=np.float64)
    L = len(u)
    pm = _square(paired, L, "paired")
    z = ctypes.c_double()
    t = None if zonly else np.empty((3, L, L), order="F")
    ctx.check(ctx.lib.messi_inside(ctx.handle, L, ptr(u), ptr(pm), float(B), int(bool(zonly)), prec_code(prec),
                                   ctypes.byref(z), ptr(t)))
    return z.value if zonly else (z.value, t)


NOALLOC_MAX_L = 512


def inside_noalloc(u, paired, B: float = 1.0, zonly: bool = True, prec=FP64, ctx: Optional[Context] = None):
    """inside() through messi_inside_noalloc: no device allocation, L <= 512."""
    ctx = ctx or context()
    u = np.ascontiguousarray(u, dtype=np.float64)
    L = len(u)
    pm = _square(paired, L, "paired")
    z = ctypes.c_double()
    t = None if zonly else np.empty((3, L, L), order="F")
    ctx.check(ctx.lib.messi_inside_noalloc(ctx.handle, L, ptr(u), ptr(pm), float(B), int(bool(zonly)),
                                           prec_code(prec), ctypes.byref(z), ptr(t)))
    return z.value if zonly else (z.value, t)


def outside(inside_t, u, paired, prec=FP64, ctx: Optional[Context] = None):
    ctx = ctx or context()
    u = np.ascontiguousarray(u, dtype=np.float64)
    L = len(u)
    it = _table3(inside_t, L, "inside table")
    pm = _square(paired, L, "paired")
    out = np.empty((3, L, L), order="F")
    ctx.check(ctx.lib.messi_outside(ctx.handle, L, ptr(it), ptr(u), ptr(pm), prec_code(prec), ptr(out)))
    return out


def bpp(inside_t, outside_t, u, paired, prec=FP64, ctx: Optional[Context] = None):
    ctx = ctx or context()
    u = np.ascontiguousarray(u, dtype=np.float64)
    L = len(u)
    it = _table3(inside_t, L, "inside table")
    ot = _table3(outside_t, L, "outside table")
    pm = _square(paired, L, "paired")
    unp = np.empty(L)
    pp = np.empty((L, L), order="F")
    ctx.check(ctx.lib.messi_bpp(ctx.handle, L, ptr(it), ptr(ot), ptr(u), ptr(pm), prec_code(prec), ptr(unp),
                                ptr(pp)))
    return unp, pp


def mea(unpaired, pairprob, alpha: float = 2.0, prec=FP64, tables: bool = False,
        ctx: Optional[Context] = None):
    """MEA consensus pairing (1-based partners, 0 unpaired) from base-pair
    posteriors; with tables, also the E (float64) and S (int) tables, L x L."""
    ctx = ctx or context()
    u = np.ascontiguousarray(unpaired, dtype=np.float64)
    L = len(u)
    pm = np.asfortranarray(pairprob, dtype=np.float64)
    if pm.shape != (L, L):
        raise ValueError(f"pairprob must be {L} x {L}")
    paired = np.empty(L, dtype=np.intc)
    E = np.empty((L, L), order="F") if tables else None
    S = np.empty((L, L), dtype=np.intc, order="F") if tables else None
    ctx.check(ctx.lib.messi_mea(ctx.handle, L, ptr(u), ptr(pm), float(alpha), prec_code(prec),
                                paired.ctypes.data_as(_ip), ptr(E),
                                S.ctypes.data_as(_ip) if tables else None))
    if tables:
        return paired.astype(np.int64), E, S.astype(np.int64)
    return paired.astype(np.int64)


def sample_structures(inside_t, paired, unpaired, nsamples: int, seed: int = 0, B: float = 1.0, prec=FP64,
                      ctx: Optional[Context] = None):
    """nsamples structures (nsamples x L, 1-based partners, 0 unpaired) drawn from a
    3 x L x L inside table, and their KH99 log priors."""
    ctx = ctx or context()
    u = np.ascontiguousarray(unpaired, dtype=np.float64)
    L = len(u)
    it = np.asfortranarray(inside_t, dtype=np.float64)
    pm = np.asfortranarray(paired, dtype=np.float64)
    if it.shape != (3, L, L) or pm.shape != (L, L):
        raise ValueError("inside must be 3 x L x L and paired L x L")
    out = np.empty((nsamples, L), dtype=np.intc)
    pri = np.empty(nsamples)
    ctx.check(ctx.lib.messi_sample_structures(ctx.handle, L, ptr(it), ptr(pm), ptr(u), prec_code(prec), int(nsamples),
                                              int(seed) & 0xFFFFFFFFFFFFFFFF, float(B), out.ctypes.data_as(_ip),
                                              ptr(pri)))
    return out.astype(np.int64), pri


def mea_timings(ctx: Optional[Context] = None) -> np.ndarray:
    """[upload, fill, traceback] milliseconds of the last MEA call (profiling on)."""
    ctx = ctx or context()
    t = np.zeros(3)
    ctx.check(ctx.lib.messi_mea_timings(ctx.handle, ptr(t)))
    return t


def _io_check(rc: int) -> None:
    if rc != 0:
        msg = load().messi_io_error()
        raise MessiError(msg.decode() if msg else f"error {rc}")


class Alignment:
    """FASTA alignment + Newick tree parsed by the library (messi_alignment_*): state
    masks, gap and base frequencies, the rooted pre-order tree."""

    def __init__(self, fasta_path: str, newick_path: str):
        L = load()
        h = _vp()
        _io_check(L.messi_alignment_read(os.fsencode(fasta_path), os.fsencode(newick_path), ctypes.byref(h)))
        self._h = h
        n, c, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.messi_alignment_info(h, ctypes.byref(n), ctypes.byref(c), ctypes.byref(m))
        self.numseqs, self.numcols, self.numnodes = n.value, c.value, m.value
        self.codes = np.empty((self.numseqs, self.numcols), dtype=np.uint8)
        self.gapfrequency = np.empty(self.numcols)
        self.obsfreqs = np.empty(4)
        L.messi_alignment_data(h, None, ptr(self.codes, _u8), ptr(self.gapfrequency), ptr(self.obsfreqs))
        self.children = np.empty((self.numnodes, 2), dtype=np.intc)
        self.seqindex = np.empty(self.numnodes, dtype=np.intc)
        self.branchlengths = np.empty(self.numnodes)
        L.messi_alignment_tree(h, ptr(self.children, _ip), ptr(self.seqindex, _ip), ptr(self.branchlengths))
        buf = ctypes.create_string_buffer(4096)
        self.names = []
        for r in range(self.numseqs):
            L.messi_alignment_name(h, r, buf, 4096)
            self.names.append(buf.value.decode())

    def data(self) -> np.ndarray:
        """(nseq, ncol, 4) 0/1 state indicators (the reference's Dataset.data)."""
        out = np.empty((self.numseqs, self.numcols, 4), order="F")
        load().messi_alignment_data(self._h, ptr(out), None, None, None)
        return out

    def device_dataset(self, maxbasepairdistance: int = 1000000, mask: Optional[np.ndarray] = None,
                       ctx: Optional[Context] = None) -> "DeviceDataset":
        ctx = ctx or context()
        m = None if mask is None else np.asfortranarray(mask > 0, dtype=np.uint8)
        maxbp = int(min(maxbasepairdistance, 2**31 - 1))
        h = _vp()
        ctx.check(ctx.lib.messi_dataset_from_alignment(ctx.handle, self._h, maxbp, ptr(m, _u8), ctypes.byref(h)))
        return DeviceDataset._adopt(ctx, h, self.numcols, self.numnodes, maxbp)

    def close(self) -> None:
        if self._h:
            load().messi_alignment_destroy(self._h)
            self._h = None

    def __del__(self):
        if _shutdown:
            return
        try:
            self.close()
        except Exception:
            pass


def read_ct(path: str):
    """(sequence, 1-based partners) of a .ct file (readctfile)."""
    L = load()
    n = ctypes.c_int()
    _io_check(L.messi_read_ct(os.fsencode(path), 0, ctypes.byref(n), None, None))
    paired = np.empty(n.value, dtype=np.intc)
    seq = ctypes.create_string_buffer(n.value + 1)
    _io_check(L.messi_read_ct(os.fsencode(path), n.value, ctypes.byref(n), paired.ctypes.data_as(_ip), seq))
    return seq.value.decode(), paired.astype(np.int64)


def parse_dbn(dbn: str) -> np.ndarray:
    """1-based partners of a dot-bracket string (parsedbn)."""
    L = load()
    n = ctypes.c_int()
    b = dbn.encode()
    _io_check(L.messi_parse_dbn(b, 0, ctypes.byref(n), None))
    paired = np.empty(n.value, dtype=np.intc)
    _io_check(L.messi_parse_dbn(b, n.value, ctypes.byref(n), paired.ctypes.data_as(_ip)))
    return paired.astype(np.int64)


class Batch:
    """Several (dataset, parameter set) members evaluated together (messi_batch_*)."""

    def __init__(self, members, ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        self.members = list(members)            # DeviceDataset objects (may repeat)
        arr = (_vp * len(self.members))(*[m.handle for m in self.members])
        h = _vp()
        self.ctx.check(self.ctx.lib.messi_batch_create(self.ctx.handle, len(self.members), arr, ctypes.byref(h)))
        self._h = h
        _live.add(self)

    def _inputs(self, sites, pairs):
        """Concatenated member-major inputs; every member's arrays are checked
        against its own node count, and all members must have the same K."""
        n = len(self.members)
        if len(sites) != n or len(pairs) != n:
            raise ValueError(f"expected inputs for {n} members, got {len(sites)} and {len(pairs)}")
        s4 = [_components(*x, m.nnodes, 4, "site") for x, m in zip(sites, self.members)]
        s16 = [_components(*x, m.nnodes, 16, "pair") for x, m in zip(pairs, self.members)]
        K4, K16 = s4[0][0], s16[0][0]
        if any(c[0] != K4 for c in s4) or any(c[0] != K16 for c in s16):
            raise ValueError("every member must have the same number of site and pair components")
        cat = lambda cs, i: np.ascontiguousarray(np.concatenate([c[i].reshape(-1) for c in cs]))
        return cat(s4, 1), cat(s4, 2), cat(s4, 3), K4, cat(s16, 1), cat(s16, 2), cat(s16, 3), K16

    def load(self, sites, pairs, prec=FP64) -> None:
        """Copy the members' mixture components to the device (messi_batch_load)."""
        P4, pi4, lw4, K4, P16, pi16, lw16, K16 = self._inputs(sites, pairs)
        self.ctx.check(self.ctx.lib.messi_batch_load(self._h, K4, ptr(P4), ptr(pi4), ptr(lw4), K16, ptr(P16),
                                                     ptr(pi16), ptr(lw16), prec_code(prec)))

    def load_eigen(self, V4, w4, Vi4, pi4, lw4, V16, w16, Vi16, pi16, lw16, t, prec=FP64) -> None:
        """messi_batch_load_eigen: the members' transition matrices built on the
        device from eigendecompositions Q = V diag(w) Vi.  Stacked inputs, one row
        per member: V, Vi (n, K, dim, dim) indexed [a, c], w (n, K, dim), pi
        (n, K, dim), log weights (n, K), branch lengths t (n, nn) or a list of
        each member's nn_m values."""
        n = len(self.members)
        tt = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in t]))
        if tt.size != sum(m.nnodes for m in self.members):
            raise ValueError("branch lengths must give every node of every member")

        def model(V, w, Vi, pi, lw, dim, what):
            V, Vi = np.asarray(V, dtype=np.float64), np.asarray(Vi, dtype=np.float64)
            if V.ndim != 4 or V.shape[0] != n or V.shape[2:] != (dim, dim) or Vi.shape != V.shape:
                raise ValueError(f"{what} eigenvectors must have shape ({n}, K, {dim}, {dim})")
            K = V.shape[1]
            w = np.ascontiguousarray(w, dtype=np.float64)
            pi = np.ascontiguousarray(pi, dtype=np.float64)
            lw = np.ascontiguousarray(lw, dtype=np.float64)
            if w.shape != (n, K, dim) or pi.shape != (n, K, dim) or lw.shape != (n, K):
                raise ValueError(f"{what}: eigenvalues, frequencies and log weights must match ({n}, {K})")
            return (K, np.ascontiguousarray(np.swapaxes(V, -1, -2)), w,
                    np.ascontiguousarray(np.swapaxes(Vi, -1, -2)), pi, lw)
        K4, a4, b4, c4, d4, e4 = model(V4, w4, Vi4, pi4, lw4, 4, "site")
        K16, a16, b16, c16, d16, e16 = model(V16, w16, Vi16, pi16, lw16, 16, "pair")
        self.ctx.check(self.ctx.lib.messi_batch_load_eigen(
            self._h, K4, ptr(a4), ptr(b4), ptr(c4), ptr(d4), ptr(e4), K16, ptr(a16), ptr(b16), ptr(c16),
            ptr(d16), ptr(e16), ptr(tt), prec_code(prec)))

    def set_tiles(self, tiles) -> None:
        """Inside tile edge per member (16, 32, or 0 for the context's); None resets."""
        if tiles is None:
            self.ctx.check(self.ctx.lib.messi_batch_set_tiles(self._h, None))
            return
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        if t.shape != (len(self.members),):
            raise ValueError(f"expected {len(self.members)} tiles, got shape {t.shape}")
        self.ctx.check(self.ctx.lib.messi_batch_set_tiles(self._h, ptr(t, _ip)))

    def set_groups(self, groups: int) -> None:
        """Member groups of a pipelined evaluation (1..4; 0 = automatic)."""
        self.ctx.check(self.ctx.lib.messi_batch_set_groups(self._h, int(groups)))

    def set_window(self, window: int) -> None:
        """Bulk window of this batch's inside schedule (1, 2, 4; 0 = the context's)."""
        self.ctx.check(self.ctx.lib.messi_batch_set_window(self._h, int(window)))

    def eval_loaded(self, B: float = 1.0, prec=FP64) -> np.ndarray:
        """log Z of every member from the components of the last load."""
        z = np.empty(len(self.members))
        self.ctx.check(self.ctx.lib.messi_batch_eval_loaded(self._h, float(B), prec_code(prec), ptr(z)))
        return z

    def eval(self, sites, pairs, B: float = 1.0, prec=FP64) -> np.ndarray:
        """sites[m] = (P4 (K,nn,4,4), pi4 (K,4), logw (K,)); pairs[m] likewise with 16 states."""
        P4, pi4, lw4, K4, P16, pi16, lw16, K16 = self._inputs(sites, pairs)
        z = np.empty(len(self.members))
        self.ctx.check(self.ctx.lib.messi_batch_eval(self._h, K4, ptr(P4), ptr(pi4), ptr(lw4), K16, ptr(P16),
                                                     ptr(pi16), ptr(lw16), float(B), prec_code(prec), ptr(z)))
        return z

    def eval_stacked(self, P4, pi4, lw4, P16, pi16, lw16, B: float = 1.0, prec=FP64) -> np.ndarray:
        """Stacked inputs: P4 (n, K4, nn, 4, 4), pi4 (n, K4, 4), lw4 (n, K4) and the
        16-state arrays likewise; every member must have nn nodes."""
        n = len(self.members)
        P4 = colmajor_P(P4)
        P16 = colmajor_P(P16)
        nn = {m.nnodes for m in self.members}
        if len(nn) != 1:
            raise ValueError("stacked inputs need members with one node count")
        nn = nn.pop()
        if P4.ndim != 5 or P4.shape[0] != n or P4.shape[2:] != (nn, 4, 4):
            raise ValueError(f"P4 must have shape ({n}, K4, {nn}, 4, 4), got {P4.shape}")
        if P16.ndim != 5 or P16.shape[0] != n or P16.shape[2:] != (nn, 16, 16):
            raise ValueError(f"P16 must have shape ({n}, K16, {nn}, 16, 16), got {P16.shape}")
        K4, K16 = P4.shape[1], P16.shape[1]
        pi4 = np.ascontiguousarray(pi4, dtype=np.float64)
        lw4 = np.ascontiguousarray(lw4, dtype=np.float64)
        pi16 = np.ascontiguousarray(pi16, dtype=np.float64)
        lw16 = np.ascontiguousarray(lw16, dtype=np.float64)
        if pi4.shape != (n, K4, 4) or lw4.shape != (n, K4) or pi16.shape != (n, K16, 16) or lw16.shape != (n, K16):
            raise ValueError("frequencies and log weights must match the stacked matrices")
        z = np.empty(n)
        self.ctx.check(self.ctx.lib.messi_batch_eval(self._h, K4, ptr(P4), ptr(pi4), ptr(lw4), K16, ptr(P16),
                                                     ptr(pi16), ptr(lw16), float(B), prec_code(prec), ptr(z)))
        return z

    def timings(self) -> np.ndarray:
        t = np.zeros(5)
        self.ctx.check(self.ctx.lib.messi_batch_timings(self._h, ptr(t)))
        return t

    def step_timings(self):
        """(ms, launches) of the inside step kernels of the last profiled
        evaluation, per tile edge (16, 32)."""
        ms = np.zeros(2)
        n = np.zeros(2, dtype=np.intc)
        self.ctx.check(self.ctx.lib.messi_batch_step_timings(self._h, ptr(ms), ptr(n, _ip)))
        return ms, n

    def close(self) -> None:
        if self._h:
            self.ctx.lib.messi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _shutdown:
            return
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """Proposal network trained and run on the device (messi_mlp_*)."""

    def __init__(self, din: int, dh: int = 128, seed: int = 0, ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        self.din, self.dh = int(din), int(dh)
        h = _vp()
        self.ctx.check(self.ctx.lib.messi_mlp_create(self.ctx.handle, self.din, self.dh, int(seed),
                                                     ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.messi_mlp_num_params(self._h, ctypes.byref(n)))
        self.num_params = int(n.value)
        _live.add(self)

    def _rows(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.din:
            raise ValueError(f"expected rows of {self.din} values, got shape {x.shape}")
        return x

    def get_params(self) -> np.ndarray:
        w = np.empty(self.num_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.messi_mlp_get_params(self._h, ptr(w, _fp)))
        return w

    def set_params(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.shape != (self.num_params,):
            raise ValueError(f"expected {self.num_params} parameters, got shape {w.shape}")
        self.ctx.check(self.ctx.lib.messi_mlp_set_params(self._h, ptr(w, _fp)))

    def forward(self, x) -> np.ndarray:
        """Rows of inputs -> rows of (mu, ell), shape (n, 2*din)."""
        x = self._rows(x)
        out = np.empty((x.shape[0], 2 * self.din), dtype=np.float32)
        self.ctx.check(self.ctx.lib.messi_mlp_forward(self._h, x.shape[0], ptr(x, _fp), ptr(out, _fp)))
        return out

    def loss_grad(self, x, t):
        x, t = self._rows(x), self._rows(t)
        if x.shape != t.shape:
            raise ValueError("inputs and targets differ in shape")
        loss = ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.messi_mlp_loss_grad(self._h, x.shape[0], ptr(x, _fp), ptr(t, _fp),
                                                        ctypes.byref(loss), ptr(g, _fp)))
        return float(loss.value), g

    def train_step(self, x, t, lr: float = 1e-3, weight_decay: float = 0.0) -> float:
        x, t = self._rows(x), self._rows(t)
        if x.shape != t.shape:
            raise ValueError("inputs and targets differ in shape")
        loss = ctypes.c_float()
        self.ctx.check(self.ctx.lib.messi_mlp_train_step(self._h, x.shape[0], ptr(x, _fp), ptr(t, _fp),
                                                         float(lr), float(weight_decay), ctypes.byref(loss)))
        return float(loss.value)

    def close(self) -> None:
        if self._h:
            self.ctx.lib.messi_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        if _shutdown:
            return
        try:
            self.close()
        except Exception:
            pass
