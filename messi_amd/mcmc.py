"""MCMC over the model parameters with a learned, Metropolis-Hastings-corrected
proposal and the adaptive random walk as baseline and fallback.

This is the caller side of the likelihood path.  The reference's MH loop
(``runchain``, src/MCMC.jl:675-1868) proposes from random walks and adaptive
Metropolis (``gettunedparams``, src/MCMC.jl:393-449: covariance of the second
half of the tuning history scaled by 2.38^2/d), evaluates
``computetotallikelihood`` (prior + structure-integrated log-likelihood,
src/FARCE.jl:2322-2350) and accepts when exp((ll' - ll) B + propratio) > u
(src/MCMC.jl:1666).  Here many chains advance together: one step evaluates
all chains' proposals in one batched device call (``lib.Batch``).

Parameter vector (12 values, unconstrained):
  0-2   log lambdaGC, log lambdaAT, log lambdaGT
  3-7   log q1..q5, with q6 = 1 (the reference fixes q6 = 1 whenever branch
        lengths are sampled, src/MCMC.jl:1661-1664, because a common factor of
        the rates and of the branch lengths leaves the likelihood unchanged)
  8-10  base-frequency logits of C, G, T against A (pi = softmax(0, ., ., .))
  11    log of a common branch-length scale s (branch b has length s * b0[b])
The target in these coordinates is  logZ + computeprior (with
samplebranchlengths) + log |Jacobian|.

Learned proposal, in whitened coordinates z = chol(C)^-1 (theta - m) with m, C
the mean and covariance of the burn-in samples:
  z' ~ N(z + mu(z), diag (kappa exp(ell(z)))^2),   (mu, ell) = network(z)
accepted with  log u < lp(z') - lp(z) + log q(z | z') - log q(z' | z).
The network (``lib.Mlp``: 12 -> 128 -> 128 -> 24, bf16 weights, FP32
accumulate) is trained on the device on the chains' own burn-in samples as a
denoiser: pairs (z + s eps, z).  The loss-optimal mean shift is then s^2 times
the score of the posterior smoothed by N(0, s^2) (Tweedie's formula), and exp(ell)
the local spread, so the proposal is a Langevin step with a learned drift and a
learned diagonal scale; s and kappa are chosen during burn-in by the expected
squared jump distance.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import model
from .model import BASEPAIRINGS

NAMES = ("log_lambdaGC", "log_lambdaAT", "log_lambdaGT", "log_q1", "log_q2", "log_q3", "log_q4",
         "log_q5", "logit_piC", "logit_piG", "logit_piT", "log_branch_scale")
DIM = 12


# ---------------------------------------------------------------------------
# Parameter map and batched model construction
# ---------------------------------------------------------------------------

def _muse_structure():
    """Single-nucleotide transitions of the 16-state model (src/RateMatrices.jl:69-148)."""
    rows = []
    for h in range(16):
        for v in range(h + 1, 16):
            if h // 4 == v // 4:
                frm, to, second = h % 4, v % 4, 1
            elif h % 4 == v % 4:
                frm, to, second = h // 4, v // 4, 0
            else:
                continue
            rows.append((h, v, frm, to, second, int(BASEPAIRINGS[v]), int(BASEPAIRINGS[h])))
    return np.array(rows, dtype=np.int64).T


_H, _V, _FRM, _TO, _SECOND, _CV, _CH = _muse_structure()
_QIDX = {(0, 1): 0, (0, 2): 1, (0, 3): 2, (1, 2): 3, (1, 3): 4, (2, 3): 5}


def _qmat(q):
    """(n, 6) exchangeabilities -> (n, 4, 4) symmetric matrices with zero diagonal."""
    n = q.shape[0]
    S = np.zeros((n, 4, 4))
    for (i, j), k in _QIDX.items():
        S[:, i, j] = S[:, j, i] = q[:, k]
    return S


def transition_matrices(Q, t):
    """Batched ``model.gettransitionmatriceslist``: Q (n, d, d), t (n, nn) branch
    lengths (t[:, 0] is the root's, unused) -> (n, nn, d, d)."""
    n, d, _ = Q.shape
    w, V = np.linalg.eig(Q)
    Vi = np.linalg.inv(V)
    tt = t[:, 1:]
    E = np.exp(w[:, None, :] * tt[:, :, None])                       # (n, nn-1, d)
    # P[b, node] = sum_j E[b, node, j] * outer(V[b, :, j], Vi[b, j, :]): one GEMM per vector
    M = (V[:, :, :, None] * Vi[:, None, :, :]).transpose(0, 2, 1, 3).reshape(n, d, d * d)
    P = np.real(E @ M).reshape(n, tt.shape[1], d, d)
    P = np.where(P <= 0.0, 1e-50, P)
    zero = tt == 0.0
    if zero.any():
        P[zero] = np.eye(d)
    out = np.zeros((n, t.shape[1], d, d))
    out[:, 1:] = P
    return out


class ParamMap:
    """The 12 coordinates, their prior and the model inputs they define (host only).

    ``base15`` is a 15-vector (``model.getparams`` layout) giving the values of
    the parameters that are not sampled (lambda gamma shape, site gamma shape,
    lambda zero weight); one site-rate and one lambda category (the 12-parameter
    model).  ``b0`` are the tree's branch lengths (pre-order, root first).
    """

    def __init__(self, base15, b0):
        self.base = np.asarray(base15, dtype=np.float64).copy()
        self.b0 = np.asarray(b0, dtype=np.float64).copy()
        bp = model.getparams(self.base, self.b0)
        # constant part of computeprior: everything but the lambda, q and branch terms
        self._prior_const = (model.computeprior(bp, samplebranchlengths=True)
                             + 0.1 * (bp.lambdaGC - 1.0) + 0.1 * (bp.lambdaAT - 1.0) + 0.1 * (bp.lambdaGT - 1.0)
                             + 0.1 * float(np.sum(bp.q)) + 0.2 * float(np.sum(self.b0[1:])))

    # --- parameter map ---------------------------------------------------
    def theta_of(self, params15) -> np.ndarray:
        """Coordinates of a 15-vector with the tree's own branch lengths (q is
        divided by q6 and the branch scale multiplied by it, which leaves every
        transition matrix unchanged)."""
        p = np.asarray(params15, dtype=np.float64)
        q = p[3:9]
        pi = np.array([p[11], p[12], p[13], 1.0 - p[11] - p[12] - p[13]])
        th = np.empty(DIM)
        th[0:3] = np.log(p[0:3])
        th[3:8] = np.log(q[:5] / q[5])
        th[8:11] = np.log(pi[1:] / pi[0])
        th[11] = np.log(q[5])
        return th

    @staticmethod
    def unpack(thetas):
        th = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
        lam = np.exp(th[:, 0:3])
        q = np.concatenate([np.exp(th[:, 3:8]), np.ones((th.shape[0], 1))], axis=1)
        lg = np.concatenate([np.zeros((th.shape[0], 1)), th[:, 8:11]], axis=1)
        lg = lg - lg.max(axis=1, keepdims=True)
        pi = np.exp(lg)
        pi /= pi.sum(axis=1, keepdims=True)
        s = np.exp(th[:, 11])
        return lam, q, pi, s

    def params_of(self, theta):
        """(15-vector, branch lengths) of one coordinate vector."""
        lam, q, pi, s = self.unpack(theta)
        p = self.base.copy()
        p[0:3] = lam[0]
        p[3:9] = q[0]
        p[11:14] = pi[0, :3]
        return p, self.b0 * s[0]

    def log_prior(self, thetas) -> np.ndarray:
        """computeprior (src/FARCE.jl:2401-2436, branch lengths sampled) plus the
        log Jacobian of the map from the coordinates."""
        th = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
        lam, q, pi, s = self.unpack(th)
        lp = (self._prior_const - 0.1 * np.sum(lam - 1.0, axis=1) - 0.1 * np.sum(q, axis=1)
              - 0.2 * s * float(np.sum(self.b0[1:])))
        jac = np.sum(th[:, 0:8], axis=1) + np.sum(np.log(pi), axis=1) + th[:, 11]
        return lp + jac

    # --- model construction (batched over vectors) -----------------------
    def rate_matrices(self, thetas):
        """(Q4 (n,4,4), pi (n,4), Q16 (n,16,16), pi16 (n,16), branch lengths (n, nn))
        of each coordinate vector: GTR (src/RateMatrices.jl:2-37), MuSE with one
        site-rate category (src/RateMatrices.jl:69-148) and its stationary
        frequencies (src/RateMatrices.jl:40-66), the tree's lengths times the scale."""
        lam, q, pi, s = self.unpack(thetas)
        n = lam.shape[0]
        t = s[:, None] * self.b0[None, :]
        S = _qmat(q)
        Q4 = S * pi[:, None, :]
        Q4[:, np.arange(4), np.arange(4)] = 0.0
        Q4[:, np.arange(4), np.arange(4)] = -Q4.sum(axis=2)
        lamf = np.concatenate([np.ones((n, 1)), lam], axis=1)     # index 0: unpaired class
        gq = S[:, _FRM, _TO]
        site = np.ones_like(gq)          # one site-rate category: both rates are 1
        m1 = lamf[:, _CV] / lamf[:, _CH]
        Q16 = np.zeros((n, 16, 16))
        Q16[:, _H, _V] = m1 * gq * site * pi[:, _TO]
        Q16[:, _V, _H] = (1.0 / m1) * gq * site * pi[:, _FRM]
        Q16[:, np.arange(16), np.arange(16)] = -Q16.sum(axis=2)
        piGC = pi[:, 1] * pi[:, 2]
        piAT = pi[:, 0] * pi[:, 3]
        piGT = pi[:, 2] * pi[:, 3]
        kappa = 1.0 / (1.0 + 2.0 * (piAT * (lam[:, 1] ** 2 - 1.0)) + 2.0 * (piGC * (lam[:, 0] ** 2 - 1.0))
                       + 2.0 * (piGT * (lam[:, 2] ** 2 - 1.0)))
        l2 = lamf ** 2
        hh, vv = np.divmod(np.arange(16), 4)
        pi16 = kappa[:, None] * l2[:, BASEPAIRINGS] * pi[:, hh] * pi[:, vv]
        return Q4, pi, Q16, pi16, t

    def components(self, thetas):
        """Stacked site and pair mixture inputs for ``lib.Batch.eval_stacked``;
        equal to ``model.site_components`` / ``model.pair_components`` of
        ``params_of`` row by row."""
        Q4, pi, Q16, pi16, t = self.rate_matrices(thetas)
        n = Q4.shape[0]
        P4 = transition_matrices(Q4, t)[:, None]
        P16 = transition_matrices(Q16, t)[:, None]
        zeros = np.zeros((n, 1))
        return P4, pi[:, None, :], zeros, P16, pi16[:, None, :], zeros.copy()

    def eigen_components(self, thetas):
        """Inputs of ``lib.Batch.load_eigen``: the eigendecompositions of Q4 and Q16
        (the device builds every branch's P(t) from them), the frequencies, zero log
        weights (one component) and the branch lengths (n, nn)."""
        Q4, pi, Q16, pi16, t = self.rate_matrices(thetas)
        n = Q4.shape[0]
        V4, w4, Vi4 = reversible_eigen(Q4, pi)
        V16, w16, Vi16 = reversible_eigen(Q16, pi16)
        zeros = np.zeros((n, 1))
        return (V4[:, None], w4[:, None], Vi4[:, None], pi[:, None, :], zeros,
                V16[:, None], w16[:, None], Vi16[:, None], pi16[:, None, :], zeros.copy(), t)


def reversible_eigen(Q, pi):
    """Eigendecomposition Q = V diag(w) Vi of rate matrices that satisfy detailed
    balance with the positive vectors pi (GTR and MuSE do): the symmetric matrix
    A = D^1/2 Q D^-1/2 (D = diag(pi)) is diagonalised by LAPACK's symmetric solver,
    A = U diag(w) U^T, so V = D^-1/2 U and Vi = U^T D^1/2 are real and exact
    inverses.  The reference decomposes Q with the general solver
    (src/FARCE.jl:412-414); both give exp(Q t) = V diag(exp(w t)) Vi."""
    r = np.sqrt(np.asarray(pi, dtype=np.float64))
    A = Q * r[:, :, None] / r[:, None, :]
    A = 0.5 * (A + np.swapaxes(A, 1, 2))
    w, U = np.linalg.eigh(A)
    V = U / r[:, :, None]
    Vi = np.swapaxes(U, 1, 2) * r[:, None, :]
    return V, w, Vi


class Posterior(ParamMap):
    """Log posterior of ``nchains`` parameter vectors per call, on the device."""

    def __init__(self, dataset, base15, nchains: int, prec="fp32", maxbasepairdistance: int = 1000000):
        from . import lib
        from .hotpath import device_dataset
        super().__init__(base15, dataset.tree.branchlengths)
        self.ds = dataset
        self.nchains = int(nchains)
        self.prec = prec
        dd = device_dataset(dataset, maxbasepairdistance)
        self.batch = lib.Batch([dd] * self.nchains)
        self.evals = 0
        self.eval_seconds = 0.0

    def log_likelihood(self, thetas, prec=None) -> np.ndarray:
        th = np.atleast_2d(thetas)
        if th.shape[0] != self.nchains:
            raise ValueError(f"expected {self.nchains} vectors, got {th.shape[0]}")
        t0 = time.perf_counter()
        pr = prec or self.prec
        self.batch.load_eigen(*self.eigen_components(th), prec=pr)
        z = self.batch.eval_loaded(1.0, pr)
        self.eval_seconds += time.perf_counter() - t0
        self.evals += th.shape[0]
        return z

    def __call__(self, thetas, prec=None) -> np.ndarray:
        th = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
        ok = np.all(np.isfinite(th), axis=1) & np.all(np.abs(th) < 50.0, axis=1)
        safe = np.where(ok[:, None], th, 0.0)
        lp = self.log_likelihood(safe, prec) + self.log_prior(safe)
        lp = np.where(ok & np.isfinite(lp), lp, -np.inf)
        return lp

    def close(self):
        self.batch.close()


# ---------------------------------------------------------------------------
# Samplers (generic in the log density: a callable on (n, d) arrays)
# ---------------------------------------------------------------------------

@dataclass
class Trace:
    samples: np.ndarray          # (steps, nchains, d)
    logp: np.ndarray             # (steps, nchains)
    accepted: np.ndarray         # (steps, nchains) bool
    kind: np.ndarray             # (steps, nchains) 0 random walk, 1 learned
    seconds: float
    evals: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def acceptance(self) -> float:
        return float(self.accepted.mean())


def rw_chain(logp: Callable, theta, lp, steps: int, chol, scale: float, rng,
             adapt_until: int = 0, target_accept: float = 0.234,
             progress: Optional[Callable[[int], None]] = None) -> Trace:
    """Random-walk Metropolis: theta' = theta + scale * chol @ eps.

    With ``adapt_until > 0`` (burn-in only) the proposal adapts during the first
    ``adapt_until`` steps as the reference's adaptive Metropolis does
    (src/MCMC.jl:393-449): every 50 steps the covariance is re-estimated from
    the second half of the history of all chains, and one scale shared by the
    chains follows a Robbins-Monro rule towards ``target_accept`` acceptance.
    The final (chol, scale) are returned in ``extra``.
    """
    theta = np.array(theta, dtype=np.float64)
    lp = np.array(lp, dtype=np.float64)
    n, d = theta.shape
    chol = np.array(chol, dtype=np.float64)
    log_scale = float(np.log(scale))
    S = np.empty((steps, n, d))
    LP = np.empty((steps, n))
    A = np.zeros((steps, n), dtype=bool)
    t0 = time.perf_counter()
    for it in range(steps):
        eps = rng.standard_normal((n, d))
        prop = theta + np.exp(log_scale) * (eps @ chol.T)
        lpp = logp(prop)
        acc = np.log(rng.random(n)) < lpp - lp
        theta = np.where(acc[:, None], prop, theta)
        lp = np.where(acc, lpp, lp)
        S[it], LP[it], A[it] = theta, lp, acc
        if progress is not None and (it + 1) % 2000 == 0:
            progress(it + 1)
        if it < adapt_until:
            log_scale += (acc.mean() - target_accept) * 2.0 / np.sqrt(it + 10.0)
            if (it + 1) % 50 == 0 and it + 1 >= 4 * d:
                hist = S[(it + 1) // 2: it + 1].reshape(-1, d)
                C = np.cov(hist.T) + 1e-12 * np.eye(d)
                try:
                    chol = np.linalg.cholesky(C)
                except np.linalg.LinAlgError:
                    pass
    secs = time.perf_counter() - t0
    return Trace(S, LP, A, np.zeros((steps, n), dtype=np.int8), secs,
                 extra={"chol": chol, "scale": float(np.exp(log_scale))})


def eval_points(logp: Callable, pts, nb: int) -> np.ndarray:
    """Evaluate any number of points with a log density that takes exactly nb rows."""
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    out = np.empty(len(pts))
    for s in range(0, len(pts), nb):
        chunk = pts[s:s + nb]
        if len(chunk) < nb:
            chunk = np.concatenate([chunk, np.repeat(chunk[-1:], nb - len(chunk), axis=0)])
        out[s:s + nb] = logp(chunk)[:min(nb, len(pts) - s)]
    return out


def _fd_derivatives(logp, x, h, nb):
    """Value, gradient and Hessian by central differences with steps h (d,)."""
    d = len(x)
    E = np.diag(h)
    pts = [x]
    for i in range(d):
        pts += [x + E[i], x - E[i]]
    pairs = [(i, j) for i in range(d) for j in range(i + 1, d)]
    for i, j in pairs:
        pts += [x + E[i] + E[j], x + E[i] - E[j], x - E[i] + E[j], x - E[i] - E[j]]
    f = eval_points(logp, np.array(pts), nb)
    f0, fp, fm = f[0], f[1:1 + 2 * d:2], f[2:2 + 2 * d:2]
    g = (fp - fm) / (2 * h)
    H = np.diag((fp - 2 * f0 + fm) / (h * h))
    for k, (i, j) in enumerate(pairs):
        a, b, c, e = f[1 + 2 * d + 4 * k: 5 + 2 * d + 4 * k]
        H[i, j] = H[j, i] = (a - b - c + e) / (4 * h[i] * h[j])
    return f0, g, H


def laplace(logp: Callable, theta0, nb: int, iters: int = 12, log: Callable = print):
    """Posterior mode and the inverse of the negative Hessian there, by damped
    Newton steps on finite differences (every batch of evaluations holds nb
    points).  The difference step of each coordinate follows the current
    curvature (half a standard deviation), so that the differences stay well
    above the evaluation's rounding."""
    x = np.asarray(theta0, dtype=np.float64).copy()
    d = len(x)
    h = np.full(d, 1e-2)
    A = None
    for it in range(iters):
        f0, g, H = _fd_derivatives(logp, x, h, nb)
        w, U = np.linalg.eigh(-H)
        w = np.maximum(w, max(w.max(), 1.0) * 1e-8)
        A = (U * w) @ U.T
        # candidate steps for several damping levels, evaluated in one batch
        lams = [0.0, 0.1, 1.0, 10.0, 100.0]
        steps = [np.linalg.solve(A + lm * np.diag(np.diag(A)), g) for lm in lams]
        cand = eval_points(logp, np.array([x + s for s in steps]), nb)
        k = int(np.argmax(cand))
        dec = float(g @ steps[0])
        log(f"mode search {it}: log p {f0:.4f}, Newton decrement {dec:.3g}, best damping {lams[k]}")
        if cand[k] > f0:
            x = x + steps[k]
        h = np.clip(0.5 / np.sqrt(np.diag(A)), 1e-5, 0.1)
        if dec < 1e-3 or cand[k] <= f0:
            break
    f0, g, H = _fd_derivatives(logp, x, h, nb)
    w, U = np.linalg.eigh(-H)
    w = np.maximum(w, max(w.max(), 1.0) * 1e-8)
    cov = (U / w) @ U.T
    return x, cov


class Whitening:
    """z = chol(C)^-1 (theta - m) from a pool of samples."""

    def __init__(self, pool):
        pool = np.asarray(pool, dtype=np.float64).reshape(-1, np.shape(pool)[-1])
        self.m = pool.mean(axis=0)
        C = np.cov(pool.T) + 1e-12 * np.eye(pool.shape[1])
        self.L = np.linalg.cholesky(C)
        self.Li = np.linalg.inv(self.L)

    def z(self, theta):
        return (np.asarray(theta) - self.m) @ self.Li.T

    def theta(self, z):
        return self.m + np.asarray(z) @ self.L.T


class Mixture:
    """Mixture of multivariate Student-t components in whitened coordinates: an
    independence proposal that can jump between the modes the burn-in visited.
    Fitted by EM (Gaussian responsibilities, k-means++ starts) to burn-in
    samples; each component's scale is inflated so that the proposal is wider
    than the fitted modes.  logq and sample are exact for the mixture itself,
    so a Metropolis-Hastings move with it is exact whatever the fit quality."""

    def __init__(self, w, means, chols, nu: float = 7.0):
        self.w = np.asarray(w, dtype=np.float64)
        self.means = np.asarray(means, dtype=np.float64)
        self.chols = np.asarray(chols, dtype=np.float64)
        self.nu = float(nu)
        self.inv = np.linalg.inv(self.chols)
        self.logdet = np.sum(np.log(np.abs(np.diagonal(self.chols, axis1=1, axis2=2))), axis=1)

    @classmethod
    def fit(cls, z, k: int, rng, iters: int = 150, inflate: float = 1.3, nu: float = 7.0) -> "Mixture":
        z = np.asarray(z, dtype=np.float64)
        n, d = z.shape
        k = max(1, min(k, n // (4 * d)))
        # k-means++ starts
        cent = [z[rng.integers(n)]]
        for _ in range(1, k):
            d2 = np.min(((z[:, None, :] - np.array(cent)[None]) ** 2).sum(-1), axis=1)
            cent.append(z[rng.choice(n, p=d2 / d2.sum())])
        means = np.array(cent)
        covs = np.repeat(np.cov(z.T)[None] + 1e-3 * np.eye(d), k, axis=0)
        w = np.full(k, 1.0 / k)
        for _ in range(iters):
            ll = np.empty((n, k))
            for j in range(k):
                Lc = np.linalg.cholesky(covs[j])
                r = np.linalg.solve(Lc, (z - means[j]).T).T
                ll[:, j] = np.log(w[j]) - np.sum(np.log(np.diag(Lc))) - 0.5 * np.sum(r * r, axis=1)
            ll -= ll.max(axis=1, keepdims=True)
            g = np.exp(ll)
            g /= g.sum(axis=1, keepdims=True)
            nk = g.sum(axis=0) + 1e-9
            w = nk / n
            means = (g.T @ z) / nk[:, None]
            for j in range(k):
                r = z - means[j]
                covs[j] = (g[:, j, None] * r).T @ r / nk[j] + 1e-3 * np.eye(d)
        chols = np.array([np.linalg.cholesky(c) * inflate for c in covs])
        return cls(w, means, chols, nu)

    def flattened(self, a: float) -> "Mixture":
        """The same components with weights (1 - a) w + a / K: rarely visited modes
        keep enough proposal mass to be entered and left."""
        k = len(self.w)
        return Mixture((1.0 - a) * self.w + a / k, self.means, self.chols, self.nu)

    def logq(self, z) -> np.ndarray:
        z = np.atleast_2d(np.asarray(z, dtype=np.float64))
        d = z.shape[1]
        nu = self.nu
        c = math.lgamma(0.5 * (nu + d)) - math.lgamma(0.5 * nu) - 0.5 * d * math.log(nu * math.pi)
        out = np.empty((z.shape[0], len(self.w)))
        for j in range(len(self.w)):
            r = (z - self.means[j]) @ self.inv[j].T
            q = np.sum(r * r, axis=1)
            out[:, j] = np.log(self.w[j]) + c - self.logdet[j] - 0.5 * (nu + d) * np.log1p(q / nu)
        m = out.max(axis=1, keepdims=True)
        return (m + np.log(np.exp(out - m).sum(axis=1, keepdims=True)))[:, 0]

    def sample(self, n: int, rng) -> np.ndarray:
        d = self.means.shape[1]
        comp = rng.choice(len(self.w), size=n, p=self.w / self.w.sum())
        g = rng.standard_normal((n, d))
        scale = np.sqrt(self.nu / rng.chisquare(self.nu, n))
        y = np.einsum("nij,nj->ni", self.chols[comp], g) * scale[:, None]
        return self.means[comp] + y


def _gauss_logq(x, mean, ell):
    r = (x - mean) * np.exp(-ell)
    return -np.sum(ell + 0.5 * r * r, axis=-1)


def learned_chain(logp: Callable, theta, lp, steps: int, net, wh: Whitening, rng, rw_chol, rw_scale: float,
                  rw_prob: float = 0.1, kappa: float = 1.0, mix: Optional[Mixture] = None,
                  ind_prob: float = 0.0, progress: Optional[Callable[[int], None]] = None,
                  mix_wh: Optional[Whitening] = None) -> Trace:
    """MH with the learned proposal z' ~ N(z + mu(z), diag (kappa exp(ell(z)))^2);
    each chain takes a random-walk step instead with probability ``rw_prob``, and
    with probability ``ind_prob`` an independence step from the fitted mixture
    ``mix`` (every kernel leaves the target invariant, so the random mixture of
    kernels does too).  With ``net`` None there is no learned move: the chain
    takes the independence step with probability ``ind_prob`` and the random walk
    otherwise (the method without the network).  The mixture lives in the
    coordinates of ``mix_wh`` when given (else ``wh``, the network's).
    Trace.kind: 0 random walk, 1 learned, 2 independence."""
    mwh = mix_wh if mix_wh is not None else wh
    theta = np.array(theta, dtype=np.float64)
    lp = np.array(lp, dtype=np.float64)
    n, d = theta.shape
    S = np.empty((steps, n, d))
    LP = np.empty((steps, n))
    A = np.zeros((steps, n), dtype=bool)
    K = np.zeros((steps, n), dtype=np.int8)
    t0 = time.perf_counter()
    for it in range(steps):
        u = rng.random(n)
        use_rw = u < rw_prob if net is not None else u >= ind_prob
        use_ind = (u >= rw_prob) & (u < rw_prob + ind_prob) if net is not None else ~use_rw
        if mix is None:
            use_ind = np.zeros(n, dtype=bool)
            use_rw = use_rw if net is not None else np.ones(n, dtype=bool)
        z = wh.z(theta)
        if net is not None:
            out = net.forward(z.astype(np.float32)).astype(np.float64)
            mu, ell = out[:, :d], np.clip(out[:, d:], -8.0, 4.0) + np.log(kappa)
            zp = z + mu + np.exp(ell) * rng.standard_normal((n, d))
            outp = net.forward(zp.astype(np.float32)).astype(np.float64)
            mup, ellp = outp[:, :d], np.clip(outp[:, d:], -8.0, 4.0) + np.log(kappa)
            log_ratio = _gauss_logq(z, zp + mup, ellp) - _gauss_logq(zp, z + mu, ell)
        else:
            zp = z.copy()
            log_ratio = np.zeros(n)
        prop = wh.theta(zp)
        if mix is not None and use_ind.any():
            zi = mix.sample(n, rng)
            li = mix.logq(mwh.z(theta)) - mix.logq(zi)   # q(z) / q(z'): the whitening's Jacobian cancels
            prop = np.where(use_ind[:, None], mwh.theta(zi), prop)
            log_ratio = np.where(use_ind, li, log_ratio)
        rwp = theta + rw_scale * (rng.standard_normal((n, d)) @ np.asarray(rw_chol).T)
        prop = np.where(use_rw[:, None], rwp, prop)
        log_ratio = np.where(use_rw, 0.0, log_ratio)
        lpp = logp(prop)
        acc = np.log(rng.random(n)) < lpp - lp + log_ratio
        theta = np.where(acc[:, None], prop, theta)
        lp = np.where(acc, lpp, lp)
        kind = np.where(use_rw, 0, np.where(use_ind, 2, 1)).astype(np.int8)
        S[it], LP[it], A[it], K[it] = theta, lp, acc, kind
        if progress is not None and (it + 1) % 2000 == 0:
            progress(it + 1)
    return Trace(S, LP, A, K, time.perf_counter() - t0)


def _pairs(z, rng, kmin, kmax, count):
    T, n, _ = z.shape
    k = rng.integers(kmin, kmax + 1, count)
    i = (rng.random(count) * (T - k)).astype(np.int64)
    c = rng.integers(0, n, count)
    return z[i, c].astype(np.float32), z[i + k, c].astype(np.float32)


def train_proposal(net, samples, wh: Whitening, rng, tau: float, train_steps: int = 1500,
                   batch: int = 512, lr: float = 1e-3, check_every: int = 50, jitter: float = 0.05,
                   weight_decay: float = 1e-2, noise: float = 0.0) -> dict:
    """Adam on the device over pairs (z_i, z_{i+k}), k ~ U[tau, 2 tau], of the
    samples (steps, nchains, d).  For a reversible chain the k-step transition
    density satisfies detailed balance, so a proposal fitted to it is accepted
    often while moving as far as k steps.  A quarter of the chains is held out:
    every ``check_every`` steps the loss on fixed held-out pairs is measured and
    the network ends with the weights of the best check (early stopping: the
    burn-in's autocorrelated samples are few independent points, and the
    Gaussian log-likelihood rewards memorising them with a shrinking scale).
    Inputs are jittered by N(0, jitter^2) and the weights decay (AdamW) for the
    same reason."""
    z = wh.z(samples)                                 # (steps, nchains, d)
    T, n, d = z.shape
    kmin = int(max(1, min(round(tau), T // 4)))
    kmax = int(max(kmin + 1, min(2 * kmin, T // 2)))
    nval = max(1, n // 4) if n > 1 else 0
    ztr, zval = (z[:, :n - nval], z[:, n - nval:]) if nval else (z, z)
    if noise > 0:
        # denoising pairs (z + noise eps, z): the loss-optimal mean shift is
        # noise^2 times the score of the noise-smoothed posterior (Tweedie), so
        # z + mu(z) + sqrt(2) exp(ell) eps is a Langevin step with a learned drift
        def pairs(zz, g, count):
            i = g.integers(0, zz.shape[0], count)
            c = g.integers(0, zz.shape[1], count)
            t = zz[i, c]
            return (t + noise * g.standard_normal(t.shape)).astype(np.float32), t.astype(np.float32)
    else:
        def pairs(zz, g, count):
            return _pairs(zz, g, kmin, kmax, count)
    vx, vt = pairs(zval, np.random.default_rng(12345), 2048)
    losses = np.empty(train_steps)
    best = (np.inf, net.get_params(), 0)
    val_hist = []
    for it in range(train_steps):
        x, t = pairs(ztr, rng, batch)
        if jitter > 0 and noise == 0:
            x = (x + jitter * rng.standard_normal(x.shape)).astype(np.float32)
        rate = lr * (0.1 + 0.9 * 0.5 * (1.0 + np.cos(np.pi * it / train_steps)))
        losses[it] = net.train_step(x, t, lr=rate, weight_decay=weight_decay)
        if (it + 1) % check_every == 0 or it + 1 == train_steps:
            vl = net.loss_grad(vx, vt)[0]
            val_hist.append((it + 1, vl))
            if vl < best[0]:
                best = (vl, net.get_params(), it + 1)
    net.set_params(best[1])
    return {"train_loss": losses, "val_loss": val_hist, "best_step": best[2], "best_val_loss": best[0],
            "lag": (kmin, kmax)}


# ---------------------------------------------------------------------------
# Diagnostics
# ---------------------------------------------------------------------------

def _autocov(x):
    """Autocovariance of each column of x (T, m) by FFT, lags 0..T-1."""
    T = x.shape[0]
    xc = x - x.mean(axis=0)
    nfft = 1 << (2 * T - 1).bit_length()
    f = np.fft.rfft(xc, n=nfft, axis=0)
    ac = np.fft.irfft(f * np.conj(f), n=nfft, axis=0)[:T]
    return ac / T


def ess(samples) -> np.ndarray:
    """Multi-chain effective sample size of each parameter (samples: (T, nchains, d)),
    with Geyer's initial monotone sequence on the chain-averaged autocorrelation
    (Gelman et al., Bayesian Data Analysis 3rd ed., section 11.5)."""
    x = np.asarray(samples, dtype=np.float64)
    T, m, d = x.shape
    out = np.empty(d)
    for j in range(d):
        y = x[:, :, j]
        ac = _autocov(y)                               # (T, m)
        W = np.mean(ac[0] * T / (T - 1))
        means = y.mean(axis=0)
        B = T * np.var(means, ddof=1) if m > 1 else 0.0
        var_plus = (T - 1) / T * W + B / T
        if var_plus <= 0:
            out[j] = float(T * m)
            continue
        rho = 1.0 - (W - np.mean(ac, axis=1)) / var_plus
        rho[0] = 1.0
        # Geyer: sums of adjacent pairs, positive and monotone
        s = 0.0
        prev = np.inf
        for t in range(0, T - 1, 2):
            p = rho[t] + rho[t + 1]
            if p <= 0:
                break
            p = min(p, prev)
            prev = p
            s += p
        tau = -1.0 + 2.0 * s
        out[j] = T * m / max(tau, 1.0 / np.log10(max(T * m, 10)))
    return out


def split_rhat(samples) -> np.ndarray:
    """Split potential scale reduction factor of each parameter (samples:
    (T, nchains, d)); Gelman et al., Bayesian Data Analysis 3rd ed., (11.4)."""
    x = np.asarray(samples, dtype=np.float64)
    h = x.shape[0] // 2
    y = np.concatenate([x[:h], x[h:2 * h]], axis=1)
    W = y.var(axis=0, ddof=1).mean(axis=0)
    Bv = h * y.mean(axis=0).var(axis=0, ddof=1)
    return np.sqrt(((h - 1) / h * W + Bv / h) / W)


def integrated_time(samples) -> float:
    """Largest integrated autocorrelation time over the parameters."""
    x = np.asarray(samples)
    return float(np.max(x.shape[0] * x.shape[1] / ess(x)))


def mcse(samples, ess_values) -> np.ndarray:
    """Monte Carlo standard error of each posterior mean (samples: (T, nchains, d)):
    the larger of sd / sqrt(ESS) and the spread of the chain means over
    sqrt(nchains).  The second treats the chains as independent replicates and
    stays honest when chains are stuck in different modes (split R-hat well
    above 1), where autocorrelation-based ESS estimates are too optimistic."""
    x = np.asarray(samples, dtype=np.float64)
    T, m, d = x.shape
    se_ess = x.reshape(-1, d).std(axis=0) / np.sqrt(np.asarray(ess_values, dtype=np.float64))
    if m < 2:
        return se_ess
    se_rep = x.mean(axis=0).std(axis=0, ddof=1) / np.sqrt(m)
    return np.maximum(se_ess, se_rep)


def compare_means(a, ess_a, b, ess_b, conservative: bool = True):
    """z-scores of the difference of posterior means of two sample sets (each
    (T, nchains, d)); with conservative=False the standard errors are
    sd / sqrt(ESS) only."""
    fa = np.asarray(a).reshape(-1, np.shape(a)[-1])
    fb = np.asarray(b).reshape(-1, np.shape(b)[-1])
    if conservative:
        se = np.sqrt(mcse(a, ess_a) ** 2 + mcse(b, ess_b) ** 2)
    else:
        se = np.sqrt(fa.var(axis=0) / ess_a + fb.var(axis=0) / ess_b)
    return (fa.mean(axis=0) - fb.mean(axis=0)) / se


def _esjd(samples, wh: Whitening) -> float:
    """Smallest over coordinates of the mean squared jump per step, in whitened
    coordinates (rejections count as zero jumps)."""
    z = wh.z(samples)
    return float(np.min(np.mean(np.diff(z, axis=0) ** 2, axis=(0, 1))))


def run_experiment(logp: Callable, theta0, net, rng, burn: int = 2000, steps: int = 1250,
                   long_steps: int = 4000, train_steps: int = 1500, rw_prob: float = 0.1,
                   noises=(0.8, 1.1), kappas=(1.3, 1.6), log: Callable = print,
                   laplace_init: bool = True, mode_logp: Optional[Callable] = None,
                   ind_prob: float = 0.15, mix_components: int = 4, mix_inflate: float = 1.3,
                   mix_nu: float = 7.0, mix_flatten: float = 0.5,
                   share_candidates=(0.85, 0.3, 0.0), share_steps: int = 500) -> dict:
    """Learned proposal against the adaptive random walk, each with a burn-in of
    ``burn`` steps from the same start, then ``steps`` production steps each and
    a long random-walk chain as the yardstick for the posterior means.

    Learned method: adaptive random walk for burn/2 steps; then for each
    denoising noise level the network is trained on every burn-in sample so far
    and the learned proposal runs once per scale factor kappa, the settings
    being scored by their expected squared jump distance; a last training on
    all burn-in samples at the best noise level fixes the network, and the best
    kappa is kept for production.  A mixture of Student-t components fitted to
    the whitened random-walk burn-in (and refitted after an exploration segment)
    gives independence moves (probability ``ind_prob``) that jump between the
    modes the burn-in visited; it is shared with the method without the network
    (random walk plus the same independence moves), so the two differ only in
    the network's moves, which take the share of the random walk (one of
    ``share_candidates``, 0 included) whose burn-in segment of ``share_steps``
    steps, independence moves included, gave the largest minimum effective
    sample size.  Random-walk method: adaptive random walk for
    all ``burn`` steps.  Production kernels are fixed (no adaptation), so each
    production chain is an exact MH chain.
    """
    theta0 = np.asarray(theta0, dtype=np.float64)
    n, d = theta0.shape
    t_start = time.perf_counter()
    if laplace_init:
        # start as the reference does after its maximum-likelihood step: at the
        # mode, here with the curvature there as the first proposal covariance
        mode, cov = laplace(mode_logp or logp, theta0[0], n, log=log)
        chol0 = np.linalg.cholesky(cov)
        theta0 = mode + 0.5 * rng.standard_normal((n, d)) @ chol0.T
        log(f"mode and curvature: {time.perf_counter() - t_start:.1f} s")
    else:
        chol0 = np.eye(d) * 0.01
    lp0 = logp(theta0)
    burn_rw = burn // 2
    per_trial = max(10, (burn - burn_rw) // (len(noises) * len(kappas)))
    tb = rw_chain(logp, theta0, lp0, burn_rw, chol0, 2.38 / np.sqrt(d), rng, adapt_until=burn_rw)
    log(f"burn-in (random walk): {burn_rw} steps x {n} chains, acceptance {tb.acceptance:.3f}, "
        f"{tb.seconds:.1f} s")
    rw_chol, rw_scale = tb.extra["chol"], tb.extra["scale"]
    segs = [tb.samples[burn_rw // 2:]]
    theta, lp = tb.samples[-1], tb.logp[-1]
    w_init = net.get_params()
    train_s = 0.0
    trials = []
    for nz in noises:
        pool = np.concatenate(segs, axis=0)
        wh = Whitening(pool)
        net.set_params(w_init)
        t0 = time.perf_counter()
        tinfo = train_proposal(net, pool, wh, rng, 1.0, train_steps=train_steps, noise=nz)
        train_s += time.perf_counter() - t0
        for k in kappas:
            ta = learned_chain(logp, theta, lp, per_trial, net, wh, rng, rw_chol, rw_scale, rw_prob=rw_prob,
                               kappa=k)
            acc = float(ta.accepted[ta.kind == 1].mean()) if (ta.kind == 1).any() else 0.0
            score = _esjd(np.concatenate([theta[None], ta.samples]), wh)
            trials.append({"noise": nz, "kappa": k, "acceptance_learned_moves": acc, "esjd_min": score,
                           "held_out_loss": tinfo["best_val_loss"]})
            log(f"burn-in (learned) noise {nz} kappa {k}: {per_trial} steps, acceptance {acc:.3f}, "
                f"min ESJD {score:.4f}")
            segs.append(ta.samples)
            theta, lp = ta.samples[-1], ta.logp[-1]
    best = max(trials, key=lambda t: t["esjd_min"])
    pool = np.concatenate(segs, axis=0)
    wh = Whitening(pool)
    net.set_params(w_init)
    t0 = time.perf_counter()
    tinfo = train_proposal(net, pool, wh, rng, 1.0, train_steps=train_steps, noise=best["noise"])
    train_s += time.perf_counter() - t0
    log(f"final training: noise {best['noise']}, kappa {best['kappa']}, pool {pool.shape[0]} x {n}, "
        f"held-out loss {tinfo['best_val_loss']:.3f} at step {tinfo['best_step']}")
    learned_burn_s = time.perf_counter() - t_start
    # random-walk method: the rest of its own burn-in
    trb = rw_chain(logp, tb.samples[-1], tb.logp[-1], burn - burn_rw, rw_chol, rw_scale, rng,
                   adapt_until=burn - burn_rw)
    rw_chol2, rw_scale2 = trb.extra["chol"], trb.extra["scale"]
    rw_burn_s = tb.seconds + trb.seconds
    log(f"burn-in (random walk, rest): {burn - burn_rw} steps, acceptance {trb.acceptance:.3f}")
    # The independence mixture, shared by the learned method and the method
    # without the network (so the two differ only in the network's moves): fitted
    # on the random-walk burn-in, used in an exploration segment with the
    # independence moves (modes the random walk barely visited get samples), and
    # refitted on both.  The method without the network then runs the
    # independence moves with probability ind_prob and the random walk otherwise.
    mix_res = None
    mixm = whm = None
    if ind_prob > 0.0:
        t_m = time.perf_counter()
        pre = burn - burn_rw - per_trial
        tm0 = rw_chain(logp, tb.samples[-1], tb.logp[-1], pre, rw_chol, rw_scale, rng, adapt_until=pre)
        m_chol, m_scale = tm0.extra["chol"], tm0.extra["scale"]
        mpool = np.concatenate([segs[0], tm0.samples[pre // 2:]], axis=0)
        whm = Whitening(mpool)
        mixm = Mixture.fit(whm.z(mpool).reshape(-1, d), mix_components, rng, inflate=mix_inflate,
                           nu=mix_nu).flattened(mix_flatten)
        tme = learned_chain(logp, tm0.samples[-1], tm0.logp[-1], per_trial, None, whm, rng, m_chol, m_scale,
                            mix=mixm, ind_prob=max(ind_prob, 0.5))
        mixm = Mixture.fit(whm.z(np.concatenate([mpool, tme.samples], axis=0)).reshape(-1, d), mix_components,
                           rng, inflate=mix_inflate, nu=mix_nu).flattened(mix_flatten)
        mix_burn_s = tb.seconds + time.perf_counter() - t_m
        tmx = learned_chain(logp, tme.samples[-1], tme.logp[-1], steps, None, whm, rng, m_chol, m_scale,
                            mix=mixm, ind_prob=ind_prob,
                            progress=lambda i: log(f"mixture chain (no network): {i} of {steps} steps"))
        mind_acc = float(tmx.accepted[tmx.kind == 2].mean()) if (tmx.kind == 2).any() else 0.0
        log(f"mixture chain (no network): {steps} steps, acceptance {tmx.acceptance:.3f}, {mind_acc:.3f} "
            f"(independence moves), {tmx.seconds:.1f} s")
        mix_res = (tmx, mix_burn_s, mind_acc)

    # Learned method: the same independence moves plus the network's moves, which
    # replace most of the random walk when their burn-in jump distance beat the
    # random walk's and a small share of it otherwise.
    rw_esjd = _esjd(np.concatenate([tb.samples[-1][None], trb.samples]), wh)
    share_trials = []
    if mixm is not None:
        # The network's share of the moves that are not independence moves is
        # chosen in burn-in, with the independence moves present: a segment at
        # each candidate share (0 = the method without the network), scored by
        # its smallest effective sample size over the parameters (the jump
        # distance favoured the network's larger local moves even where they
        # mixed the slowest parameter worse).
        learned_burn_s += mix_res[1] - tb.seconds      # its share of the shared mixture's burn-in
        t_s = time.perf_counter()
        for share in share_candidates:
            p = (1.0 - ind_prob) * share
            ts = learned_chain(logp, theta, lp, share_steps, net if p > 0.0 else None, wh, rng, rw_chol,
                               rw_scale, rw_prob=1.0 - ind_prob - p, kappa=best["kappa"], mix=mixm,
                               ind_prob=ind_prob, mix_wh=whm)
            score = float(ess(ts.samples).min())
            share_trials.append({"share": share, "learned_prob": p, "min_ess": score,
                                 "esjd_min": _esjd(np.concatenate([theta[None], ts.samples]), wh),
                                 "acceptance": ts.acceptance})
            log(f"burn-in share trial {share}: {share_steps} steps, acceptance {ts.acceptance:.3f}, "
                f"min ESS {score:.1f}")
            theta, lp = ts.samples[-1], ts.logp[-1]
        learned_burn_s += time.perf_counter() - t_s
        pick = max(share_trials, key=lambda t: t["min_ess"])
        p_learned = pick["learned_prob"]
        l_rw, l_ind = 1.0 - ind_prob - p_learned, ind_prob
    else:
        p_learned, l_rw, l_ind = 1.0 - rw_prob, rw_prob, 0.0
    log(f"learned move share {p_learned:.2f} (burn-in min ESJD: learned {best['esjd_min']:.4f}, "
        f"random walk {rw_esjd:.4f})")
    tl = learned_chain(logp, theta, lp, steps, net, wh, rng, rw_chol, rw_scale, rw_prob=l_rw,
                       kappa=best["kappa"], mix=mixm, ind_prob=l_ind, mix_wh=whm,
                       progress=lambda i: log(f"learned chain: {i} of {steps} steps"))
    learned_acc = float(tl.accepted[tl.kind == 1].mean()) if (tl.kind == 1).any() else 0.0
    ind_acc = float(tl.accepted[tl.kind == 2].mean()) if (tl.kind == 2).any() else 0.0
    log(f"learned chain: {steps} steps, acceptance {learned_acc:.3f} (learned moves), {ind_acc:.3f} "
        f"(independence moves), {tl.seconds:.1f} s")
    tr = rw_chain(logp, trb.samples[-1], trb.logp[-1], steps, rw_chol2, rw_scale2, rng,
                  progress=lambda i: log(f"random-walk chain: {i} of {steps} steps"))
    log(f"random-walk chain: {steps} steps, acceptance {tr.acceptance:.3f}, {tr.seconds:.1f} s")
    tlong = rw_chain(logp, trb.samples[-1], trb.logp[-1], long_steps, rw_chol2, rw_scale2, rng,
                     progress=lambda i: log(f"long random-walk chain: {i} of {long_steps} steps"))
    log(f"long random-walk chain: {long_steps} steps, {tlong.seconds:.1f} s")
    e_l, e_r, e_long = ess(tl.samples), ess(tr.samples), ess(tlong.samples)
    z = compare_means(tl.samples, e_l, tlong.samples, e_long)
    z_rw = compare_means(tr.samples, e_r, tlong.samples, e_long)
    z_ess = compare_means(tl.samples, e_l, tlong.samples, e_long, conservative=False)
    res = {
        "nchains": n, "burn": burn, "steps": steps, "long_steps": long_steps,
        "train_seconds": train_s, "learned_burn_seconds": learned_burn_s, "trials": trials,
        "chosen": {"noise": best["noise"], "kappa": best["kappa"], "learned_share": p_learned,
                   "burn_esjd_learned": best["esjd_min"], "burn_esjd_random_walk": rw_esjd,
                   "share_trials": share_trials},
        "learned": {"seconds": tl.seconds, "acceptance_learned_moves": learned_acc,
                    "acceptance_independence_moves": ind_acc, "independence_prob": l_ind,
                    "acceptance": tl.acceptance, "ess": e_l.tolist(), "min_ess": float(e_l.min()),
                    "min_ess_per_s": float(e_l.min() / tl.seconds),
                    "median_ess_per_s": float(np.median(e_l) / tl.seconds)},
        "random_walk": {"seconds": tr.seconds, "acceptance": tr.acceptance, "ess": e_r.tolist(),
                        "min_ess": float(e_r.min()), "min_ess_per_s": float(e_r.min() / tr.seconds),
                        "median_ess_per_s": float(np.median(e_r) / tr.seconds)},
        "long_chain": {"seconds": tlong.seconds, "ess": e_long.tolist(),
                       "mean": tlong.samples.reshape(-1, d).mean(axis=0).tolist(),
                       "sd": tlong.samples.reshape(-1, d).std(axis=0).tolist()},
        "learned_mean": tl.samples.reshape(-1, d).mean(axis=0).tolist(),
        "split_rhat": {"learned": split_rhat(tl.samples).tolist(), "random_walk": split_rhat(tr.samples).tolist(),
                       "long": split_rhat(tlong.samples).tolist()},
        "mean_z_learned_vs_long": z.tolist(), "max_abs_z_learned": float(np.max(np.abs(z))),
        "mean_z_learned_vs_long_ess_only": z_ess.tolist(),
        "mcse": {"learned": mcse(tl.samples, e_l).tolist(), "long": mcse(tlong.samples, e_long).tolist()},
        "mean_z_rw_vs_long": z_rw.tolist(), "max_abs_z_rw": float(np.max(np.abs(z_rw))),
    }
    res["ess_per_s_gain_min"] = res["learned"]["min_ess_per_s"] / res["random_walk"]["min_ess_per_s"]
    res["ess_per_s_gain_median"] = res["learned"]["median_ess_per_s"] / res["random_walk"]["median_ess_per_s"]
    # ESS per second counting each method's own burn-in (and, for the learned one,
    # its training) as well as its production chain
    res["learned"]["burn_seconds"] = learned_burn_s
    res["learned"]["min_ess_per_s_incl_burn"] = float(e_l.min() / (tl.seconds + learned_burn_s))
    res["random_walk"]["burn_seconds"] = rw_burn_s
    res["random_walk"]["min_ess_per_s_incl_burn"] = float(e_r.min() / (tr.seconds + rw_burn_s))
    res["ess_per_s_gain_min_incl_burn"] = (res["learned"]["min_ess_per_s_incl_burn"]
                                           / res["random_walk"]["min_ess_per_s_incl_burn"])
    traces = {"learned": tl, "random_walk": tr, "long": tlong, "burn": tb}
    if mix_res is not None:
        tmx, mix_burn_s, mind_acc = mix_res
        e_m = ess(tmx.samples)
        z_m = compare_means(tmx.samples, e_m, tlong.samples, e_long)
        res["mixture_no_network"] = {
            "seconds": tmx.seconds, "burn_seconds": mix_burn_s, "acceptance": tmx.acceptance,
            "acceptance_independence_moves": mind_acc, "ess": e_m.tolist(), "min_ess": float(e_m.min()),
            "min_ess_per_s": float(e_m.min() / tmx.seconds), "median_ess_per_s": float(np.median(e_m) / tmx.seconds),
            "min_ess_per_s_incl_burn": float(e_m.min() / (tmx.seconds + mix_burn_s)),
            "mean": tmx.samples.reshape(-1, d).mean(axis=0).tolist(), "mean_z_vs_long": z_m.tolist(),
            "max_abs_z": float(np.max(np.abs(z_m))), "split_rhat": split_rhat(tmx.samples).tolist()}
        mr = res["mixture_no_network"]
        res["network_gain_over_mixture_min"] = res["learned"]["min_ess_per_s"] / mr["min_ess_per_s"]
        res["network_gain_over_mixture_median"] = res["learned"]["median_ess_per_s"] / mr["median_ess_per_s"]
        res["mixture_gain_over_rw_min"] = mr["min_ess_per_s"] / res["random_walk"]["min_ess_per_s"]
        traces["mixture"] = tmx
    res["traces"] = traces
    return res
