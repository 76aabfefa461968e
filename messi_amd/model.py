"""Substitution models and transition matrices (host side, FP64, NumPy).

This is the model layer that sits above the device kernels: it turns a
parameter vector into the per-branch transition matrices P(t) that the
pruning kernels consume.  Every function restates the reference's own
definition:

* ``gtr``                 -- src/RateMatrices.jl:2-37 (unnormalised GTR, Q[i,j] = q_ij * pi_j)
* ``calculatemusefreqs``  -- src/RateMatrices.jl:40-66
* ``muse``                -- src/RateMatrices.jl:69-148 (16-state dinucleotide MuSE model)
* ``gettransitionmatriceslist`` -- src/FARCE.jl:389-433 (eigen, clamp <= 0 to 1e-50, t == 0 -> I,
  index 0 (the root) is an unused zero matrix)
* ``ModelParameters`` / ``getparams`` -- src/FARCE.jl:460-657 (15-vector layout, lambda categories
  from discretizegamma3, the single-category special case at 645-652)
* ``discretizegamma``     -- src/FARCE.jl:1531-1540
* ``discretizegamma3``    -- src/FARCE.jl:1566-1583
* ``computeprior``        -- src/FARCE.jl:2401-2436
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# Pair class of dinucleotide state (h-1)*4+v (1-based in the reference): 1 = GC/CG, 2 = AU/UA, 3 = GU/UG.
# src/RateMatrices.jl:49 and :71.
BASEPAIRINGS = np.array([0, 0, 0, 2, 0, 0, 1, 0, 0, 1, 0, 3, 2, 0, 3, 0], dtype=np.int64)


def gtr(q1, q2, q3, q4, q5, q6, piv=(0.25, 0.25, 0.25, 0.25)) -> np.ndarray:
    """Unnormalised GTR rate matrix, src/RateMatrices.jl:2-37."""
    S = np.zeros((4, 4))
    S[0, 1] = S[1, 0] = q1
    S[0, 2] = S[2, 0] = q2
    S[0, 3] = S[3, 0] = q3
    S[1, 2] = S[2, 1] = q4
    S[1, 3] = S[3, 1] = q5
    S[2, 3] = S[3, 2] = q6
    S = S * np.asarray(piv, dtype=np.float64)[None, :]
    Q = S.copy()
    np.fill_diagonal(Q, 0.0)
    Q[np.diag_indices(4)] = -S.sum(axis=1) + np.diag(S)
    return Q


def calculatemusefreqs(obsfreqs, lambdaGC, lambdaAT, lambdaGT) -> np.ndarray:
    """Stationary distribution of the MuSE model, src/RateMatrices.jl:40-66."""
    f = np.asarray(obsfreqs, dtype=np.float64)
    piGC = f[1] * f[2]
    piAT = f[0] * f[3]
    piGT = f[2] * f[3]
    kappa = 1.0 / (1.0 + 2.0 * (piAT * (lambdaAT * lambdaAT - 1.0))
                   + 2.0 * (piGC * (lambdaGC * lambdaGC - 1.0))
                   + 2.0 * (piGT * (lambdaGT * lambdaGT - 1.0)))
    lam2 = {0: 1.0, 1: lambdaGC * lambdaGC, 2: lambdaAT * lambdaAT, 3: lambdaGT * lambdaGT}
    out = np.zeros(16)
    for h in range(4):
        for v in range(4):
            idx = h * 4 + v
            out[idx] = kappa * lam2[int(BASEPAIRINGS[idx])] * f[h] * f[v]
    return out


def muse(lambdaGC, lambdaAT, lambdaGT, q1, q2, q3, q4, q5, q6,
         obsfreqs=(0.25, 0.25, 0.25, 0.25), siteRate1=1.0, siteRate2=1.0) -> np.ndarray:
    """16-state MuSE rate matrix, src/RateMatrices.jl:69-148.

    Only single-nucleotide changes have non-zero rate.  State index is
    (nucleotide of the first column)*4 + (nucleotide of the second column);
    a change of the first column is scaled by ``siteRate1`` and of the second
    by ``siteRate2``.  A move into a paired class is multiplied by its lambda,
    a move out of one divided by it.
    """
    f = np.asarray(obsfreqs, dtype=np.float64)
    gq = gtr(q1, q2, q3, q4, q5, q6, np.ones(4))
    lam = {1: lambdaGC, 2: lambdaAT, 3: lambdaGT}
    Q = np.zeros((16, 16))
    for h in range(16):
        for v in range(h + 1, 16):
            if h // 4 == v // 4:          # same first nucleotide: second column changes
                frm, to = h % 4, v % 4
                site = siteRate2
            elif h % 4 == v % 4:          # same second nucleotide: first column changes
                frm, to = h // 4, v // 4
                site = siteRate1
            else:
                continue
            m1 = 1.0
            m2 = 1.0
            cv, ch = int(BASEPAIRINGS[v]), int(BASEPAIRINGS[h])
            if cv:
                m1 *= lam[cv]
                m2 /= lam[cv]
            if ch:
                m1 /= lam[ch]
                m2 *= lam[ch]
            m1 *= gq[frm, to]
            m2 *= gq[frm, to]
            m1 *= site
            m2 *= site
            Q[h, v] = m1 * f[to]
            Q[v, h] = m2 * f[frm]
    np.fill_diagonal(Q, 0.0)
    Q[np.diag_indices(16)] = -Q.sum(axis=1)
    return Q


def gettransitionmatriceslist(branchlengths: Sequence[float], Q: np.ndarray) -> np.ndarray:
    """P(t) for every node, src/FARCE.jl:410-433 (with ``absmat`` 389-408).

    Returns an array (nnodes, dim, dim) with P[n][a, d] = P(a -> d) over the
    branch above node n.  Node 0 (the root) gets a zero matrix, as in the
    reference (its entry is never used).  Zero-length branches give the
    identity; otherwise entries <= 0 are clamped to 1e-50.
    """
    bl = np.asarray(branchlengths, dtype=np.float64)
    dim = Q.shape[0]
    w, V = np.linalg.eig(Q)
    Vi = np.linalg.inv(V)
    out = np.zeros((len(bl), dim, dim))
    if len(bl) > 1:
        t = bl[1:]
        E = np.exp(w[None, :] * t[:, None])                     # (n-1, dim)
        P = np.real(np.einsum("ij,nj,jk->nik", V, E, Vi))
        P = np.where(P <= 0.0, 1e-50, P)
        zero = t == 0.0
        if zero.any():
            P[zero] = np.eye(dim)
        out[1:] = P
    return out


def discretizegamma(shape: float, scale: float, numcategories: int) -> np.ndarray:
    """Equal-weight gamma rate categories (category medians), src/FARCE.jl:1531-1540."""
    if numcategories == 1:
        return np.array([1.0])
    from scipy.stats import gamma
    catwidth = 1.0 / numcategories
    vs = [catwidth / 2.0 + i * catwidth for i in range(numcategories)]
    return np.array([gamma.ppf(v, a=shape, scale=scale) for v in vs])


def discretizegamma3(lambdazeroweight: float, meanGC: float, meanAT: float, meanGT: float, gammavar: float,
                     numcategories: int):
    """Lambda categories, src/FARCE.jl:1566-1583: a rate-0 category of weight
    lambdazeroweight, then numcategories equal-weight gamma categories
    (discretizegamma(gammavar, 1/gammavar, n)), each lambda's rates scaled by its
    mean minus one.  Returns (weights, ratesGC, weights, ratesAT, weights, ratesGT)."""
    rates = np.concatenate([[0.0], discretizegamma(gammavar, 1.0 / gammavar, numcategories)
                            if numcategories > 0 else np.zeros(0)])
    w = np.full(numcategories + 1, (1.0 - lambdazeroweight) / numcategories if numcategories > 0 else np.inf)
    w[0] = lambdazeroweight
    return w, rates * (meanGC - 1.0), w.copy(), rates * (meanAT - 1.0), w.copy(), rates * (meanGT - 1.0)


@dataclass
class MuseSpecificParameters:
    """One lambda category, src/FARCE.jl:662-674."""
    lambdacat: int
    prob: float
    lambdarate: float

    @property
    def logprob(self) -> float:
        return float(np.log(self.prob))


@dataclass
class ModelParameters:
    """The fields of src/FARCE.jl:460-501 that the likelihood path reads."""
    freqs: np.ndarray
    lambdaGC: float
    lambdaAT: float
    lambdaGT: float
    q: np.ndarray                      # q1..q6
    branchlengths: np.ndarray          # per node in pre-order, root first (root entry unused)
    lambdazeroweight: float = 0.5
    lambdaratesGC: np.ndarray = field(default_factory=lambda: np.zeros(1))
    lambdaratesAT: np.ndarray = field(default_factory=lambda: np.zeros(1))
    lambdaratesGT: np.ndarray = field(default_factory=lambda: np.zeros(1))
    lambdaweightsGC: np.ndarray = field(default_factory=lambda: np.ones(1))
    siteRates: np.ndarray = field(default_factory=lambda: np.ones(1))
    siteWeights: np.ndarray = field(default_factory=lambda: np.ones(1))
    lambdaGammaShapeGC: float = 1.0
    lambdaGammaScaleGC: float = 1.0
    siteGammaShape: float = 1.0
    states: Optional[np.ndarray] = None  # per-column site categories (1-based), set by samplestates

    @property
    def siteCats(self) -> int:
        return len(self.siteRates)

    def copy(self) -> "ModelParameters":
        import copy
        return copy.deepcopy(self)


def getparams(params: Sequence[float], branchlengths: Sequence[float], siteCats: int = 1,
              lambdaCats: int = 1, fixGU: bool = False, fixGCAU: bool = False) -> ModelParameters:
    """15-vector -> ModelParameters, src/FARCE.jl:609-657.

    Layout: [lambdaGC, lambdaAT, lambdaGT, q1..q6, lambdaGammaShape, siteGammaShape,
    pi1, pi2, pi3, lambdazeroweight].  The lambda categories come from
    discretizegamma3 with lambdaCats - 1 gamma categories and the gamma parameter
    lambdaGammaScaleGC, which this getparams leaves at its constructor default
    of 1 (src/FARCE.jl:629; the shape params[10] enters only the prior).  fixGCAU
    copies the AT rates to GC, fixGU zeroes the GT rates (src/FARCE.jl:631-636).
    With one lambda category every lambda rate is lambda - 1 with weight 1
    (src/FARCE.jl:645-652), whatever fixGU and fixGCAU asked for.  The site rates
    are siteCats equal-weight gamma categories (src/FARCE.jl:641-643).
    """
    p = np.asarray(params, dtype=np.float64)
    if lambdaCats < 1:
        raise ValueError("lambdaCats must be at least 1")
    freqs = np.array([p[11], p[12], p[13], 1.0 - p[11] - p[12] - p[13]])
    mp = ModelParameters(freqs=freqs, lambdaGC=p[0], lambdaAT=p[1], lambdaGT=p[2], q=p[3:9].copy(),
                         branchlengths=np.asarray(branchlengths, dtype=np.float64).copy(),
                         lambdazeroweight=p[14])
    mp.lambdaGammaShapeGC = p[9]
    mp.siteGammaShape = p[10]
    (mp.lambdaweightsGC, mp.lambdaratesGC, _, mp.lambdaratesAT, _, mp.lambdaratesGT) = discretizegamma3(
        mp.lambdazeroweight, mp.lambdaGC, mp.lambdaAT, mp.lambdaGT, mp.lambdaGammaScaleGC, lambdaCats - 1)
    if fixGCAU:
        mp.lambdaratesGC = mp.lambdaratesAT.copy()
    if fixGU:
        mp.lambdaratesGT = np.zeros(len(mp.lambdaratesGT))
    mp.siteWeights = np.ones(siteCats) / siteCats
    mp.siteRates = discretizegamma(mp.siteGammaShape, 1.0 / mp.siteGammaShape, siteCats)
    if len(mp.lambdaweightsGC) == 1:
        mp.lambdaratesGC = np.array([mp.lambdaGC - 1.0])
        mp.lambdaratesAT = np.array([mp.lambdaAT - 1.0])
        mp.lambdaratesGT = np.array([mp.lambdaGT - 1.0])
        mp.lambdaweightsGC = np.array([1.0])
    return mp


def getmusespecificparamsarray(params: ModelParameters) -> List[MuseSpecificParameters]:
    """src/FARCE.jl:676-688."""
    return [MuseSpecificParameters(i + 1, float(w), float(r))
            for i, (w, r) in enumerate(zip(params.lambdaweightsGC, params.lambdaratesGC))]


def site_components(params: ModelParameters):
    """Mixture components of the 4-state site model, src/FARCE.jl:742-773 and 1198-1220.

    Yields (P4 list over nodes, pi4, log weight) per site-rate category.
    """
    Q = gtr(*params.q, params.freqs)
    out = []
    for r, w in zip(params.siteRates, params.siteWeights):
        out.append((gettransitionmatriceslist(params.branchlengths, Q * r), params.freqs.copy(),
                    float(np.log(w))))
    return out


def pair_components(params: ModelParameters):
    """Mixture components of the 16-state pair model, src/FelsensteinCUDA.jl:277-419.

    Components are ordered siteCat1-major, then siteCat2, then lambda category,
    with log weight log w_j + log w_k + log omega_m (src/FelsensteinCUDA.jl:413,417).
    """
    out = []
    musep = getmusespecificparamsarray(params)
    for j, (r1, w1) in enumerate(zip(params.siteRates, params.siteWeights)):
        for k, (r2, w2) in enumerate(zip(params.siteRates, params.siteWeights)):
            for m in musep:
                gc = params.lambdaratesGC[m.lambdacat - 1] + 1.0
                at = params.lambdaratesAT[m.lambdacat - 1] + 1.0
                gt = params.lambdaratesGT[m.lambdacat - 1] + 1.0
                Q = muse(gc, at, gt, *params.q, params.freqs, r1, r2)
                pi16 = calculatemusefreqs(params.freqs, gc, at, gt)
                out.append((gettransitionmatriceslist(params.branchlengths, Q), pi16,
                            float(np.log(w1) + np.log(w2) + m.logprob)))
    return out


def computeprior(params: ModelParameters, samplebranchlengths: bool = False,
                 fixLambdaGT: bool = False) -> float:
    """Log prior, src/FARCE.jl:2401-2436 (shape/scale terms of the unused lambda
    gamma for AT and GT are at their constructor default of 1.0)."""
    from scipy.stats import beta
    pr = -(params.lambdaGC - 1.0) * 0.1 - (params.lambdaAT - 1.0) * 0.1
    if not fixLambdaGT:
        pr += -(params.lambdaGT - 1.0) * 0.1
    pr += float(-0.1 * np.sum(params.q))
    if samplebranchlengths:
        pr += float(-0.2 * np.sum(params.branchlengths[1:]))
    pr += -0.1 * params.lambdaGammaShapeGC - 0.1 * params.lambdaGammaScaleGC
    pr += -0.1 * 1.0 - 0.1 * 1.0 - 0.1 * 1.0 - 0.1 * 1.0
    pr += -0.1 * params.siteGammaShape
    pr += float(beta.logpdf(params.lambdazeroweight, 2.0, 2.0))
    return pr
