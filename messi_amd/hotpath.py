"""The reference's hot-path entry points, with their names and argument meaning,
routed to libmessi (HIP, gfx950).  There is no CPU fallback in this package.

===============================  ==========================================================
Function here                    Reference
===============================  ==========================================================
``felsensteincuda``              src/FelsensteinCUDA.jl:214-516
``computeunpairedlikelihoods``   src/FARCE.jl:1198-1220
``coevolutionall``               src/FARCE.jl:1736-1858 (integratesiterates=true, GPU branch)
``computeinsideKH99``            src/InsideCUDA.jl:23-53
``computeoutsideKH99``           src/InsideCUDA.jl:55-61
``computebasepairprobs``         src/InsideOutsideAlgorithm.jl:486-533
``computestructureintegratedlikelihood``  src/FARCE.jl:2322-2345
``getPosteriorDecodingConsensusStructure``  src/RNATools.jl:76-85
``computeposteriordecodingcuda``  src/InsideCUDA.jl:183-215
``iterativeposteriordecoding``   src/RNATools.jl:6-42
``consensus``                    src/MESSI.jl:1097-1115 (posteriors and decoding in one device pass)
``felsensteincudapaired``        src/FelsensteinCUDA.jl:721-912 (without its leftnumpairs offset)
``computemodellikelihood``       src/FARCE.jl:2439-2515 (fixed structure, site rates integrated)
``samplestructure``              src/InsideOutsideAlgorithm.jl:623-627
``samplestructures``             the same, many samples per device launch
``calculateKH99prior``           src/InsideOutsideAlgorithm.jl:86-144 (host, for structures given by the caller)
``maskgapped``                   src/RNATools.jl:441-451 (host, in place; ``gapmask`` gives the same rule as a
                                 dataset pair mask, so the device never prunes the masked pairs)
``maskgapped_structure``         src/RNATools.jl:453-466
``felsensteinposteriorsiterateconditionalscuda``  src/FelsensteinCUDA.jl:518-717
===============================  ==========================================================

``prec`` selects the plain FP64 path (default) or the scaled FP32 path.
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import numpy as np

from . import lib
from .model import (ModelParameters, MuseSpecificParameters, computeprior, pair_components,
                    site_components)


class KH99:
    """Marker for the KH99 grammar (src/InsideOutsideAlgorithm.jl:30-53); its rule
    probabilities are compiled into the library."""

    RULES = (("S", "LS", 0.868534), ("S", "s", 0.117609877998), ("S", "dFd", 0.013856122002),
             ("F", "dFd", 0.787640), ("F", "LS", 0.21236), ("L", "s", 0.894603), ("L", "dFd", 0.105397))


# Device datasets built with a pair mask, kept per dataset (least recently used first out).
MASKED_CACHE = 4


def _cached(dataset, attr: str, key, make, limit: int):
    """Dataset-attached cache of device datasets keyed by content: at most ``limit``
    entries, the least recently used one closed when a new one arrives."""
    cache = dataset.__dict__.setdefault(attr, {})
    if key in cache:
        cache[key] = cache.pop(key)   # most recently used last
        return cache[key]
    while len(cache) >= limit:
        old = cache.pop(next(iter(cache)))
        old.close()
    cache[key] = make()
    return cache[key]


def mask_key(mask: np.ndarray):
    """Content key of a pair mask: its shape and the packed bits of mask > 0
    (column-major, as the library reads it)."""
    m = np.asfortranarray(np.asarray(mask) > 0)
    return m.shape, np.packbits(m.reshape(-1, order="F")).tobytes()


def device_dataset(dataset, maxbasepairdistance: int = 1000000, mask: Optional[np.ndarray] = None
                   ) -> "lib.DeviceDataset":
    """Device-resident copy of a Dataset for one distance bound and pair mask,
    cached on the dataset by the bound and the mask's content."""
    maxbp = int(min(maxbasepairdistance, 2**31 - 1))
    if mask is None:
        return _cached(dataset, "_device", maxbp, lambda: lib.DeviceDataset(dataset, maxbp), 1 << 30)
    return _cached(dataset, "_device_masked", (maxbp, mask_key(mask)),
                   lambda: lib.DeviceDataset(dataset, maxbp, mask), MASKED_CACHE)


def gapmask(gapfrequency, cutoff: float) -> np.ndarray:
    """Pair mask of maskgapped! (src/RNATools.jl:441-451): pair (i, j) is allowed
    unless gapfrequency[i] >= cutoff or gapfrequency[j] >= cutoff (uint8 L x L)."""
    g = np.asarray(gapfrequency, dtype=np.float64)
    ok = g < cutoff
    return np.asfortranarray(np.logical_and.outer(ok, ok).astype(np.uint8))


def maskgapped(probs: np.ndarray, gapfrequency, cutoff: float, v: float = 0.0) -> np.ndarray:
    """maskgapped! (src/RNATools.jl:441-451), in place: probs[i, j] = v when
    gapfrequency[i] >= cutoff or gapfrequency[j] >= cutoff (the reference's
    callers pass v = -Inf for a pair log-likelihood matrix).  Returns probs."""
    g = np.asarray(gapfrequency, dtype=np.float64)
    if probs.shape != (len(g), len(g)):
        raise ValueError("probs must be L x L with L = len(gapfrequency)")
    bad = g >= cutoff
    probs[bad, :] = v
    probs[:, bad] = v
    return probs


def maskgapped_structure(paired, gapfrequency, cutoff: float) -> np.ndarray:
    """maskgapped for a structure (src/RNATools.jl:453-466): a copy of the 1-based
    partners with every pair (i, j) removed whose i or j has gap frequency >= cutoff."""
    s = np.array(paired, dtype=np.int64)
    g = np.asarray(gapfrequency, dtype=np.float64)
    i = np.flatnonzero(s > np.arange(1, len(s) + 1))
    j = s[i] - 1
    drop = (g[i] >= cutoff) | (g[j] >= cutoff)
    s[i[drop]] = 0
    s[j[drop]] = 0
    return s


def felsensteincuda_gapmasked(dataset, params: ModelParameters, gapcutoff: float,
                              maxbasepairdistance: int = 1000000, prec=lib.FP64) -> np.ndarray:
    """felsensteincuda followed by maskgapped! at ``gapcutoff`` (the ML objective's
    pair matrix, src/MESSI.jl:309-310), with the masked pairs never pruned: the
    device dataset carries the gap mask, so those entries come back -inf."""
    mask = gapmask(dataset.gapfrequency, gapcutoff)
    M, _ = felsensteincuda(dataset, params, None, maxbasepairdistance, prec=prec, mask=mask)
    return M


def _stack(components):
    P = np.stack([c[0] for c in components])
    pi = np.stack([c[1] for c in components])
    lw = np.array([c[2] for c in components])
    return P, pi, lw


def computeunpairedlikelihoods(dataset, params: ModelParameters, prec=lib.FP64,
                               maxbasepairdistance: int = 1000000) -> np.ndarray:
    """Site log-likelihoods mixed over site-rate categories."""
    dd = device_dataset(dataset, maxbasepairdistance)
    return dd.site_loglik(*_stack(site_components(params)), prec=prec)


def felsensteincuda(dataset, params: ModelParameters, museparams: Optional[List[MuseSpecificParameters]] = None,
                    maxbasepairdistance: int = 1000000, keepmuseconditionals: bool = False,
                    prec=lib.FP64, mask: Optional[np.ndarray] = None
                    ) -> Tuple[np.ndarray, List[np.ndarray]]:
    """All-pairs 16-state pruning with the site-category x lambda-category mixture.

    Returns (L x L matrix, per-component matrices if ``keepmuseconditionals``).
    The lambda categories come from ``params`` (``museparams`` is accepted for
    signature parity; with one lambda category it is ``getmusespecificparamsarray``).
    """
    dd = device_dataset(dataset, maxbasepairdistance, mask)
    comps = pair_components(params)
    P, pi, lw = _stack(comps)
    if keepmuseconditionals:
        mat, per = dd.pair_loglik(P, pi, lw, prec=prec, percomp=True)
        return np.asarray(mat), per
    return np.asarray(dd.pair_loglik(P, pi, lw, prec=prec)), []


def coevolutionall(dataset, params: ModelParameters, parallel: bool = True, keepmatrices: bool = False,
                   integratesiterates: bool = True, maxbasepairdistance: int = 1000000,
                   usecuda: bool = True, prec=lib.FP64):
    if not (integratesiterates and usecuda):
        raise NotImplementedError("only the integrated-site-rate device path is provided")
    return felsensteincuda(dataset, params, None, maxbasepairdistance, keepmatrices, prec=prec)


def computeinsideKH99(unpairedlogprobs, pairedlogprobs, B: float = 1.0, zonly: bool = True,
                      usecudain: bool = True, prec=lib.FP64):
    """log Z, or the (3, L, L) inside table [S, F, L] when ``zonly`` is false."""
    if not usecudain:
        raise NotImplementedError("the CPU inside algorithm is not part of this library")
    r = lib.inside(unpairedlogprobs, pairedlogprobs, B, zonly, prec)
    return r if zonly else r[1]


def computeoutsideKH99(inside, unpairedlogprobs, pairedlogprobs, usecuda: bool = True, prec=lib.FP64):
    if not usecuda:
        raise NotImplementedError("the CPU outside algorithm is not part of this library")
    return lib.outside(inside, unpairedlogprobs, pairedlogprobs, prec)


def computebasepairprobs(inside, outside, unpairedlogprobs, pairedlogprobs, grammar: KH99 = None,
                         prec=lib.FP64):
    """(unpaired probabilities (L), pair probabilities (L x L)), row-normalised."""
    return lib.bpp(inside, outside, unpairedlogprobs, pairedlogprobs, prec)


def computestructureintegratedlikelihood(dataset, params: ModelParameters, maxbasepairdistance: int = 1000000,
                                         samplebranchlengths: bool = False, fixLambdaGT: bool = False,
                                         integratesiterates: bool = True, usecuda: bool = True,
                                         prec=lib.FP64) -> float:
    """log prior + log Z of the structure-integrated model (one device evaluation)."""
    p = params
    if fixLambdaGT:
        p = params.copy()
        p.lambdaGT = 1.0
        p.lambdaratesGT = np.array([0.0])
    dd = device_dataset(dataset, maxbasepairdistance)
    z = dd.eval(_stack(site_components(p)), _stack(pair_components(p)), 1.0, prec)
    return computeprior(p, samplebranchlengths, fixLambdaGT) + z


def loglikelihood(dataset, params: ModelParameters, maxbasepairdistance: int = 1000000, prec=lib.FP64) -> float:
    """log Z only (the eval without the prior)."""
    dd = device_dataset(dataset, maxbasepairdistance)
    return dd.eval(_stack(site_components(params)), _stack(pair_components(params)), 1.0, prec)


def posterior(dataset, params: ModelParameters, maxbasepairdistance: int = 1000000, prec=lib.FP64):
    """(log Z, unpaired probabilities, pair probabilities) in one device pass."""
    dd = device_dataset(dataset, maxbasepairdistance)
    return dd.posterior(_stack(site_components(params)), _stack(pair_components(params)), prec)


def getPosteriorDecodingConsensusStructure(pairedposteriorprobs, unpairedposteriorprobs, usecuda: bool = False,
                                           alpha: float = 2.0) -> np.ndarray:
    """MEA consensus pairing: 1-based partner per position, 0 when unpaired.

    Both settings run on the device.  usecuda=True keeps the semantics of the
    reference's CUDA branch (computeposteriordecodingcuda: FP32, the posterior
    matrix read at i*L + k, alpha*p + E fused); usecuda=False those of its CPU
    branch's iterativeposteriordecoding (FP64), called with the arguments in the
    order of that function's signature (the reference's CPU branch passes them
    swapped, src/RNATools.jl:83 versus :6, and fails with a MethodError)."""
    return lib.mea(unpairedposteriorprobs, pairedposteriorprobs, alpha, lib.FP32 if usecuda else lib.FP64)


def computeposteriordecodingcuda(singleprobsin, pairprobsin, alpha: float = 2.0):
    """(E, S) tables of the FP32 decoding, L x L (S: -1, -2 or the 1-based split)."""
    _, E, S = lib.mea(singleprobsin, pairprobsin, alpha, lib.FP32, tables=True)
    return E, S


def iterativeposteriordecoding(singleprobs, pairprobs, alpha: float = 2.0):
    """(E, S) tables of the FP64 decoding, computed on the device."""
    _, E, S = lib.mea(singleprobs, pairprobs, alpha, lib.FP64, tables=True)
    return E, S


def consensus(dataset, params: ModelParameters, maxbasepairdistance: int = 1000000, prec=lib.FP64,
              usecuda: bool = True, alpha: float = 2.0):
    """(log Z, MEA consensus pairing) with the posteriors kept on the device;
    usecuda selects the decoding semantics as in getPosteriorDecodingConsensusStructure."""
    dd = device_dataset(dataset, maxbasepairdistance)
    return dd.posterior_consensus(_stack(site_components(params)), _stack(pair_components(params)), prec,
                                  lib.FP32 if usecuda else lib.FP64, alpha)


def samplestructures(inside, pairedlogprobs, unpairedlogprobs, nsamples: int, seed: int = 0, B: float = 1.0,
                     prec=lib.FP64):
    """nsamples KH99 structures from (S, 1, L) of an inside table: (structures
    (nsamples x L, 1-based partners), KH99 log priors)."""
    return lib.sample_structures(inside, pairedlogprobs, unpairedlogprobs, nsamples, seed, B, prec)


def samplestructure(seed: int, inside, pairedlogprobs, unpairedlogprobs, x: int, y: int, paired,
                    grammar: KH99 = None, B: float = 1.0, prec=lib.FP64) -> float:
    """One stochastic traceback written into ``paired`` (in place, 1-based partners);
    returns the structure's KH99 log prior.  The reference takes an RNG; here the
    draw is keyed by ``seed``.  Only the whole sequence (x = 1, y = L) is supported,
    as in every reference call."""
    L = len(unpairedlogprobs)
    if (x, y) != (1, L):
        raise ValueError("samplestructure samples the whole sequence only (x = 1, y = L)")
    S, pri = lib.sample_structures(inside, pairedlogprobs, unpairedlogprobs, 1, seed, B, prec)
    paired[:] = S[0]
    return float(pri[0])


def calculateKH99prior(pairedsites) -> float:
    """KH99 log prior of a structure (1-based partners, 0 unpaired): the log
    probability of its unique derivation."""
    import math
    p = [int(v) for v in pairedsites]
    L = len(p)
    r = {("S", "sS"): math.log(0.868534 * 0.894603), ("S", "dFdS"): math.log(0.868534 * 0.105397),
         ("S", "s"): math.log(0.117609877998), ("S", "dFd"): math.log(0.013856122002),
         ("F", "dFd"): math.log(0.787640), ("F", "sS"): math.log(0.21236 * 0.894603),
         ("F", "dFdS"): math.log(0.21236 * 0.105397)}
    ll = 0.0
    stack = [("S", 1, L)] if L >= 1 else []
    while stack:
        sym, s, e = stack.pop()
        if s < 1 or s > L:
            continue
        k = p[s - 1]
        if k == 0:
            if s == e and sym == "S":
                ll += r[("S", "s")]
            elif s < e:
                ll += r[(sym, "sS")]
                stack.append(("S", s + 1, e))
        elif k > s:
            if k == e:
                ll += r[(sym, "dFd")]
                stack.append(("F", s + 1, k - 1))
            else:
                ll += r[(sym, "dFdS")]
                stack.append(("F", s + 1, k - 1))
                stack.append(("S", k + 1, e))
    return ll


def structure_dataset(dataset, paired) -> "lib.DeviceDataset":
    """Device dataset whose pair list is exactly the pairs of one structure (the
    reference's getindexdict(dataset, paired), src/FelsensteinCUDA.jl:77-142);
    cached on the dataset by the structure's content."""
    p = np.ascontiguousarray(paired, dtype=np.int64)
    L = len(p)

    def make():
        mask = np.zeros((L, L), dtype=np.uint8, order="F")
        x = np.flatnonzero(p > np.arange(1, L + 1))
        mask[x, p[x] - 1] = 1
        return lib.DeviceDataset(dataset, L, mask)

    return _cached(dataset, "_device_structure", p.tobytes(), make, 8)


def felsensteincudapaired(dataset, paired, params: ModelParameters, museparams=None, prec=lib.FP64) -> float:
    """Sum of the pair log-likelihoods of the structure's pairs (1-based partners)."""
    dd = structure_dataset(dataset, paired)
    _, (_, pairsum) = dd.structure_loglik(_stack(site_components(params)), _stack(pair_components(params)),
                                         paired, prec)
    return pairsum


def samplestates(seed: int, dataset, params: ModelParameters, paired, B: float = 1.0, prec=lib.FP64):
    """samplestates (src/FARCE.jl:2534-2538): draws the site category of every
    column given the structure paired (1-based partners), one category per
    unpaired column and one category pair per base pair, from the conditionals of
    computeuncachedlikelihood (src/FARCE.jl:1336-1486) raised to B; sets
    params.states and returns (unpaired log-likelihoods, paired log-likelihoods at
    the openers, states).  The reference's rng becomes a seed (the draw of column
    x is keyed by (seed, x)); its museconditionals output is not produced."""
    dd = structure_dataset(dataset, paired)
    states, unp, pll = dd.sample_states(_stack(site_components(params)), _stack(pair_components(params)),
                                        paired, params.siteCats, seed=seed, B=B, prec=prec)
    params.states = states
    return unp, pll, states


def felsensteinposteriorsiterateconditionalscuda(singleprobs, pairprobs, dataset, params: ModelParameters,
                                                 museparams: Optional[List[MuseSpecificParameters]] = None,
                                                 maxbasepairdistance: int = 1000000,
                                                 keepmuseconditionals: bool = False, prec=lib.FP64) -> np.ndarray:
    """(siteCats, L) log posterior weights of the site-rate categories of every
    column, src/FelsensteinCUDA.jl:518-717 (the lambda categories come from params;
    museparams is accepted for signature parity)."""
    dd = device_dataset(dataset, maxbasepairdistance)
    return dd.site_rate_posteriors(_stack(site_components(params)), _stack(pair_components(params)),
                                   params.siteCats, singleprobs, pairprobs, prec)


def computemodellikelihood(dataset, params: ModelParameters, paired, samplebranchlengths: bool = False,
                           useprior: bool = True, fixLambdaGT: bool = False, prec=lib.FP64) -> float:
    """Fixed-structure log posterior: prior + unpaired site log-likelihoods + pair
    log-likelihoods of the structure, in one device call (the site rates are
    integrated, as in the reference's GPU branch)."""
    p = params
    if fixLambdaGT:
        p = params.copy()
        p.lambdaGT = 1.0
        p.lambdaratesGT = np.array([0.0])
    dd = structure_dataset(dataset, paired)
    ll, _ = dd.structure_loglik(_stack(site_components(p)), _stack(pair_components(p)), paired, prec)
    return (computeprior(p, samplebranchlengths, fixLambdaGT) if useprior else 0.0) + ll
