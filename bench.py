"""Benchmark: full-likelihood evaluations per second at 1,500 nt x 50 sequences.

One step = one batched evaluation through the C ABI (messi_batch_eval_loaded): for each
of B parameter vectors of the 50 x 1500 alignment (B independent MCMC chains /
proposals), the site log-likelihoods (4-state pruning), the all-pairs 16-state
pair matrix and the KH99 inside partition function.  All members share each
tree-level launch and each tile-diagonal launch of the inside schedule.  The
alignment, its dedup lists and the transition matrices of each parameter vector
are resident in HBM when the timed region starts (messi_batch_load, untimed);
the rate with the matrices copied from the host on every call is reported
beside it as evals_per_s_host_inputs, and the rate from parameter vectors
(host eigendecompositions, transition matrices built on the device, then the
evaluation, all timed: what an MCMC step pays) as evals_per_s_from_params.  ``value`` is evaluations per second
of the scaled FP32 path at batch B; the plain FP64 path, and both paths at
batch 1 (one chain), are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
       python bench.py --instances      # 64 uneven instances: one batch vs 64 calls
       python bench.py --mcmc           # learned MCMC proposal vs random walk (ESS/s)
       python bench.py --genome         # 30 x 9000 nt posterior, full span and D=500
       python bench.py --mea            # MEA consensus decoding at L = 1500 and 9000
       python bench.py --sample         # structures sampled from the inside table (1024 per evaluation)
       python bench.py --structure      # fixed-structure log-likelihood (true structure)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3}   # MI355X dense TFLOP/s (vector == matrix for these types)


def _components(params):
    from messi_amd import model
    from messi_amd.hotpath import _stack
    return _stack(model.site_components(params)), _stack(model.pair_components(params))


def _perturbed(a, ds, k, seed=0, siteCats=1, lambdaCats=1):
    from messi_amd import model
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        v = a.params.copy()
        if lambdaCats > 1:
            v[14] = 0.3     # lambda-zero weight of the lambda mixture
        if i:
            v[0:3] = np.maximum(1.0, v[0:3] * np.exp(rng.normal(0, 0.05, 3)))
            v[3:9] = v[3:9] * np.exp(rng.normal(0, 0.05, 6))
        out.append(model.getparams(v, ds.tree.branchlengths, siteCats=siteCats, lambdaCats=lambdaCats))
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_eval(ds, params):
    """One full evaluation by the oracle (site log-likelihoods, all L(L-1)/2 pair
    log-likelihoods, the inside recursion): (log Z, seconds per part)."""
    import oracle
    from messi_amd import model
    t0 = time.perf_counter()
    u = oracle.site_loglik(ds, model.site_components(params))
    t1 = time.perf_counter()
    M = oracle.pair_loglik(ds, model.pair_components(params))
    t2 = time.perf_counter()
    z = oracle.inside(u, M)[0, 0, -1]
    t3 = time.perf_counter()
    return z, (t1 - t0, t2 - t1, t3 - t2)


def cpu_baseline(ds, params, L):
    """The oracle (FP64 C restatement of the reference CPU path, whose own code is
    serial) timed on one full evaluation: single-threaded, as the reference runs,
    and with its OpenMP loops on every core the process may use."""
    import oracle
    n_all = oracle.get_threads()
    oracle.set_threads(1)
    try:
        z1, t1 = _oracle_eval(ds, params)
    finally:
        oracle.set_threads(n_all)
    zN, tN = _oracle_eval(ds, params)
    npairs = L * (L - 1) // 2
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": 1.0 / sum(t1), "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"one full evaluation: all {L} site columns, all {npairs} column pairs, the full "
                      f"inside recursion at L={L} (site {t1[0]:.2f} s, pairs {t1[1]:.1f} s, inside {t1[2]:.1f} s)",
            "all_cores": {"value": 1.0 / sum(tN), "cores": n_all,
                          "seconds": {"site": tN[0], "pairs": tN[1], "inside": tN[2]},
                          "same_logZ": bool(zN == z1)},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": aff}


def time_batch(b, sites, pairs, prec, steps, warmup, tune: bool = False):
    """Time evaluations with the mixture components resident in HBM (loaded once,
    outside the timed region), then the same with host inputs copied every call
    (the PCIe-inclusive rate), then a profiling pass with phase events.  With
    tune, the autotuner first picks the batch's schedule (untimed)."""
    from messi_amd import autotune, lib
    ctx = lib.context()
    ctx.set_profiling(False)
    b.load(sites, pairs, prec)
    b.schedule = None
    if tune:
        ch = autotune.tune_batch(b, prec, reps=3)
        b.schedule = {"tile32_from_length": ch["threshold"], "window": ch["window"], "groups": ch["groups"],
                      "tile_candidates_ms": ch["candidates"], "window_candidates_ms": ch["window_candidates"],
                      "group_candidates_ms": ch["group_candidates"]}
    for _ in range(warmup):
        z = b.eval_loaded(1.0, prec)
    t0 = time.perf_counter()
    for _ in range(steps):
        z = b.eval_loaded(1.0, prec)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        b.eval(sites, pairs, 1.0, prec)
    ms_host = (time.perf_counter() - t0) * 1e3 / steps
    ctx.set_profiling(True)
    ph = np.zeros(5)
    steps_ms, steps_n = np.zeros(2), np.zeros(2)
    nprof = max(3, min(10, steps))
    for _ in range(nprof):
        b.eval_loaded(1.0, prec)
        ph += b.timings()
        sm, sn = b.step_timings()
        steps_ms += sm
        steps_n += sn
    ph /= nprof
    ctx.set_profiling(False)
    # HIP events around the inside step kernels alone: their mean launch duration,
    # and the launches per evaluation, per tile edge (16, 32)
    b.step_launch_us = float(steps_ms.sum() / steps_n.sum() * 1e3) if steps_n.sum() else None
    b.step_launches = (steps_n / nprof).tolist()
    return ms, z, ph, ms_host


def time_from_params(b, pm, thetas, prec, steps, warmup):
    """Evaluations from parameter vectors: the host builds each vector's rate
    matrices and their eigendecompositions, the device every branch's transition
    matrix (messi_batch_load_eigen) and then the evaluation; all of it timed."""
    for _ in range(warmup):
        b.load_eigen(*pm.eigen_components(thetas), prec=prec)
        z = b.eval_loaded(1.0, prec)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.load_eigen(*pm.eigen_components(thetas), prec=prec)
        z = b.eval_loaded(1.0, prec)
    return (time.perf_counter() - t0) * 1e3 / steps, z


def main_instances(args):
    """64 uneven instances (L ~ U[80,600], N ~ U[10,60]) in one batch vs 64 calls,
    with the inside tile edges chosen by the autotuner for the batch."""
    from messi_amd import lib, model, synth
    from messi_amd.hotpath import device_dataset
    als = synth.config4()
    dss = [synth.dataset_of(a) for a in als]
    dds = [device_dataset(d) for d in dss]
    comps = [_components(model.getparams(a.params, d.tree.branchlengths)) for a, d in zip(als, dss)]
    sites, pairs = [c[0] for c in comps], [c[1] for c in comps]
    from messi_amd import autotune
    b = lib.Batch(dds)
    ctx = lib.context()
    res = {}
    for prec in ("fp64", "fp32"):
        # schedule from the autotuner (timed on this batch, outside the timed region)
        b.load(sites, pairs, prec)
        tuned = autotune.tune_batch(b, prec)
        ms_b, zb, ph, ms_bh = time_batch(b, sites, pairs, prec, args.steps, args.warmup)
        tile_of = autotune.tiles_for([d.numcols for d in dss], tuned["threshold"])

        def sequential():
            out = []
            for dd, s, p, t in zip(dds, sites, pairs, tile_of):
                ctx.set_tile(t)
                out.append(dd.eval(s, p, 1.0, prec))
            ctx.set_tile(16)
            return out
        for _ in range(args.warmup):
            sequential()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            zs = sequential()
        ms_s = (time.perf_counter() - t0) * 1e3 / args.steps
        res[prec] = {"batched_evals_per_s": 64e3 / ms_b, "batched_evals_per_s_host_inputs": 64e3 / ms_bh,
                     "sequential_evals_per_s": 64e3 / ms_s,
                     "speedup_host_inputs": ms_s / ms_bh,
                     "max_rel_diff": float(np.max(np.abs(zb - zs) / np.abs(zs))),
                     "batched_phase_ms": {"site": ph[0], "pair": ph[1], "inside": ph[2]},
                     "autotuner": {"tile32_from_length": tuned["threshold"], "candidates_ms": tuned["candidates"]}}
    out = {"metric": "total evals/s, 64 uneven instances in one batch vs 64 sequential calls",
           "value": res["fp32"]["batched_evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "64 instances, L ~ U[80,600], N ~ U[10,60], seed 4",
                      "L": [d.numcols for d in dss], "N": [d.numseqs for d in dss]},
           "variants": res}
    print(json.dumps(out), flush=True)


def _device_memory():
    """(free, total) bytes of device 0 from the HIP runtime."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0:
            return int(free.value), int(total.value)
    except OSError:
        pass
    return None, None


def main_genome(args):
    """30 sequences x 9000 nt: posterior (site, pair, inside, outside, bpp) at full
    span and at maximum base-pair distance 500, FP64 and FP32; errors of the FP32
    path against FP64, phase times and device memory."""
    from messi_amd import lib, model, synth

    t0 = time.perf_counter()

    def log(*m):
        print("[genome]", *m, f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

    a = synth.config3()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    site, pair = _components(p)
    ctx = lib.context()
    free0, total = _device_memory()
    out = {"metric": "genome-scale posterior, 30 x 9000 nt: FP32 vs FP64 error and time", "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic (seeded Yule tree, 30% paired, helices <= 12)",
           "config": {"workload": "config3: 30 sequences x 9000 nt, posterior at full span and at D=500",
                      "L": ds.numcols, "nseq": ds.numseqs}, "runs": {}}
    for D in args.genome_spans:
        tb = time.perf_counter()
        dd = lib.DeviceDataset(ds, D)
        build_s = time.perf_counter() - tb
        st = dd.stats()
        log(f"D={D}: dataset built in {build_s:.1f} s, stats {st}")
        r = {"dataset_build_s": build_s, "stats": st}
        res = {}
        for prec in ("fp64", "fp32"):
            ctx.set_profiling(False)
            tw = time.perf_counter()
            dd.posterior(site, pair, prec)            # allocates the tables
            first_s = time.perf_counter() - tw
            # the pair posteriors land in a page-locked buffer allocated once, as a
            # caller that computes posteriors repeatedly would keep it
            ppbuf = lib.pinned_empty((ds.numcols, ds.numcols), order="F")
            ctx.set_profiling(True)
            tw = time.perf_counter()
            z, unp, pp = dd.posterior(site, pair, prec, out=ppbuf)
            wall = time.perf_counter() - tw
            ph = dd.timings()
            ctx.set_profiling(False)
            free, _ = _device_memory()
            res[prec] = (z, unp, pp)
            r[prec] = {"logZ": z, "posterior_s": wall, "first_call_s": first_s,
                       "phase_s": {k: float(v) / 1e3 for k, v in zip(("site", "pair", "inside", "outside", "bpp"), ph)},
                       "device_bytes_in_use": (total - free) if free is not None else None}
            log(f"D={D} {prec}: logZ {z:.6f}, posterior {wall:.2f} s, phases (ms) {np.round(ph, 1).tolist()}")
        z64, u64, p64 = res["fp64"]
        z32, u32, p32 = res["fp32"]
        r["fp32_logZ_rel_err"] = abs(z32 - z64) / abs(z64)
        r["fp32_bpp_max_abs_err"] = float(max(np.max(np.abs(u32 - u64)), np.max(np.abs(p32 - p64))))
        r["fp64_row_sum_max_dev"] = float(np.max(np.abs(u64 + p64.sum(axis=1) - 1.0)))
        log(f"D={D}: fp32 logZ rel err {r['fp32_logZ_rel_err']:.2e}, bpp max abs err {r['fp32_bpp_max_abs_err']:.2e}")
        out["runs"][str(D)] = r
        del res, z64, u64, p64, z32, u32, p32
        dd.close()
    out["device_total_bytes"] = total
    first = out["runs"][str(args.genome_spans[0])]
    out["value"] = first["fp32"]["posterior_s"]
    print(json.dumps(out), flush=True)


def _structure_mixture(L, n, rng):
    """Posteriors (getpairprobs, src/RNATools.jl:87-102) of n samples of one random
    nested structure (35 % of sites paired, helices of 4-8 bp), each sample keeping
    every pair with probability 0.75."""
    from messi_amd import synth
    partner = synth._structure(L, 0.35, 4, 8, rng)
    single = np.zeros(L)
    pair = np.zeros((L, L), order="F")
    opens = np.flatnonzero(partner > np.arange(L))
    for _ in range(n):
        keep = opens[rng.random(len(opens)) < 0.75]
        pair[keep, partner[keep]] += 1.0
        pair[partner[keep], keep] += 1.0
        u = np.ones(L)
        u[keep] = 0.0
        u[partner[keep]] = 0.0
        single += u
    return single / n, pair / n


def main_mea(args):
    """MEA consensus decoding (fill + traceback) of L x L posteriors at L = 1500
    and 9000, FP64 (the Julia loop's semantics) and FP32 (the CUDA kernel's), with
    the posteriors resident in HBM (device event times) and with the host upload."""
    from messi_amd import lib
    ctx = lib.context()
    rng = np.random.default_rng(args.seed)
    runs = {}
    for L in args.mea_lengths:
        single, pair = _structure_mixture(L, 8, rng)
        terms = L * (L - 1) * (L - 2) // 6     # (i, k, j) with i < k < j
        r = {}
        paired = {}
        for prec in ("fp64", "fp32"):
            ctx.set_profiling(False)
            for _ in range(args.warmup):
                lib.mea(single, pair, 2.0, prec)
            ctx.set_profiling(True)
            fill, tb, wall = [], [], []
            for _ in range(args.steps):
                t0 = time.perf_counter()
                paired[prec] = lib.mea(single, pair, 2.0, prec)
                wall.append(time.perf_counter() - t0)
                t = lib.mea_timings()
                fill.append(t[1])
                tb.append(t[2])
            ctx.set_profiling(False)
            f, b, w = float(np.mean(fill)), float(np.mean(tb)), float(np.mean(wall)) * 1e3
            peak = PEAK[prec]
            ach = 2 * terms / (f * 1e-3) / 1e12
            r[prec] = {"fill_ms": f, "traceback_ms": b, "decodings_per_s": 1e3 / (f + b),
                       "decodings_per_s_host_inputs": 1e3 / w, "wall_ms_host_inputs": w,
                       "fill_tflops_equiv": ach, "fill_frac_of_vector_peak": ach / peak,
                       "pairs": int(np.count_nonzero(paired[prec]) // 2)}
            print(f"[mea] L={L} {prec}: fill {f:.3f} ms, traceback {b:.3f} ms, host-input wall {w:.2f} ms",
                  file=sys.stderr, flush=True)
        r["fp32_fp64_same_pairing"] = bool(np.array_equal(paired["fp32"], paired["fp64"]))
        runs[str(L)] = r
    L0 = str(args.mea_lengths[0])
    lead = runs[L0]["fp32"]
    out = {"metric": f"MEA consensus decodings/s at L={L0} (fill + traceback, posteriors resident)",
           "value": lead["decodings_per_s"], "unit": "decodings/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": lead["fill_ms"] + lead["traceback_ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic posteriors: 8 samples of one random nested structure, each pair kept with p=0.75",
           "config": {"workload": f"MEA decoding, alpha 2, L in {args.mea_lengths}"},
           "roofline": {"bound": "vector", "achieved": lead["fill_tflops_equiv"], "peak": PEAK["fp32"],
                        "unit": "TFLOP/s", "frac": lead["fill_frac_of_vector_peak"], "traffic": None,
                        "kernel": "mea_step_kernel<float,16>: 2 ops (add, max) per (i, k, j) term"},
           "runs": runs}
    if not args.no_cpu:
        import oracle
        L = int(L0)
        single, pair = _structure_mixture(L, 8, np.random.default_rng(args.seed))
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < 10.0 or n == 0:
            _, S = oracle.mea(single, pair, 2.0, "fp64")
            oracle.mea_traceback(S)
            n += 1
        dt = (time.perf_counter() - t0) / n
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "decodings/s", "cores": 1, "kind": "port",
                               "sample": f"{n} full FP64 decodings at L={L} (oracle fill + traceback)"}
    print(json.dumps(out), flush=True)


def main_sample(args):
    """Structure sampling (stochastic traceback) on the 50 x 1500 alignment: one
    step = one evaluation plus args.samples structures drawn from its inside
    table on the device; the sampling time is the step time minus the
    evaluation-only step time.  FP64 and FP32 tables."""
    from messi_amd import lib, model, synth
    from messi_amd.hotpath import device_dataset
    a = synth.config2()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    site, pair = _components(p)
    dd = device_dataset(ds)
    L = ds.numcols
    n = args.samples
    res = {}
    for prec in ("fp64", "fp32"):
        for _ in range(args.warmup):
            dd.sample_posterior(site, pair, n, seed=1, prec=prec)
        t0 = time.perf_counter()
        for k in range(args.steps):
            dd.sample_posterior(site, pair, 0, seed=k, prec=prec)
        ev = (time.perf_counter() - t0) * 1e3 / args.steps
        t0 = time.perf_counter()
        for k in range(args.steps):
            z, S, pri = dd.sample_posterior(site, pair, n, seed=k, prec=prec)
        tot = (time.perf_counter() - t0) * 1e3 / args.steps
        smp = max(tot - ev, 1e-6)
        res[prec] = {"eval_ms": ev, "eval_plus_sampling_ms": tot, "sampling_ms": smp,
                     "structures_per_s": n * 1e3 / smp, "mean_pairs": float((S > 0).sum(1).mean() / 2),
                     "mean_logprior": float(np.mean(pri))}
        print(f"[sample] {prec}: eval {ev:.2f} ms, eval + {n} samples {tot:.2f} ms", file=sys.stderr, flush=True)
    out = {"metric": f"structures sampled per second at 1500nt x 50seq ({n} per evaluation)",
           "value": res["fp32"]["structures_per_s"], "unit": "structures/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["fp32"]["sampling_ms"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded Yule tree, GTR/MuSE simulation, 40% paired, 5% gaps)",
           "config": {"workload": f"config2: 50 sequences x 1500 nt; {n} stochastic tracebacks per evaluation",
                      "L": L, "samples": n}, "variants": res}
    if not args.no_cpu:
        import oracle
        u = oracle.site_loglik(ds, model.site_components(p))
        M = oracle.pair_loglik(ds, model.pair_components(p), maxbp=L)
        ins = oracle.inside(u, M)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 10.0 or k == 0:
            oracle.sample_structures(ins, M, 8, seed=5, first=8 * k)
            k += 1
        dt = (time.perf_counter() - t0) / (8 * k)
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "structures/s", "cores": 1, "kind": "port",
                               "sample": f"{8 * k} oracle tracebacks at L={L} from the oracle's FP64 inside table "
                                         f"(the table itself not timed)"}
    print(json.dumps(out), flush=True)


def main_siterate(args):
    """Site-rate posteriors (felsensteinposteriorsiterateconditionalscuda) at 50 x 1500:
    one call per step from base-pair posteriors already on the host (the call copies
    them in), for the 1 x 1 and 3 x 5 category mixtures, both precisions."""
    from messi_amd import synth, model, hotpath
    a = synth.config2()
    ds = synth.dataset_of(a)
    L = ds.numcols
    dd = hotpath.device_dataset(ds)
    rng = np.random.default_rng(0)
    pp = np.triu(rng.random((L, L)) / L, 1)
    pp = pp + pp.T
    single = 1.0 - pp.sum(axis=1)
    res = {}
    for sc_, lc_ in ((1, 1), (3, 5)):
        p = model.getparams(a.params.copy(), ds.tree.branchlengths, siteCats=sc_, lambdaCats=lc_)
        sc = hotpath._stack(model.site_components(p))
        pc = hotpath._stack(model.pair_components(p))
        for prec in ("fp64", "fp32"):
            for _ in range(args.warmup):
                dd.site_rate_posteriors(sc, pc, sc_, single, pp, prec)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                dd.site_rate_posteriors(sc, pc, sc_, single, pp, prec)
            res[f"{sc_}x{lc_}_{prec}"] = (time.perf_counter() - t0) * 1e3 / args.steps
    print(json.dumps({"metric": "site-rate posteriors, ms per call (host posteriors in, weights out)",
                      "unit": "ms", "config": {"workload": "50 sequences x 1500 nt, full span",
                                               "L": L, "steps": args.steps}, "ms": res}), flush=True)


def main_copy_calibration(args):
    """One device-to-device copy of 256 MiB (torch clone): the reference kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the same counters of the likelihood
    kernels to bytes (a copy reads and writes its size once)."""
    import torch
    n = 256 << 20
    x = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    torch.cuda.synchronize()
    y = x.clone()
    torch.cuda.synchronize()
    print(json.dumps({"mode": "copy calibration", "bytes_read": n, "bytes_written": n,
                      "checksum": int(y[:4096].sum().item())}), flush=True)


def main_structure(args):
    """Fixed-structure log-likelihood on the 50 x 1500 alignment with its true
    structure: one step = site pruning of every column, pair pruning of the
    structure's pairs only (a structure-masked pair list) and the sums, with the
    transition matrices copied from the host each call."""
    from messi_amd import hotpath, lib, model, synth
    a = synth.config2()
    ds = synth.dataset_of(a)
    p = model.getparams(a.params, ds.tree.branchlengths)
    site, pair = _components(p)
    true = np.asarray(a.pairedsites)
    paired = np.where(true >= 0, true + 1, 0)
    dd = hotpath.structure_dataset(ds, paired)
    res = {}
    for prec in ("fp64", "fp32"):
        for _ in range(args.warmup):
            dd.structure_loglik(site, pair, paired, prec)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ll, parts = dd.structure_loglik(site, pair, paired, prec)
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        res[prec] = {"ms_per_eval": ms, "evals_per_s": 1e3 / ms, "loglik": ll, "parts": parts}
    # state sampling given the structure (samplestates), 4 site categories: one step =
    # per-category site pruning, per-component pair pruning of the structure's pairs,
    # one draw per unpaired column and per pair
    p4 = model.getparams(a.params, ds.tree.branchlengths, siteCats=4)
    site4, pair4 = _components(p4)
    for prec in ("fp64", "fp32"):
        for _ in range(args.warmup):
            dd.sample_states(site4, pair4, paired, 4, seed=0, prec=prec)
        t0 = time.perf_counter()
        for k in range(args.steps):
            dd.sample_states(site4, pair4, paired, 4, seed=k, prec=prec)
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        res[prec]["states_4_categories"] = {"ms_per_draw": ms, "draws_per_s": 1e3 / ms}
    out = {"metric": "fixed-structure log-likelihood evals/s at 1500nt x 50seq",
           "value": res["fp32"]["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["fp32"]["ms_per_eval"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded Yule tree, GTR/MuSE simulation, 40% paired, 5% gaps); its true structure",
           "config": {"workload": "config2: 50 sequences x 1500 nt, fixed structure "
                                  f"({int(np.count_nonzero(paired) // 2)} pairs), host inputs each call",
                      "L": ds.numcols, "pairs": int(np.count_nonzero(paired) // 2)},
           "variants": res,
           "fp32_rel_err": abs(res["fp32"]["loglik"] - res["fp64"]["loglik"]) / abs(res["fp64"]["loglik"]),
           "dataset_stats": dd.stats()}
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        u = oracle.site_loglik(ds, model.site_components(p))
        mask = np.zeros((ds.numcols, ds.numcols), dtype=bool)
        x = np.flatnonzero(paired > np.arange(1, ds.numcols + 1))
        mask[x, paired[x] - 1] = True
        M = oracle.pair_loglik(ds, model.pair_components(p), mask=mask)
        ll = float(np.sum(u[paired == 0]) + np.sum(M[x, paired[x] - 1]))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "evals/s", "cores": 1, "kind": "port",
                               "sample": f"one full oracle evaluation (all sites, the {len(x)} structure pairs)",
                               "loglik": ll}
        # the oracle's state draw with 4 categories: per-component site and pair
        # log-likelihoods, then the draws
        t0 = time.perf_counter()
        sitec = np.stack([oracle.site_loglik(ds, [c]) for c in model.site_components(p4)])
        pc4 = model.pair_components(p4)
        pairc = np.full((len(pc4), ds.numcols), -np.inf)
        for k, (P, pi, _) in enumerate(pc4):
            Mk = oracle.pair_loglik(ds, [(P, pi, 0.0)], mask=mask)
            pairc[k, x] = Mk[x, paired[x] - 1]
        oracle.sample_states(sitec, pairc, [c[2] for c in pc4], paired, 4, 1.0, seed=0)
        out["cpu_baseline_states"] = {"value": 1.0 / (time.perf_counter() - t0), "unit": "draws/s", "cores": 1,
                                      "kind": "port", "sample": "one full state draw, 4 site categories"}
    print(json.dumps(out), flush=True)


def main_mcmc(args):
    """Learned proposal vs adaptive random walk on the 50 x 1500 alignment:
    16 chains, adaptive random-walk burn-in, network trained on the device on the
    second half of the burn-in, then learned / random-walk / long random-walk
    chains from the same states.  ESS/s counts each chain's wall time (host
    model construction, device evaluation, network inference)."""
    from messi_amd import lib, mcmc, synth

    a = synth.config2()
    ds = synth.dataset_of(a)
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    post = mcmc.Posterior(ds, a.params, args.chains, prec=args.prec)
    net = lib.Mlp(mcmc.DIM, 128, seed=args.seed)
    th0 = post.theta_of(a.params) + rng.normal(0, 0.05, (args.chains, mcmc.DIM))

    def log(*m):
        print("[mcmc]", *m, f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

    res = mcmc.run_experiment(post, th0, net, rng, burn=args.burn, steps=args.mcmc_steps,
                              long_steps=args.long_steps, train_steps=args.train_steps, log=log,
                              mode_logp=lambda th: post(th, prec="fp64"), ind_prob=args.ind_prob,
                              mix_components=args.mix_components, mix_flatten=args.mix_flatten,
                              mix_inflate=args.mix_inflate)
    tr = res.pop("traces")
    if args.save_traces:
        np.savez_compressed(args.save_traces, learned=tr["learned"].samples, learned_kind=tr["learned"].kind,
                            learned_acc=tr["learned"].accepted, random_walk=tr["random_walk"].samples,
                            long=tr["long"].samples[::10], burn=tr["burn"].samples)
    z = np.abs(np.array(res["mean_z_learned_vs_long"]))
    z_ess = np.abs(np.array(res["mean_z_learned_vs_long_ess_only"]))
    z_rw = np.abs(np.array(res["mean_z_rw_vs_long"]))
    mr = res.get("mixture_no_network")
    z_mix = np.abs(np.array(mr["mean_z_vs_long"])) if mr else None
    out = {"metric": "MCMC ESS/s, learned proposal vs adaptive random walk (min ESS over 12 parameters)",
           "value": res["ess_per_s_gain_min"], "unit": "x (learned ESS/s / random-walk ESS/s)", "n_gpus": 1,
           "higher_is_better": True, "dtype": args.prec.replace("fp", "f"), "data": "synthetic",
           "config": {"workload": f"50 sequences x 1500 nt, 12 parameters, {args.chains} chains, burn-in "
                                  f"{args.burn}, {args.mcmc_steps} production steps for each of learned, mixture "
                                  f"without network and random walk, long chain {args.long_steps}, network "
                                  f"12-128-128-24 trained {args.train_steps} Adam steps", "seed": args.seed},
           "ess_per_s_gain_min_incl_burn_and_training": res["ess_per_s_gain_min_incl_burn"],
           "ess_per_s_gain_median": res["ess_per_s_gain_median"],
           "network_gain_over_mixture_min": res.get("network_gain_over_mixture_min"),
           "network_gain_over_mixture_median": res.get("network_gain_over_mixture_median"),
           "mixture_gain_over_rw_min": res.get("mixture_gain_over_rw_min"),
           "means_within_2_mcse": int(np.sum(z <= 2.0)), "means_within_3_mcse": int(np.sum(z <= 3.0)),
           "means_within_2_mcse_ess_only": int(np.sum(z_ess <= 2.0)),
           "means_within_2_mcse_random_walk": int(np.sum(z_rw <= 2.0)),
           "means_within_2_mcse_mixture_no_network": int(np.sum(z_mix <= 2.0)) if mr else None,
           "max_split_rhat": {k: float(np.max(v)) for k, v in res["split_rhat"].items()},
           "evals": post.evals, "device_eval_seconds": post.eval_seconds,
           "param_names": list(mcmc.NAMES), "result": res}
    net.close()
    post.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--groups", type=int, default=0,
                    help="member groups of the pipelined batch evaluation (1..4, 0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--instances", action="store_true")
    ap.add_argument("--mcmc", action="store_true")
    ap.add_argument("--genome", action="store_true")
    ap.add_argument("--mea", action="store_true")
    ap.add_argument("--sample", action="store_true")
    ap.add_argument("--structure", action="store_true")
    ap.add_argument("--siterate", action="store_true", help="time the site-rate posteriors")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 256 MiB device copy (reference for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--mea-lengths", type=int, nargs="+", default=[1500, 9000])
    ap.add_argument("--tile", type=int, default=0, help="inside tile edge 16 or 32 (default: library's)")
    ap.add_argument("--window", type=int, default=0, help="inside bulk window 1, 2 or 4 (default: library's)")
    ap.add_argument("--single-launch", type=int, default=-1,
                    help="1: inside pass as one launch, 0: one launch per tile-diagonal (default: library's)")
    ap.add_argument("--precs", nargs="+", default=["fp64", "fp32"], help="precisions to time (profiling runs)")
    ap.add_argument("--no-posterior", action="store_true", help="skip the posterior error check (profiling runs)")
    ap.add_argument("--genome-spans", type=int, nargs="+", default=[1000000, 500])
    ap.add_argument("--chains", type=int, default=16)
    ap.add_argument("--burn", type=int, default=2000)
    ap.add_argument("--mcmc-steps", type=int, default=20000)
    ap.add_argument("--ind-prob", type=float, default=0.25)
    ap.add_argument("--save-traces", default="", help="write the chains to this .npz (MCMC mode)")
    ap.add_argument("--mix-components", type=int, default=6)
    ap.add_argument("--mix-flatten", type=float, default=0.5)
    ap.add_argument("--mix-inflate", type=float, default=1.3)
    ap.add_argument("--long-steps", type=int, default=30000)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--prec", default="fp32")
    ap.add_argument("--site-cats", type=int, default=1, help="site-rate categories of the model (K = 1 default)")
    ap.add_argument("--lambda-cats", type=int, default=1, help="lambda categories of the model")
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the likelihood path is "
              f"single-GPU in this build", file=sys.stderr)
        sys.exit(2)
    if args.tile:
        from messi_amd import lib
        lib.context().set_tile(args.tile)
    if args.window:
        from messi_amd import lib
        lib.context().set_bulk_window(args.window)
    if args.single_launch >= 0:
        from messi_amd import lib
        lib.context().set_single_launch(bool(args.single_launch))
    if args.instances:
        return main_instances(args)
    if args.mcmc:
        return main_mcmc(args)
    if args.genome:
        return main_genome(args)
    if args.mea:
        return main_mea(args)
    if args.sample:
        return main_sample(args)
    if args.structure:
        return main_structure(args)
    if args.copy_calibration:
        return main_copy_calibration(args)
    if args.siterate:
        return main_siterate(args)

    from messi_amd import lib, synth
    from messi_amd.hotpath import device_dataset

    a = synth.config2()
    ds = synth.dataset_of(a)
    L = ds.numcols
    lib.context()
    dd = device_dataset(ds)
    stats = dd.stats()
    ps = _perturbed(a, ds, args.batch, siteCats=args.site_cats, lambdaCats=args.lambda_cats)
    comps = [_components(p) for p in ps]
    sites, pairs = [c[0] for c in comps], [c[1] for c in comps]
    bB = lib.Batch([dd] * args.batch)
    bB.set_groups(args.groups)
    b1 = lib.Batch([dd])
    from messi_amd import mcmc
    pm = mcmc.ParamMap(a.params, ds.tree.branchlengths)
    thetas = pm.theta_of(a.params) + np.random.default_rng(1).normal(0, 0.02, (args.batch, mcmc.DIM))
    res = {}
    tune = not (args.tile or args.window or args.groups)
    for prec in args.precs:
        ms1, z1, ph1, ms1h = time_batch(b1, sites[:1], pairs[:1], prec, args.steps, args.warmup, tune)
        msB, zB, phB, msBh = time_batch(bB, sites, pairs, prec, args.steps, args.warmup, tune)
        # from parameter vectors: the 12-parameter model (one site and one lambda category)
        k1 = args.site_cats == 1 and args.lambda_cats == 1
        ms1p = time_from_params(b1, pm, thetas[:1], prec, args.steps, args.warmup)[0] if k1 else float("nan")
        msBp = time_from_params(bB, pm, thetas, prec, args.steps, args.warmup)[0] if k1 else float("nan")
        res[prec] = {"batch1": {"ms_per_eval": ms1, "evals_per_s": 1e3 / ms1, "logZ": float(z1[0]),
                                "evals_per_s_host_inputs": 1e3 / ms1h,
                                "evals_per_s_from_params": 1e3 / ms1p,
                                "phase_ms": {"site": ph1[0], "pair": ph1[1], "inside": ph1[2]},
                                "schedule": b1.schedule, "inside_step_launch_us": b1.step_launch_us},
                     f"batch{args.batch}": {"ms_per_step": msB, "evals_per_s": args.batch * 1e3 / msB,
                                            "evals_per_s_host_inputs": args.batch * 1e3 / msBh,
                                            "evals_per_s_from_params": args.batch * 1e3 / msBp,
                                            "ms_per_step_from_params": msBp,
                                            "phase_ms": {"site": phB[0], "pair": phB[1], "inside": phB[2]},
                                            "inside_step_launch_us": bB.step_launch_us,
                                            "inside_step_launches": bB.step_launches,
                                            "schedule": bB.schedule},
                     "logZ_batch": zB.tolist()}
    T_in = (L ** 3 - L) / 6
    lead = "fp32" if "fp32" in res else args.precs[0]
    r = res[lead][f"batch{args.batch}"]
    out = {
        "metric": "likelihood evals/s at 1500nt x 50seq, FP32 vs FP64",
        "value": args.batch * 1e3 / r["ms_per_step"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": lead.replace("fp", "f"),
        "data": "synthetic (seeded Yule tree, GTR/MuSE simulation, 40% paired, 5% gaps)",
        "config": {"workload": f"config2: 50 sequences x 1500 nt, {args.site_cats} site-rate and "
                               f"{args.lambda_cats} lambda categories (K4={args.site_cats}, "
                               f"K16={args.site_cats ** 2 * args.lambda_cats}), full-span pairs and inside; "
                               f"{args.batch} parameter vectors per step (one per chain)",
                   "L": L, "nseq": ds.numseqs, "batch": args.batch, "sum_U": stats["sum_U"],
                   "root_entries": stats["root_entries"],
                   "member_groups": args.groups or "automatic",
                   "bulk_window": args.window or 4},
        "variants": res,
    }
    if lead == "fp32":
        # algorithmic flops per launch (2 per bifurcation term, the batch's terms spread over
        # the step launches) over the step kernel's mean launch duration (HIP events)
        nl = r.get("inside_step_launches") or [0, 0]
        TE = 32 if nl[1] and not nl[0] else 16
        sch = r.get("schedule") or {}
        W = sch.get("window") or args.window or 4
        NT = int(round(sum(nl))) or -(-L // TE)
        flops_launch = args.batch * 2 * T_in / NT
        launch_us = r.get("inside_step_launch_us") or r["phase_ms"]["inside"] * 1e3 / NT
        achieved = flops_launch / (launch_us * 1e-6) / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp32"], "unit": "TFLOP/s",
                           "frac": achieved / PEAK["fp32"], "traffic": None,
                           "kernel": f"inside_step_kernel<XF,{TE},{W}> (batch of {args.batch}: {flops_launch:.4g} "
                                     f"flops per launch = 2 per bifurcation term, {int(T_in)} terms per evaluation, "
                                     f"{NT} launches; mean launch {launch_us:.1f} us)",
                           "inside_phase_tflops": args.batch * 2 * T_in / (r["phase_ms"]["inside"] * 1e-3) / 1e12}
        # HBM bytes per launch measured by the FETCH_SIZE / WRITE_SIZE passes of this same workload
        # (copy-calibrated); only quoted for the shape the counters were taken on
        tf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                          "inside_xf16_config2_batch16_traffic.json")
        if args.batch == 16 and TE == 16 and W == 4 and os.path.exists(tf):
            with open(tf) as f:
                tj = json.load(f)
            out["roofline"]["traffic"] = tj["hbm_bytes_per_launch"]
            out["roofline"]["traffic_note"] = ("bytes per launch, mean over the 94 launches of one batch-16 "
                                               "evaluation, from profiles/" + os.path.basename(tf))
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = {k: res["fp32"][k]["evals_per_s"] / res["fp64"][k]["evals_per_s"]
                                 for k in ("batch1", f"batch{args.batch}")}
        out["fp32_logZ_rel_err"] = (abs(res["fp32"]["batch1"]["logZ"] - res["fp64"]["batch1"]["logZ"])
                                    / abs(res["fp64"]["batch1"]["logZ"]))
    if not args.no_posterior:
        # posterior (outside + base-pair probabilities) of the first parameter vector in both paths
        post = {prec: dd.posterior(sites[0], pairs[0], prec) for prec in ("fp64", "fp32")}
        (z64, u64, p64), (z32, u32, p32) = post["fp64"], post["fp32"]
        out["fp32_posterior_logZ_rel_err"] = abs(z32 - z64) / abs(z64)
        out["fp32_bpp_max_abs_err"] = float(max(np.max(np.abs(u32 - u64)), np.max(np.abs(p32 - p64))))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ds, ps[0], L)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
