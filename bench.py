"""Benchmark: full-likelihood evaluations per second at 1,500 nt x 50 sequences.

One step = one full-likelihood evaluation through the C ABI (messi_eval): site
log-likelihoods (4-state pruning), the all-pairs 16-state pair matrix and the
KH99 inside partition function, with the alignment and dedup lists resident in
HBM and the per-branch transition matrices passed in from the host.  Both the
plain FP64 path and the scaled FP32 path are timed; ``value`` is the FP32 path
(the FP64 figure and the FP32/FP64 ratio are reported beside it).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def _components(ds, params):
    from messi_amd import model
    from messi_amd.hotpath import _stack
    return _stack(model.site_components(params)), _stack(model.pair_components(params))


def cpu_baseline(ds, params, L):
    """Oracle (single-threaded FP64 C restatement of the reference CPU path) on a
    bounded sample: all site columns, the column pairs within distance 100
    (scaled to all pairs by count), and the full inside recursion."""
    import oracle
    from messi_amd import model
    t0 = time.perf_counter()
    u = oracle.site_loglik(ds, model.site_components(params))
    t1 = time.perf_counter()
    D = 100
    M = oracle.pair_loglik(ds, model.pair_components(params), maxbp=D)
    t2 = time.perf_counter()
    npairs_sample = sum(min(D, L - 1 - x) for x in range(L))
    npairs = L * (L - 1) // 2
    pair_full = (t2 - t1) * npairs / npairs_sample
    M = np.where(np.isfinite(M), M, -np.inf)
    t3 = time.perf_counter()
    oracle.inside(u, M)
    t4 = time.perf_counter()
    per_eval = (t1 - t0) + pair_full + (t4 - t3)
    return {"value": 1.0 / per_eval, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"all {L} site columns, {npairs_sample} of {npairs} column pairs (|x-y|<=100, "
                      f"time scaled by pair count), full inside at L={L}; measured "
                      f"{(t4 - t0):.1f} s of CPU time"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prec", default="both", choices=["both", "fp64", "fp32"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the likelihood path is "
              f"single-GPU in this build", file=sys.stderr)
        sys.exit(2)

    from messi_amd import lib, model, synth
    from messi_amd.hotpath import device_dataset

    a = synth.config2()
    ds = synth.dataset_of(a)
    params = model.getparams(a.params, ds.tree.branchlengths)
    L = ds.numcols
    ctx = lib.context()
    dd = device_dataset(ds)
    stats = dd.stats()
    site, pair = _components(ds, params)
    precs = ["fp64", "fp32"] if args.prec == "both" else [args.prec]
    res = {}
    for prec in precs:
        ctx.set_profiling(False)
        for _ in range(args.warmup):
            z = dd.eval(site, pair, 1.0, prec)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            z = dd.eval(site, pair, 1.0, prec)
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        # Phase device times from HIP events on the library's stream.
        ctx.set_profiling(True)
        ph = np.zeros(5)
        nprof = max(3, min(10, args.steps))
        for _ in range(nprof):
            dd.eval(site, pair, 1.0, prec)
            ph += dd.timings()
        ph /= nprof
        ctx.set_profiling(False)
        res[prec] = {"ms_per_step": ms, "evals_per_s": 1e3 / ms, "logZ": z,
                     "phase_ms": {"site": ph[0], "pair": ph[1], "inside": ph[2]}}
    T_in = (L ** 3 - L) / 6
    main_prec = "fp32" if "fp32" in res else "fp64"
    r = res[main_prec]
    inside_ms = r["phase_ms"]["inside"]
    achieved = 2 * T_in / (inside_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK[main_prec], "unit": "TFLOP/s",
            "frac": achieved / PEAK[main_prec], "traffic": None,
            "kernel": "inside_diag_kernel (inside phase, 2 flops per bifurcation term, "
                      f"{int(T_in)} terms per eval)"}
    out = {
        "metric": "likelihood evals/s at 1500nt x 50seq, FP32 vs FP64",
        "value": r["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if main_prec == "fp32" else "f64",
        "data": "synthetic (seeded Yule tree, GTR/MuSE simulation, 40% paired, 5% gaps)",
        "config": {"workload": "config2: 50 sequences x 1500 nt, K=1, full-span pairs and inside",
                   "L": L, "nseq": ds.numseqs, "sum_U": stats["sum_U"], "root_entries": stats["root_entries"]},
        "roofline": roof,
        "variants": res,
    }
    if "fp64" in res and "fp32" in res:
        out["fp32_over_fp64"] = res["fp32"]["evals_per_s"] / res["fp64"]["evals_per_s"]
        out["fp32_logZ_rel_err"] = abs(res["fp32"]["logZ"] - res["fp64"]["logZ"]) / abs(res["fp64"]["logZ"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ds, params, L)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
