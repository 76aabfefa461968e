"""Schedule autotuner for batched evaluation (the inside tile edge per member).

The inside schedule runs members of one tile edge (16 or 32) in one launch
sequence of tile-diagonals.  Which edge is faster depends on a member's length
and on what else shares the launches: short instances are latency-bound (few
tiles per diagonal; the shorter 2T-1 step wavefront of T = 16 wins), long ones
can prefer T = 32 (half the launches, 8x the work per split-point product).
The tuner times the batch itself with every candidate assignment "T = 32 from
length X on" (X over the distinct member lengths, plus all-16 and all-32) and
keeps the fastest; results are cached in a JSON table keyed by the batch's
length profile and precision, so a later batch of the same shape reuses them.

Usage:
    b = lib.Batch(members); b.load(sites, pairs, "fp32")
    choice = autotune.tune_batch(b, "fp32")      # sets b's tiles, returns timings
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional

import numpy as np

from . import lib

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")


def _key(lengths, prec) -> str:
    Ls = sorted(int(x) for x in lengths)
    return f"{lib.prec_code(prec)}:{len(Ls)}:{Ls[0]}-{Ls[-1]}:{int(np.sum(Ls))}"


def load_table(path: str = TABLE) -> Dict[str, dict]:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_table(table: Dict[str, dict], path: str = TABLE) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tiles_for(lengths, threshold: Optional[int]):
    """T = 32 for members of length >= threshold (None: all 16)."""
    return [32 if (threshold is not None and L >= threshold) else 16 for L in lengths]


def _time(batch, prec, reps: int) -> float:
    batch.eval_loaded(1.0, prec)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        batch.eval_loaded(1.0, prec)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e3


def tune_batch(batch, prec, reps: int = 3, table: Optional[Dict[str, dict]] = None,
               max_candidates: int = 8, path: Optional[str] = None) -> dict:
    """Time the candidate tile assignments on the loaded batch, set the fastest
    and return {"threshold", "ms", "candidates": [(threshold, ms), ...]}.  With a
    table (dict), a cached choice for the same length profile is applied without
    timing, and a new choice is stored in it (and written to path if given)."""
    lengths = [m.L for m in batch.members]
    key = _key(lengths, prec)
    if table is not None and key in table:
        ch = table[key]
        batch.set_tiles(tiles_for(lengths, ch["threshold"]))
        return dict(ch, cached=True)
    distinct = sorted(set(lengths))
    if len(distinct) > max_candidates - 2:
        qs = np.quantile(distinct, np.linspace(0, 1, max_candidates - 2)[1:-1])
        inner = sorted(set(int(min(d for d in distinct if d >= q)) for q in qs))
    else:
        inner = distinct[1:]
    cands = [None, distinct[0]] + [x for x in inner if x != distinct[0]]
    results = []
    for thr in cands:
        batch.set_tiles(tiles_for(lengths, thr))
        results.append((thr, _time(batch, prec, reps)))
    best = min(results, key=lambda r: r[1])
    batch.set_tiles(tiles_for(lengths, best[0]))
    ch = {"threshold": best[0], "ms": best[1], "candidates": results}
    if table is not None:
        table[key] = {"threshold": best[0], "ms": best[1]}
        if path:
            save_table(table, path)
    return dict(ch, cached=False)
