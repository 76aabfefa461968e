"""Schedule autotuner for batched evaluation.

Three choices of a batch's schedule are tuned by timing the batch itself, one
after the other (coordinate search), every result being bit-for-bit the same
whatever is chosen:

* the inside tile edge per member (16 or 32): members of one edge share one
  launch sequence of tile-diagonals.  Short or one-off instances are latency
  bound (the in-tile wavefront of 2T-1 dependent steps runs once per launch);
  long ones and, in the scaled path, single evaluations can prefer T = 32 (half
  the launches, 8x the work per split-point product).  Candidates are "T = 32
  from length X on" over the distinct member lengths, plus all-16 and all-32;
* the bulk window (1, 2 or 4 launches per accumulator visit), for the batch;
* the member groups of a pipelined evaluation (1 .. 4 streams), for batches of
  two or more members.

Results are cached in a JSON table keyed by the batch's length profile and
precision, so a later batch of the same shape reuses them.

Usage:
    b = lib.Batch(members); b.load(sites, pairs, "fp32")
    choice = autotune.tune_batch(b, "fp32")      # sets b's schedule, returns timings
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


def apply(batch, lengths, choice: dict) -> None:
    """Set a batch's schedule from a choice of tune_batch."""
    batch.set_tiles(tiles_for(lengths, choice["threshold"]))
    batch.set_window(int(choice.get("window", 0)))
    batch.set_groups(int(choice.get("groups", 0)))


def tune_batch(batch, prec, reps: int = 3, table: Optional[Dict[str, dict]] = None,
               max_candidates: int = 8, path: Optional[str] = None, windows=(4, 2, 1),
               groups=(0, 1, 2, 3, 4)) -> dict:
    """Time the candidate schedules on the loaded batch (tile assignments, then
    bulk windows, then member groups, each at the best of the ones before), set
    the fastest and return {"threshold", "window", "groups", "ms", "candidates":
    [(threshold, ms), ...], "window_candidates": [(window, ms), ...],
    "group_candidates": [(groups, ms), ...]}.  With a table (dict), a cached
    choice for the same length profile is applied without timing, and a new
    choice is stored in it (and written to path if given)."""
    lengths = [m.L for m in batch.members]
    key = _key(lengths, prec)
    if table is not None and key in table:
        ch = table[key]
        apply(batch, lengths, ch)
        return dict(ch, cached=True)
    distinct = sorted(set(lengths))
    if len(distinct) > max_candidates - 2:
        qs = np.quantile(distinct, np.linspace(0, 1, max_candidates - 2)[1:-1])
        inner = sorted(set(int(min(d for d in distinct if d >= q)) for q in qs))
    else:
        inner = distinct[1:]
    cands = [None, distinct[0]] + [x for x in inner if x != distinct[0]]
    batch.set_window(0)
    batch.set_groups(0)
    results = []
    for thr in cands:
        batch.set_tiles(tiles_for(lengths, thr))
        results.append((thr, _time(batch, prec, reps)))
    best = min(results, key=lambda r: r[1])
    batch.set_tiles(tiles_for(lengths, best[0]))
    wres = []
    for w in windows:
        batch.set_window(w)
        wres.append((w, _time(batch, prec, reps)))
    bw = min(wres, key=lambda r: r[1])
    batch.set_window(bw[0])
    gres = []
    n = len(lengths)
    for g in groups:
        if g > n:
            continue
        batch.set_groups(g)
        gres.append((g, _time(batch, prec, reps)))
    bg = min(gres, key=lambda r: r[1]) if gres else (0, bw[1])
    batch.set_groups(bg[0])
    ch = {"threshold": best[0], "window": bw[0], "groups": bg[0], "ms": bg[1], "candidates": results,
          "window_candidates": wres, "group_candidates": gres}
    if table is not None:
        table[key] = {"threshold": best[0], "window": bw[0], "groups": bg[0], "ms": bg[1]}
        if path:
            save_table(table, path)
    return dict(ch, cached=False)
