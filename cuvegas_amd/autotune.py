"""Autotuner for the batched sampling launch: picks the workgroup size and the
samples-per-lane run length by timing calibration iterations of the actual
batch, and caches the choice per (integrand, item shapes, precision, device).

The candidates span the trade-offs of the kernel: larger workgroups amortise
the per-task LDS bin flush over more windows but raise the per-task minimum
work (small items waste lanes); longer runs amortise the cross-lane scan and
the cube walk but coarsen the window granularity.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
import uuid
from typing import Dict, Sequence, Tuple

from .batch import Batch, BatchProgram

DEFAULT_CANDIDATES: Tuple[Tuple[int, int, int], ...] = (
    (256, 16, 4), (256, 16, 8), (512, 16, 2), (512, 32, 2), (256, 32, 4), (1024, 16, 2))

_PROGRAMS: Dict[tuple, BatchProgram] = {}


def _cache_path() -> str:
    base = os.environ.get("CUVEGAS_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "cuvegas_amd"))
    return os.path.join(base, "autotune.json")


def _signature(source, items, precision, n_intervals, max_dim, device) -> str:
    h = hashlib.sha256()
    h.update(source.encode())
    # every field of an item: dimension, budget, parameter offset and an
    # explicit n_strat all change the cube count, the windows or the cost
    h.update(json.dumps([[int(x) for x in it] for it in items]).encode())
    h.update(f"{precision}|{n_intervals}|{max_dim}|{device}".encode())
    return h.hexdigest()[:32]


def program(source: str, max_dim: int, n_intervals: int, precision: str, block: int,
            run_len: int, task_windows: int = 8) -> BatchProgram:
    key = (source, max_dim, n_intervals, precision, block, run_len, task_windows)
    if key not in _PROGRAMS:
        _PROGRAMS[key] = BatchProgram(source, max_dim, n_intervals, precision, run_len=run_len,
                                      options=f"block={block} task_windows={task_windows}")
    return _PROGRAMS[key]


def autotune_batch(source: str, items: Sequence, theta: Sequence[float], *, max_dim: int = 12,
                   n_intervals: int = 256, precision: str = "fp32", iters: int = 2,
                   candidates=DEFAULT_CANDIDATES, use_cache: bool = True, device: str = "",
                   verbose: bool = False) -> dict:
    """Returns {"block", "run_len", "task_windows", "ms_per_iter", "table"} for the
    fastest candidate (block size, samples per lane run, windows per wave per task)."""
    sig = _signature(source, items, precision, n_intervals, max_dim, device)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if sig in cache:
            return dict(cache[sig], cached=True)
    table = []
    for block, run_len, task_windows in candidates:
        prog = program(source, max_dim, n_intervals, precision, block, run_len, task_windows)
        b = Batch(prog, items, theta, max_it=iters, skip=0, seed=0)
        t0 = time.perf_counter()
        b.run(iters, use_graph=True)
        r = b.results()
        wall = time.perf_counter() - t0
        b.close()
        row = {"block": block, "run_len": run_len, "task_windows": task_windows,
               "ms_per_iter": r["loop_ms"] / iters,
               "evals_per_s": float(r["nevals"].sum()) / (r["loop_ms"] * 1e-3), "wall_s": wall}
        table.append(row)
        if verbose:
            print("[autotune]", row, file=sys.stderr, flush=True)
    best = min(table, key=lambda r: r["ms_per_iter"])
    out = {"block": best["block"], "run_len": best["run_len"], "task_windows": best["task_windows"],
           "ms_per_iter": best["ms_per_iter"], "table": table}
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            cache[sig] = out
            # whole-file replace: readers never see a partial file; the name is
            # unique per call, so threads of one process do not share it
            tmp = f"{path}.{os.getpid()}.{uuid.uuid4().hex}.tmp"
            with open(tmp, "w") as f:
                json.dump(cache, f)
            os.replace(tmp, path)
        except OSError:
            pass
    return dict(out, cached=False)
