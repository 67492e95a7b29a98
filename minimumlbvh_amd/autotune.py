"""Autotuner for the launch shapes of libmlbvh (sort tile, persistent-trace
grid and chunk).  Results never depend on these choices; only time does.

tune() times every radix-sort tile shape on a sort of representative size
and a grid of persistent-trace shapes on a batched trace, with HIP events on
the current stream, keeps the fastest, and applies it to the context
(mlbvh_set_tuning).  The choice can be cached in a JSON file keyed by device
name.
"""
from __future__ import annotations

import json
import os
from typing import Optional

import numpy as np
import torch

from . import lbvh

SORT_CONFIGS = [(256, 16), (256, 32), (512, 16), (512, 24), (1024, 16)]
TRACE_SHAPES = [(bpc, chunk) for bpc in (4, 8, 16) for chunk in (64, 128, 256)]


def _time(fn, reps=5, warmup=2):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        times.append(s.elapsed_time(e))
    return float(np.median(times))


def get_tuning(device: int = 0):
    import ctypes

    ctx = lbvh.context(device)
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lbvh.check(ctx.lib.mlbvh_get_tuning(ctx.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return {"sort_config": a.value, "trace_blocks_per_cu": b.value, "trace_chunk": c.value}


def tune_sort(n: int = 1 << 24, key_bits: int = 64, device: int = 0, reps: int = 5):
    """Returns (best config index, {config: ms})."""
    cur = get_tuning(device)
    dev = f"cuda:{device}"
    g = torch.Generator(device=dev).manual_seed(1)
    dtype = torch.int64 if key_bits == 64 else torch.int32
    lo, hi = (-(2**63), 2**63 - 1) if key_bits == 64 else (-(2**31), 2**31 - 1)
    src = torch.randint(lo, hi, (n,), dtype=dtype, device=dev, generator=g)
    keys, tmp = torch.empty_like(src), torch.empty_like(src)
    sorter = lbvh.OnesweepSort(device)
    res = {}
    for cfg in range(len(SORT_CONFIGS)):
        lbvh.set_tuning(cfg, cur["trace_blocks_per_cu"], cur["trace_chunk"], device)

        def run():
            keys.copy_(src)
            sorter.sort(keys, None, tmp, None, n, 0, key_bits)

        copy_ms = _time(lambda: keys.copy_(src), reps)
        res[cfg] = max(_time(run, reps) - copy_ms, 1e-6)
    best = min(res, key=res.get)
    lbvh.set_tuning(best, cur["trace_blocks_per_cu"], cur["trace_chunk"], device)
    return best, res


def tune_trace(bvh: "lbvh.BatchedBVH", triangles, rays, device: int = 0, reps: int = 3):
    """Times the persistent batched trace over TRACE_SHAPES; returns (best, {shape: ms})."""
    cur = get_tuning(device)
    hits = torch.empty((rays.shape[0], 8), dtype=torch.int32, device=rays.device)
    res = {}
    for bpc, chunk in TRACE_SHAPES:
        lbvh.set_tuning(cur["sort_config"], bpc, chunk, device)
        res[(bpc, chunk)] = _time(lambda: bvh.trace(triangles, rays, hits=hits), reps, 1)
    best = min(res, key=res.get)
    lbvh.set_tuning(cur["sort_config"], best[0], best[1], device)
    return best, res


def tune(device: int = 0, cache_file: Optional[str] = None, sort_n: int = 1 << 24) -> dict:
    """Tune the sort tile (and load/save the choice from cache_file)."""
    name = torch.cuda.get_device_name(device)
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            data = json.load(f)
        if data.get("device") == name:
            lbvh.set_tuning(data["sort_config"], data["trace_blocks_per_cu"], data["trace_chunk"], device)
            return data
    best, res = tune_sort(sort_n, 64, device)
    out = dict(get_tuning(device))
    out["device"] = name
    out["sort_ms"] = {str(k): round(v, 4) for k, v in res.items()}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(out, f, indent=1)
    return out
