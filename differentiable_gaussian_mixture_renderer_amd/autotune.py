"""Autotuner for the batched splat: picks the tile shape {8x8, 16x16, 32x8}
and the number of Gaussians staged in LDS per round {64, 128, 256} by timing
forward+backward on the actual workload.

The LDS batch size does not change results; the tile shape does (binning is
per tile, and a Gaussian still reaches every pixel of a tile its rectangle
touches), so accuracy checks compare against the oracle run with the chosen
tile shape.  Results are cached per workload signature in a JSON file.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Sequence, Tuple

import torch

from .batching import BatchSplatter

TILES = ((8, 8), (16, 16), (32, 8))
LDS_BATCHES = (64, 128, 256)
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "dgmr_mi355x", "autotune.json")


def signature(n: int, cameras: Sequence) -> str:
    res = sorted((c.width, c.height) for c in cameras)
    return f"n={n};views={len(cameras)};px={sum(w * h for w, h in res)};res={res[0]}-{res[-1]}"


def _time(bs: BatchSplatter, params, cameras, grads, reps: int) -> float:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        bs.forward(*params, cameras)
        bs.backward(grads)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def autotune(params, cameras: Sequence, reps: int = 2, use_cache: bool = True) -> Tuple[Tuple[Tuple[int, int], int], Dict[str, float]]:
    """Returns ((tile, lds_batch), {"WxH/B": ms}) for forward+backward of the batch."""
    sig = signature(params[2].shape[0], cameras)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if sig in cache:
            e = cache[sig]
            return ((tuple(e["tile"]), e["batch"]), e["table"])
    dev = params[2].device
    grads = [torch.full((3, c.height, c.width), 1.0 / (3 * c.height * c.width), device=dev) for c in cameras]
    table = {}
    best = None
    for tile in TILES:
        for b in LDS_BATCHES:
            bs = BatchSplatter(tile=tile, batch=b, device=dev)
            bs.forward(*params, cameras)  # warm-up, sizes the workspace
            bs.backward(grads)
            ms = _time(bs, params, cameras, grads, reps)
            table[f"{tile[0]}x{tile[1]}/{b}"] = round(ms, 3)
            if best is None or ms < best[0]:
                best = (ms, tile, b)
            del bs
    choice = (best[1], best[2])
    if use_cache:
        cache[sig] = {"tile": list(best[1]), "batch": best[2], "table": table}
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return choice, table
