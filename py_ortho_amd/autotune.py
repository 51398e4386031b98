"""Autotuner of the batched plane sweep: for every image-size class of a block, the tile shape
(32 x 32, 64 x 16 or 64 x 8: LDS footprint and registers against halo overhead) and the number of
hypothesis chunks per view (parallelism against merge traffic), chosen by timing each candidate on
the device.

Each candidate runs the sweep of up to `views_per_class` views of the class over the first `planes`
hypotheses; the fastest per pixel-hypothesis wins.  Each chunk's set-up is spread over its planes,
so the tuning sweep must be long enough for the chunk counts it compares: over 16 planes 600x800
views ranked 2 chunks first, over 48 planes 12 and over 96 planes 16 (4 % faster per
pixel-hypothesis than 6; 900x1200 views 12 chunks, 1200x1600 and 1800x2400 8, the last two within
0.5 % of each other from 4 to 16).  Results can be kept in a JSON file keyed by class, sources,
window and plane count.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional, Sequence, Tuple

import torch

from . import batch

CANDIDATES: Tuple[Tuple[int, int], ...] = ((32, 1), (32, 2), (32, 4), (32, 6), (32, 8), (32, 12), (32, 16),
                                           (16, 1), (16, 2), (8, 1))


def _key(cls, s_max, r, k, planes) -> str:
    return f"{cls[0]}x{cls[1]}|S{s_max}|r{r}|k{k}|D{planes}"


def _time_batch(B: batch.Batch, reps: int) -> float:
    B.run()  # warm-up: code objects, caches
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        s.record()
        B.run()
        e.record()
        e.synchronize()
        best = min(best, s.elapsed_time(e))
    return best


def tune_block(blk, candidates: Sequence[Tuple[int, int]] = CANDIDATES, planes: int = 96, views_per_class: int = 2,
               reps: int = 2, cache_path: Optional[str] = None):
    """-> ({(H, W): (tile_h, chunks)}, {(H, W): {(tile_h, chunks): ns per pixel-hypothesis}})."""
    o = blk.opt
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    classes: Dict[Tuple[int, int], list] = {}
    for k in range(blk.n):
        classes.setdefault(blk.size_class(k), []).append(k)
    choice, table = {}, {}
    D = min(planes, o.num_depths)
    for cls, ks in sorted(classes.items()):
        ks = ks[:views_per_class]
        key = _key(cls, max(len(blk.nbrs[k]) for k in ks), o.window, o.top_k, o.num_depths)
        if key in cache:
            choice[cls] = tuple(cache[key])
            continue
        px_h = sum(blk.pixels()[k] for k in ks) * D
        scratch = [torch.empty_like(blk.raw[k]) for k in ks]
        table[cls] = {}
        for th, ch in candidates:
            B = batch.Batch()
            try:
                B.set_sweep([blk.sweep_view(k, th, ch, depth=scratch[i], planes=blk.planes[k][:D].contiguous())
                             for i, k in enumerate(ks)], o.window, o.top_k)
                table[cls][(th, ch)] = _time_batch(B, reps) * 1e6 / px_h
            finally:
                B.close()
        choice[cls] = min(table[cls], key=table[cls].get)
        cache[key] = list(choice[cls])
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return choice, table
