"""Many views' depth maps in one launch: the batched plane sweep and PatchMatch (om_batch).

The reference computes depth maps one view per host thread (densify.py:596).  Here a batch plans
all views once -- descriptor tables and workspaces in HBM -- and then runs every view's sweep in one
persistent work-queue launch over (view, tile, hypothesis-chunk) items, followed by the PatchMatch
passes of all views (4 launches per colour pass for the whole batch).  The whole sequence can be
captured in a HIP graph and replayed.  Results equal the per-view calls (dense.sweep,
dense.sweep_finalize, dense.pm_refine) bit for bit.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib, dense
from ._lib import OmPmView, OmSweepDesc, check


@dataclass
class SweepView:
    """One view of a batched sweep (arguments of dense.sweep)."""
    ref: torch.Tensor
    mu: torch.Tensor
    sd: torch.Tensor
    srcs: Sequence[torch.Tensor]
    ray: torch.Tensor
    geo: np.ndarray
    hyp: dict
    state: Optional[dense.SweepState] = None
    depth: Optional[torch.Tensor] = None      # receives the parabola depth when given
    tile_h: int = 0                           # 8, 16 or 32 rows (0 = 32) unless the batch sets one for all
    chunks: int = 0                           # hypothesis chunks (0 = auto) unless the batch sets one for all


@dataclass
class PMView:
    """One view of a batched PatchMatch (arguments of dense.pm_refine)."""
    ref: torch.Tensor
    ray: torch.Tensor
    Cr: np.ndarray
    srcs: Sequence[torch.Tensor]
    src_depths: Sequence[Optional[torch.Tensor]]
    par: np.ndarray
    depth: torch.Tensor
    normal: torch.Tensor
    score: torch.Tensor
    active: torch.Tensor
    geo_tol: float
    dmin: float
    dmax: float
    seed: int
    depth_from: Optional[torch.Tensor] = None  # copied into depth at the start of every run
    score_from: Optional[torch.Tensor] = None  # copied into score at the start of every run


class Batch:
    """A planned batch (om_batch).  set_sweep / set_pm plan the work (allocating the batch's device
    tables), run() enqueues it on the current stream, capture() records it in a HIP graph and
    replay() launches the graph.  The tensors passed in must stay alive with the batch."""

    def __init__(self):
        self._lib = _lib.lib()
        self._h = self._lib.om_batch_create()
        if not self._h:
            raise MemoryError("om_batch_create")
        self._keep: list = []
        self.n_pm = 0

    def close(self):
        if self._h:
            self._lib.om_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sweep(self, views: Sequence[SweepView], r: int, top_k: int, chunks: int = 0, tile_h: int = 0,
                  grid_blocks: int = 0):
        n = len(views)
        descs = (OmSweepDesc * n)()
        outs = (ctypes.c_void_p * n)()
        for v, V in enumerate(views):
            H, W = int(V.ref.shape[0]), int(V.ref.shape[1])
            if V.state is None:
                V.state = dense.SweepState(H, W, V.ref.device)
            mode, a, b, off, n_all = dense._hyp_inputs(V.hyp, H, W)
            d, keep = dense._sweep_desc(V.ref, V.mu, V.sd, list(V.srcs), V.ray, V.geo, r, top_k, V.state, mode, a, b,
                                        off, 0, n_all, int(V.chunks), int(V.tile_h))
            descs[v] = d
            self._keep += [keep, V]
            outs[v] = V.depth.data_ptr() if V.depth is not None else None
        check(self._lib.om_batch_set_sweep(self._h, descs, n, outs, int(chunks), int(tile_h), int(grid_blocks)),
              "batch sweep")
        self._keep.append(descs)

    def set_pm(self, views: Sequence[PMView], r: int, step: int, top_k: int, geo_w: float, iters: int,
               init_normals: bool = False):
        n = len(views)
        arr = (OmPmView * n)()
        for v, V in enumerate(views):
            H, W = int(V.ref.shape[0]), int(V.ref.shape[1])
            S = len(V.srcs)
            if S < 1 or S > 16:
                raise ValueError("PatchMatch needs 1..16 sources")
            o = arr[v]
            o.ref = dense._need(V.ref, "ref", shape=(H, W))
            o.H, o.W = H, W
            o.ray = dense._need(V.ray, "ray", shape=(3, H, W))
            for i in range(3):
                o.Cr[i] = float(V.Cr[i])
            o.S = S
            par = np.asarray(V.par, np.float64)
            for s in range(S):
                o.imgs[s] = dense._need(V.srcs[s], "src")
                o.deps[s] = None if V.src_depths[s] is None else dense._need(V.src_depths[s], "src depth")
                o.ih[s], o.iw[s] = int(V.srcs[s].shape[0]), int(V.srcs[s].shape[1])
                for t in range(18):
                    o.par[s][t] = float(par[s, t])
            o.depth = dense._need(V.depth, "depth", shape=(H, W))
            o.normal = dense._need(V.normal, "normal", shape=(3, H, W))
            o.score = dense._need(V.score, "score", shape=(H, W))
            o.active = dense._need(V.active, "active", torch.uint8, shape=(H, W))
            o.geo_tol, o.dmin, o.dmax = float(V.geo_tol), float(V.dmin), float(V.dmax)
            o.seed = int(V.seed) & 0xFFFFFFFF
            o.depth_from = None if V.depth_from is None else dense._need(V.depth_from, "depth_from", shape=(H, W))
            o.score_from = None if V.score_from is None else dense._need(V.score_from, "score_from", shape=(H, W))
            self._keep.append(V)
        check(self._lib.om_batch_set_pm(self._h, arr, n, int(r), int(step), int(top_k), float(geo_w), int(iters),
                                        int(bool(init_normals))), "batch pm")
        self.n_pm = n

    def run(self):
        check(self._lib.om_batch_run(self._h, dense._stream()), "batch run")

    def capture(self, stream: Optional[torch.cuda.Stream] = None):
        """Records run() into a HIP graph (on a side stream; torch's stream is not captured)."""
        s = stream or torch.cuda.Stream()
        torch.cuda.current_stream().synchronize()
        check(self._lib.om_batch_capture(self._h, s.cuda_stream), "batch capture")
        self._cap_stream = s

    def replay(self):
        check(self._lib.om_batch_replay(self._h, dense._stream()), "batch replay")

    def pm_changed(self) -> List[int]:
        out = (ctypes.c_int64 * max(self.n_pm, 1))()
        check(self._lib.om_batch_pm_changed(self._h, out), "batch pm_changed")
        return [int(out[i]) for i in range(self.n_pm)]

    def pm_evaluated(self) -> List[int]:
        """Propagation cost evaluations each view performed in the last run (-1: no memo)."""
        out = (ctypes.c_int64 * max(self.n_pm, 1))()
        check(self._lib.om_batch_pm_evaluated(self._h, out), "batch pm_evaluated")
        return [int(out[i]) for i in range(self.n_pm)]


def init_normals(depth: torch.Tensor, ray: torch.Tensor) -> torch.Tensor:
    """densify._init_normals (densify.py:481-496) on the device, float32 (3, H, W)."""
    H, W = int(depth.shape[0]), int(depth.shape[1])
    out = torch.empty((3, H, W), dtype=torch.float32, device=depth.device)
    check(_lib.lib().om_init_normals(dense._need(depth, "depth", shape=(H, W)), dense._need(ray, "ray", shape=(3, H, W)),
                                     H, W, out.data_ptr(), dense._stream()), "init_normals")
    return out
