"""Depth maps of a whole block of views in one batch on one GPU.

The reference builds a block's depth maps in two thread-pool passes (densify.py:540-650): pass 1
runs the plane sweep of every view (one view per host thread, densify.py:596), pass 2 repairs each
view with PatchMatch, reading its neighbours' pass-1 maps for the geometric term
(densify.py:620-634).  Here both passes of all views are one om_batch: a persistent work-queue
sweep over every (view, tile, hypothesis chunk), the parabola, then the PatchMatch passes of all
views, optionally captured once in a HIP graph and replayed.

Per view k the sweep tests `num_depths` fronto-parallel planes, uniform in inverse depth over the
view's range [dmin, dmax] (the reference's plane set, densify.py:405-412, with a per-view instead
of a per-block band), with window radius `window` and top-k mean NCC; the parabola gives the pass-1
depth (densify.py:354-361).  PatchMatch then starts from a copy of that depth, normals from
densify._init_normals and the sweep's best score, and refines every pixel (all pixels active) for
`pm_iters` iterations against the same neighbours, whose pass-1 maps are the source depths.

Every result equals the per-view calls (dense.sweep + dense.sweep_finalize + dense.pm_refine) bit
for bit; `run_single` runs exactly those per-view calls for comparison.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import batch, dense, formats, fusion

__all__ = ["BlockOptions", "BlockDepth"]


@dataclass
class BlockOptions:
    num_depths: int = 192
    window: int = 5
    top_k: int = 2
    pm_iters: int = 6
    pm_window: int = 4
    pm_step: int = 2
    pm_geo_weight: float = 0.5
    geo_tol: float = 0.01
    # PatchMatch's pixels and tolerance.  "all": every pixel with geo_tol (the 48-view block
    # workload, all of it one HIP graph).  "inconsistent": the reference's pass-2 inputs built from
    # the sweep's maps (densify.py:611-634): the pixels that are valid in the image and inconsistent
    # with fewer than min_views - 1 neighbours, below min_ncc or without a depth, and the median of
    # the neighbours' relative tolerances (densify.py:574-582, with the view's median sweep depth in
    # place of its tie points' median depth, which a block of images alone does not have).
    repair: str = "all"
    min_ncc: float = 0.3
    min_views: int = 0          # 0: 3 when the median view has >= 3 neighbours, else 2 (densify.py:560)
    px_tol: float = 1.5         # reprojection tolerance of the consistency check, pixels


def _geo18(cam_s, cam_r) -> np.ndarray:
    b = np.asarray(cam_s.R) @ (np.asarray(cam_r.C) - np.asarray(cam_s.C))
    return np.r_[np.asarray(cam_s.R).ravel(), b, [cam_s.f, cam_s.k1, cam_s.k2, cam_s.k3, cam_s.cx, cam_s.cy]]


def _par18(cam) -> np.ndarray:
    return np.r_[np.asarray(cam.R).ravel(), np.asarray(cam.C), [cam.f, cam.k1, cam.k2, cam.k3, cam.cx, cam.cy]]


def inverse_depth_planes(dmin: float, dmax: float, n: int) -> np.ndarray:
    """n inverse depths from 1/dmax to 1/dmin, evenly spaced (float32 of the float64 values)."""
    lo, hi = 1.0 / dmax, 1.0 / dmin
    step = (hi - lo) / max(n - 1, 1)
    return (lo + np.arange(n) * step).astype(np.float32)


class BlockDepth:
    """imgs[k]: (h, w) float32 device image (NaN = masked); cams[k]: camera with R, C, f, k1, k2, k3,
    cx, cy, w, h; nbrs[k]: source view indices of view k (1..16); ranges[k]: (dmin, dmax)."""

    def __init__(self, imgs: Sequence[torch.Tensor], cams: Sequence, nbrs: Sequence[Sequence[int]],
                 ranges: Sequence[Tuple[float, float]], opt: Optional[BlockOptions] = None):
        self.opt = opt or BlockOptions()
        o = self.opt
        n = len(imgs)
        if not (len(cams) == len(nbrs) == len(ranges) == n):
            raise ValueError("imgs, cams, nbrs and ranges need one entry per view")
        self.n = n
        self.cams = list(cams)
        self.nbrs = [list(map(int, nb)) for nb in nbrs]
        self.ranges = [(float(a), float(b)) for a, b in ranges]
        for k, nb in enumerate(self.nbrs):
            if not 1 <= len(nb) <= 16 or any(j == k or not 0 <= j < n for j in nb):
                raise ValueError(f"view {k}: needs 1..16 neighbours among the other views")
        for im in imgs:
            dense._need(im, "image")
        self.imgs = list(imgs)
        self.ref = [torch.where(torch.isfinite(im), im, torch.zeros_like(im)).contiguous() for im in imgs]
        self.stats = [dense.ref_stats(r, o.window) for r in self.ref]
        self.rays = [dense.rays_world(int(c.h), int(c.w), c.R, c.cx, c.cy, c.f, c.k1, c.k2, c.k3) for c in self.cams]
        dev = imgs[0].device
        self.planes_host = [inverse_depth_planes(a, b, o.num_depths) for a, b in self.ranges]
        self.planes = [torch.from_numpy(p).to(dev) for p in self.planes_host]
        self.raw = [torch.empty(im.shape, dtype=torch.float32, device=dev) for im in imgs]
        self.depth = [torch.empty(im.shape, dtype=torch.float32, device=dev) for im in imgs]
        self.normal = [torch.empty((3,) + tuple(im.shape), dtype=torch.float32, device=dev) for im in imgs]
        self.score = [torch.empty(im.shape, dtype=torch.float32, device=dev) for im in imgs]
        self.active = [torch.ones(im.shape, dtype=torch.uint8, device=dev) for im in imgs]
        self.states = [dense.SweepState(int(im.shape[0]), int(im.shape[1]), dev) for im in imgs]
        self._batch: Optional[batch.Batch] = None
        self._pm = False

    # ---- accounting
    def pixels(self) -> List[int]:
        return [int(im.shape[0]) * int(im.shape[1]) for im in self.imgs]

    def pixel_hypotheses(self) -> Tuple[int, int]:
        """(sweep, PatchMatch) pixel-hypotheses of one run: H W D and H W 11 iters per view."""
        px = sum(self.pixels())
        return px * self.opt.num_depths, px * 11 * self.opt.pm_iters

    def size_class(self, k: int) -> Tuple[int, int]:
        return int(self.imgs[k].shape[0]), int(self.imgs[k].shape[1])

    # ---- views of the batch
    def sweep_view(self, k: int, tile_h: int = 0, chunks: int = 0, depth: Optional[torch.Tensor] = None,
                   planes: Optional[torch.Tensor] = None) -> batch.SweepView:
        mu, sd = self.stats[k]
        return batch.SweepView(self.ref[k], mu, sd, [self.imgs[j] for j in self.nbrs[k]], self.rays[k],
                               np.stack([_geo18(self.cams[j], self.cams[k]) for j in self.nbrs[k]]),
                               {"mode": "const", "inv": self.planes[k] if planes is None else planes},
                               state=self.states[k], depth=self.raw[k] if depth is None else depth,
                               tile_h=tile_h, chunks=chunks)

    def pm_view(self, k: int) -> batch.PMView:
        a, b = self.ranges[k]
        return batch.PMView(self.imgs[k], self.rays[k], np.asarray(self.cams[k].C, np.float64),
                            [self.imgs[j] for j in self.nbrs[k]], [self.raw[j] for j in self.nbrs[k]],
                            np.stack([_par18(self.cams[j]) for j in self.nbrs[k]]), self.depth[k], self.normal[k],
                            self.score[k], self.active[k], self.opt.geo_tol, a, b, self.seed(k),
                            depth_from=self.raw[k], score_from=self.states[k].best)

    @staticmethod
    def seed(k: int) -> int:
        return (k * 2654435761) % 2 ** 32  # the reference's per-view seed (densify.py:634)

    # ---- the reference's repair inputs (BlockOptions.repair == "inconsistent")
    def min_views(self) -> int:
        o = self.opt
        return o.min_views or (3 if np.median([len(nb) for nb in self.nbrs]) >= 3 else 2)

    def rel_tols(self, k: int, dmed: float) -> Dict[int, float]:
        """Relative depth tolerance per neighbour of view k (densify.py:574-582): the depth change
        that moves a point by px_tol pixels between the two views, in [0.01, 0.05]."""
        C = [np.asarray(c.C, np.float64) for c in self.cams]
        f = float(self.cams[k].f)
        return {j: float(np.clip(self.opt.px_tol * dmed / max(f * float(np.linalg.norm(C[k] - C[j])), 1e-9),
                                 0.01, 0.05)) for j in self.nbrs[k]}

    def repair_inputs(self):
        """Per view (active, geo_tol, score) from the sweep's maps (self.raw, self.states): the
        consistency count against the neighbours' maps (om_consistency), the score rounded to
        float16 and widened as densify stores it, active = finite(image) & (count < min_views - 1 |
        score < min_ncc | ~finite(depth)) (densify.py:619) and geo_tol = median(rel_tols)
        (densify.py:632)."""
        o, mv = self.opt, self.min_views()
        dm = {k: self.raw[k] for k in range(self.n)}
        out = []
        for k in range(self.n):
            D = self.raw[k]
            fin = D[torch.isfinite(D)]
            dmed = float(fin.median()) if fin.numel() else 1.0
            tol = self.rel_tols(k, dmed)
            cnt = fusion.consistency(k, D, self.nbrs[k], dm, self.cams, o.px_tol, tol)
            S = formats.to_f32(formats.to_f16(self.states[k].best))
            active = torch.isfinite(self.imgs[k]) & ((cnt < mv - 1) | (S < o.min_ncc) | ~torch.isfinite(D))
            out.append((active.to(torch.uint8).contiguous(), float(np.median(list(tol.values()))), S))
        return out

    # ---- batched run
    def plan(self, tune: Optional[Dict[Tuple[int, int], Tuple[int, int]]] = None, pm: bool = True) -> batch.Batch:
        """Plans the batch; tune maps a size class (H, W) to (tile_h, chunks) (autotune.tune_block).
        With repair == "inconsistent" the batch holds the sweep only: run() then builds the repair
        inputs from its maps and runs PatchMatch as a second batch."""
        if self._batch is not None:
            self._batch.close()
        if self.opt.repair not in ("all", "inconsistent"):
            raise ValueError(f"BlockOptions.repair must be 'all' or 'inconsistent' (got {self.opt.repair!r})")
        o = self.opt
        B = batch.Batch()
        sv = []
        for k in range(self.n):
            th, ch = (tune or {}).get(self.size_class(k), (0, 0))
            sv.append(self.sweep_view(k, th, ch))
        B.set_sweep(sv, o.window, o.top_k)
        self._pm = pm and o.pm_iters > 0
        if self._pm and o.repair == "all":
            B.set_pm([self.pm_view(k) for k in range(self.n)], o.pm_window, o.pm_step, o.top_k, o.pm_geo_weight,
                     o.pm_iters, init_normals=True)
        self._batch = B
        return B

    def run(self):
        if self._batch is None:
            self.plan()
        self._batch.run()
        if self._pm and self.opt.repair == "inconsistent":
            self._run_repair()

    def _run_repair(self):
        """PatchMatch of every view over the reference's repair inputs, as one batch."""
        o = self.opt
        inputs = self.repair_inputs()
        views = []
        for k, (active, geo_tol, S) in enumerate(inputs):
            self.active[k].copy_(active)
            self.score[k].copy_(S)
            v = self.pm_view(k)
            v.geo_tol, v.score_from = geo_tol, None
            views.append(v)
        B = batch.Batch()
        try:
            B.set_pm(views, o.pm_window, o.pm_step, o.top_k, o.pm_geo_weight, o.pm_iters, init_normals=True)
            B.run()
        finally:
            torch.cuda.current_stream().synchronize()
            B.close()

    def capture(self):
        if self.opt.repair != "all":
            raise ValueError("capture() records the 'all' repair mode; 'inconsistent' builds its inputs on the host")
        if self._batch is None:
            self.plan()
        self._batch.capture()

    def replay(self):
        self._batch.replay()

    def close(self):
        if self._batch is not None:
            self._batch.close()
            self._batch = None

    # ---- the plain per-view form, for comparison
    def run_single(self, per_hypothesis: bool = False):
        """Every view on its own: the fused single-view sweep (or, with per_hypothesis, one
        sweep_hypothesis launch per plane, the reference's loop shape densify.py:339-342), the
        parabola, then the single-view PatchMatch of every view."""
        o = self.opt
        for k in range(self.n):
            v = self.sweep_view(k)
            st = self.states[k]
            st.reset()
            if per_hypothesis:
                plane = torch.empty_like(v.ref)
                for i in range(o.num_depths):
                    plane.fill_(float(self.planes_host[k][i]))
                    dense.sweep_hypothesis(v.ref, v.mu, v.sd, v.srcs, v.ray, v.geo, plane, o.window, o.top_k, i,
                                           *st.tensors())
            else:
                dense.sweep(v.ref, v.mu, v.sd, v.srcs, v.ray, v.geo, v.hyp, o.window, o.top_k, st)
            self.raw[k].copy_(dense.sweep_finalize(v.ref, v.mu, v.sd, v.srcs, v.ray, v.geo, v.hyp, o.window, o.top_k,
                                                   st))
        if o.pm_iters <= 0:
            return
        inputs = self.repair_inputs() if o.repair == "inconsistent" else None
        for k in range(self.n):
            P = self.pm_view(k)
            self.depth[k].copy_(self.raw[k])
            geo_tol = o.geo_tol
            if inputs is None:
                self.score[k].copy_(self.states[k].best)
            else:
                active, geo_tol, S = inputs[k]
                self.active[k].copy_(active)
                self.score[k].copy_(S)
            self.normal[k].copy_(batch.init_normals(self.raw[k], self.rays[k]))
            dense.pm_refine(P.ref, P.ray, P.Cr, P.srcs, P.src_depths, P.par, self.depth[k], self.normal[k],
                            self.score[k], self.active[k], o.pm_window, o.pm_step, o.top_k, o.pm_geo_weight,
                            geo_tol, o.pm_iters, P.dmin, P.dmax, P.seed)
