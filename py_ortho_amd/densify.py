"""Densification on the MI355X: the drop-in for pyOrthomosaic's `densify.densify`.

`densify(ar, rec, gains, biases, out_dir, opt)` takes the reference's arguments and returns its
`DenseCloud` (densify.py:499-662), with every array after image decode resident in HBM:

1. views: each decoded uint8 image goes to the device once (3 bytes a pixel) and becomes the
   matching grey and the fusion colours there (`om_depth_view`);
2. pass 1: the depth map of every view (`mvs.depth_map`: pyramid, coarse band, sweep, two
   refinement levels); the score is kept as float16 and widened again, as the reference stores it
   (densify.py:590-592, :613);
3. pass 2: per view the geometric consistency count against the neighbours' pass-1 maps
   (`om_consistency`) and the PatchMatch inputs -- the pixels to repair, normals, the neighbours'
   maps as source depths, the median relative tolerance, the depth range or the map's own span, the
   view's seed; then PatchMatch of all views in one batch (`om_batch`; each view reads only pass-1
   maps, so the views are independent) and the final filter (densify.py:609-641);
4. fusion: the reference's sequential fusion over the final maps (`om_fuse`, densify.py:656-666).

Nothing goes through files: the `.npy` depth-map set is written only when `opt.keep_depthmaps`
is set, with the reference's names and bytes.  Image decode stays on the host: `load_rgb(frame,
scale)` and `load_mask(path, w, h)` are the reference's imageio.load_rgb and masks.load_mask unless
other loaders are passed.

Pass 1 runs the views one after another on the device stream, each launch spanning the whole GPU;
pass 2's PatchMatch runs all views of a block together.
"""
from __future__ import annotations

import logging
import os
import shutil
import time
from typing import Callable, Dict, Optional

import numpy as np
import torch

from . import batch, dense, formats, fusion, mvs, plan, prep

log = logging.getLogger(__name__)

__all__ = ["densify", "DensifyRun"]


def _default_loaders():
    from orthomosaic.imageio import load_rgb  # pyOrthomosaic's decoders, when installed beside this package
    from orthomosaic.masks import load_mask
    return load_rgb, load_mask


class DensifyRun:
    """One densification of a block.  The stages can be called one by one (tests, benchmarks);
    densify() runs them all."""

    def __init__(self, ar, rec, gains: dict, biases: Optional[dict], opt, load_rgb: Optional[Callable] = None,
                 load_mask: Optional[Callable] = None, device="cuda", exact: bool = False):
        self.ar, self.rec, self.opt = ar, rec, opt
        self.exact = bool(exact)
        self.gains, self.biases = gains or {}, biases
        if load_rgb is None or load_mask is None:
            d_rgb, d_mask = _default_loaders()
            load_rgb, load_mask = load_rgb or d_rgb, load_mask or d_mask
        self.load_rgb, self.load_mask = load_rgb, load_mask
        self.device = torch.device(device)
        self.N = len(rec.used)
        self._cameras()
        self._views: Dict[int, tuple] = {}
        self.nbrs = plan.neighbours(rec, opt.neighbors)
        self._settings()
        self.raw: Dict[int, torch.Tensor] = {}
        self.score16: Dict[int, torch.Tensor] = {}
        self.low: Dict[int, torch.Tensor] = {}
        self.final: Dict[int, torch.Tensor] = {}
        self.stats = {"repaired": 0, "active": 0, "pixels": 0}

    # ------------------------------------------------------------------ set-up (densify.py:517-586)
    def _cameras(self):
        rec, frames, opt = self.rec, self.ar.frames, self.opt
        self.scales, self.cams = {}, {}
        for k, i in enumerate(rec.used):
            fr = frames[i]
            long_side = max(fr.width, fr.height)
            s = min(1.0, (opt.max_image_dim or long_side) / long_side)
            w, h = int(round(fr.width * s)), int(round(fr.height * s))
            it = rec.intr[rec.cam_group[k]]
            full = mvs.Cam(rec.R[k], rec.C[k], it.f, it.k1, it.k2, it.cx, it.cy, fr.width, fr.height,
                           getattr(it, "k3", 0.0))
            self.cams[k] = full.scaled(w / fr.width, w, h)
            self.scales[k] = s

    def _settings(self):
        rec, opt = self.rec, self.opt
        tp = plan.tie_points(rec)
        own_z = {}
        for k in range(self.N):
            idx = tp.own(k)
            if len(idx):
                own_z[k] = (rec.R[k] @ (rec.X[idx] - rec.C[k]).T)[2]
        self.med_gsd = float(np.median([float(np.nanmedian(own_z[k])) / self.cams[k].f for k in sorted(own_z)]))
        scale_med = float(np.median(list(self.scales.values())))
        self.px_tol = opt.consistency_px or float(np.clip(2.0 * rec.rms_px * scale_med, 1.0, 3.0))
        n_good = np.array([len(x) for x in self.nbrs])
        self.min_views = opt.min_views or (3 if np.median(n_good) >= 3 else 2)
        self.dmed = {k: float(np.median(z)) for k, z in own_z.items()}

    def rel_tols(self, k: int) -> dict:
        """Relative depth tolerance per neighbour: the depth change that moves a point by px_tol
        pixels between the two views, in [0.01, 0.05] (densify.py:574-582)."""
        C, out = self.rec.C, {}
        for j in self.nbrs[k]:
            B = float(np.linalg.norm(C[k] - C[j]))
            out[j] = (self.opt.consistency_depth or
                      float(np.clip(self.px_tol * self.dmed.get(k, 1.0) / max(self.cams[k].f * B, 1e-9), 0.01, 0.05)))
        return out

    # ------------------------------------------------------------------ views
    def view(self, k: int):
        """(grey float32 (h, w) NaN = masked, rgb8 uint8 (h, w, 3), camera) of view k on the device,
        decoded and prepared once (densify.py:528-535)."""
        v = self._views.get(k)
        if v is None:
            i = self.rec.used[k]
            rgb = np.ascontiguousarray(self.load_rgb(self.ar.frames[i], self.scales[k]), np.uint8)
            mask = self.load_mask(self.ar.frames[i].path, rgb.shape[1], rgb.shape[0])
            b = None if self.biases is None else self.biases.get(i)
            grey, rgb8, _ = prep.prepare_view(rgb, self.gains.get(i), b, mask, levels=1, device=self.device)
            v = (grey, rgb8, self.cams[k])
            self._views[k] = v
        return v

    # ------------------------------------------------------------------ pass 1 (densify.py:587-603)
    def pass1(self):
        for k in range(self.N):
            r = mvs.depth_map(k, self.nbrs[k], self.rec, self.view, self.opt, device=self.device, exact=self.exact)
            if r is None:
                continue
            depth, score, low = r
            self.raw[k] = depth.contiguous()
            self.score16[k] = formats.to_f16(score)
            self.low[k] = low

    # ------------------------------------------------------------------ pass 2 (densify.py:609-641)
    def repair_inputs(self, k: int):
        """View k's repair inputs (densify.py:611-634): (D, S, pm) with D a copy of the pass-1 map and
        S the widened float16 score, which PatchMatch refines in place, and pm the remaining
        pm_refine arguments as a batch.PMView, or None when there is nothing to repair (PatchMatch
        off, no active pixel or no neighbour map)."""
        opt, rec = self.opt, self.rec
        D = self.raw[k].clone()
        S = formats.to_f32(self.score16[k])
        if not opt.patchmatch:
            return D, S, None
        tol = self.rel_tols(k)
        cnt = fusion.consistency(k, D, self.nbrs[k], self.raw, self.cams, self.px_tol, tol)
        grey, _, cr = self.view(k)
        # active = finite(grey) & (cnt < min_views - 1 | S < min_ncc | ~finite(D))  (densify.py:619)
        active = (torch.isfinite(grey) & ((cnt < self.min_views - 1) | (S < opt.min_ncc) | ~torch.isfinite(D)))
        active = active.to(torch.uint8).contiguous()
        src_ids = [j for j in self.nbrs[k] if j in self.raw]
        self.stats["pixels"] += int(torch.isfinite(D).sum())
        n_active = int(active.sum())
        if not n_active or not src_ids:
            return D, S, None
        self.stats["active"] += n_active
        ray = mvs.rays_world(tuple(D.shape), cr, self.device)
        normal = mvs.init_normals(D, ray)
        par = np.array([np.r_[self.cams[j].R.ravel(), self.cams[j].C, self.cams[j].f, self.cams[j].k1,
                              self.cams[j].k2, self.cams[j].k3, self.cams[j].cx, self.cams[j].cy]
                        for j in src_ids], np.float64)
        rng = plan.depth_range(rec, k)
        if rng is None:                      # the map's own span (np.nanmin / np.nanmax)
            fin = D[torch.isfinite(D)]
            rng = (float(fin.min()), float(fin.max())) if fin.numel() else (float("nan"), float("nan"))
        pm = batch.PMView(grey, ray, np.asarray(cr.C, np.float64), [self.view(j)[0] for j in src_ids],
                          [self.raw[j] for j in src_ids], par, D, normal, S, active,
                          float(np.median(list(tol.values()))), rng[0], rng[1], k * 2654435761 % 2 ** 32)
        return D, S, pm

    def _final(self, k: int, D: torch.Tensor, S: torch.Tensor) -> torch.Tensor:
        # D[(S < min_ncc) | low] = NaN with the float32 score (densify.py:638)
        bad = (S < self.opt.min_ncc) | self.low[k]
        return torch.where(bad, torch.full_like(D, float("nan")), D).contiguous()

    def repair(self, k: int) -> torch.Tensor:
        """View k's final depth map: PatchMatch over its inconsistent pixels (when opt.patchmatch),
        then NaN where the score is below min_ncc or the window is flat."""
        opt = self.opt
        D, S, pm = self.repair_inputs(k)
        if pm is not None:
            ch = dense.pm_refine(pm.ref, pm.ray, pm.Cr, pm.srcs, pm.src_depths, pm.par, pm.depth, pm.normal, pm.score,
                                 pm.active, opt.pm_window, opt.pm_step, opt.top_k, opt.pm_geo_weight, pm.geo_tol,
                                 opt.pm_iters, pm.dmin, pm.dmax, pm.seed)
            self.stats["repaired"] += int(ch)
        return self._final(k, D, S)

    def pass2(self, batched: bool = True, max_batch_pixels: int = 64 << 20):
        """Pass 2 of every view.  Each view's repair reads only pass-1 maps (its own and its
        neighbours'), so the views are independent: batched, their PatchMatch runs as one om_batch
        (all views' colour passes together) per group of up to max_batch_pixels reference pixels,
        with the same inputs and results as the per-view calls."""
        keys = sorted(self.raw)
        if not batched or not self.opt.patchmatch:
            for k in keys:
                self.final[k] = self.repair(k)
            return
        opt = self.opt
        group, px = [], 0

        def flush():
            nonlocal group, px
            todo = [(k, D, S, pm) for k, D, S, pm in group if pm is not None]
            if todo:
                B = batch.Batch()
                try:
                    B.set_pm([pm for _, _, _, pm in todo], opt.pm_window, opt.pm_step, opt.top_k, opt.pm_geo_weight,
                             opt.pm_iters)
                    B.run()
                    self.stats["repaired"] += sum(B.pm_changed())
                finally:
                    B.close()
            for k, D, S, _ in group:
                self.final[k] = self._final(k, D, S)
            group, px = [], 0

        for k in keys:
            D, S, pm = self.repair_inputs(k)
            n = int(D.numel()) if pm is not None else 0
            if group and px + n > max_batch_pixels:
                flush()
            group.append((k, D, S, pm))
            px += n
        flush()

    # ------------------------------------------------------------------ fusion (densify.py:652-666)
    def fuse(self) -> formats.DenseCloud:
        have = sorted(self.final)
        pts, cols, nv, owner = [], [], [], []
        if have:
            rgb = {k: self.view(k)[1] for k in have}
            fusion.fuse(have, self.final, None, self.cams, self.nbrs, set(have), self.rel_tols, self.opt,
                        self.px_tol, self.min_views, rgb, pts, cols, nv, owner)
        return formats.dense_cloud(pts, cols, nv, owner, self.med_gsd * max(1, self.opt.point_stride))

    def save(self, wd: str):
        """The depth-map files of densify's two passes (densify.py:590-592, :639)."""
        os.makedirs(wd, exist_ok=True)
        for k in sorted(self.raw):
            np.save(os.path.join(wd, f"p{k:05d}.npy"), self.raw[k].cpu().numpy())
            np.save(os.path.join(wd, f"s{k:05d}.npy"), self.score16[k].cpu().numpy())
            np.save(os.path.join(wd, f"t{k:05d}.npy"), self.low[k].to(torch.bool).cpu().numpy())
        for k, d in sorted(self.final.items()):
            formats.save_final(wd, k, d)


def densify(ar, rec, gains: dict, biases: Optional[dict], out_dir: str, opt, load_rgb: Optional[Callable] = None,
            load_mask: Optional[Callable] = None, device="cuda", exact: bool = False) -> formats.DenseCloud:
    """densify.densify (densify.py:499-510) on the device -> DenseCloud.  <out_dir>/depthmaps holds
    the reference's depth-map files afterwards only when opt.keep_depthmaps is set.  exact: the
    sweeps in the reference's summation order (om_sweep_rows), so the cloud is the reference's bit
    for bit for any input; otherwise the fused sweep, equal to it except where an FP64 window sum
    rounds (scores within 1e-4)."""
    wd = os.path.join(out_dir, "depthmaps")
    shutil.rmtree(wd, ignore_errors=True)
    t0 = time.time()
    run = DensifyRun(ar, rec, gains, biases, opt, load_rgb, load_mask, device, exact)
    run.pass1()
    log.info("Depth maps done in %.1fs (%d/%d images)", time.time() - t0, len(run.raw), run.N)
    run.pass2()
    if opt.keep_depthmaps:
        run.save(wd)
    cloud = run.fuse()
    log.info("Dense cloud: %d points fused in %.1fs", len(cloud.xyz), time.time() - t0)
    return cloud
