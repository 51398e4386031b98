"""Per-view depth maps on the MI355X: the drop-in for pyOrthomosaic's depth-map entry points.

Mirrors the reference's `densify` depth-map API (orthomosaic/densify.py:57-437) with every array
on the device and every hot step a libom355 kernel:

- `DepthOptions`   the reference's knobs, same names and defaults (densify.py:57-81)
- `Cam`            pinhole + radial camera, pixel centres at integers (densify.py:133-160)
- `rays_world`     (3, H, W) world rays (densify.py:288-297)
- `box`, `half`    window mean and 2x2 pyramid step (densify.py:94-103, :127-130)
- `_sweep`/`sweep` winner-take-all over inverse-depth hypotheses + parabola (densify.py:316-361)
- `depth_map`      pyramid, coarse sweep, two refinement levels (densify.py:373-437)
- `init_normals`, `pm_refine`  PatchMatch repair inputs and kernel (densify.py:481-496,
  _dense.pyx:566-666)

Hypotheses are never materialised: the sweep kernel generates them per pixel from the same
float32 expressions as the reference's `_fast.hyp_linear` / `_fast.hyp_refine`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import dense, plan

__all__ = ["DepthOptions", "Cam", "rays_world", "box", "half", "sweep", "_sweep", "depth_map", "init_normals",
           "pm_refine", "sweep_hypothesis"]

sweep_hypothesis = dense.sweep_hypothesis
pm_refine = dense.pm_refine


@dataclass
class DepthOptions:
    max_image_dim: int = 1600
    neighbors: int = 4
    num_depths: int = 64
    refine_steps: int = 3
    window: int = 5
    top_k: int = 2
    min_ncc: float = 0.3
    min_texture: float = 0.5
    min_views: int = 0
    consistency_px: float = 0.0
    consistency_depth: float = 0.0
    patchmatch: bool = False
    pm_iters: int = 2
    pm_window: int = 4
    pm_step: int = 2
    pm_geo_weight: float = 0.5
    layer_gap: float = 1.0
    keep_depthmaps: bool = False
    point_stride: int = 2
    cache_mb: int = 1200
    workers: int = 0


class Cam:
    """Pinhole + 3-term radial camera (R world->camera, C centre, pixel centres at integers)."""

    def __init__(self, R, C, f, k1, k2, cx, cy, w, h, k3=0.0):
        self.R, self.C = np.asarray(R, np.float64), np.asarray(C, np.float64)
        self.f, self.k1, self.k2, self.k3 = float(f), float(k1), float(k2), float(k3)
        self.cx, self.cy, self.w, self.h = float(cx), float(cy), int(w), int(h)

    def scaled(self, s: float, w: int, h: int) -> "Cam":
        """The same camera for an image resampled by s (pixel centres: (c + .5) s - .5)."""
        return Cam(self.R, self.C, self.f * s, self.k1, self.k2, (self.cx + 0.5) * s - 0.5,
                   (self.cy + 0.5) * s - 0.5, w, h, self.k3)

    def project_cam(self, x0, x1, x2):
        """Camera-frame points -> pixels (u, v); depth clamped to 1e-6 (check z > 0 separately)."""
        z = np.maximum(x2, 1e-6)
        a, b = x0 / z, x1 / z
        q = a * a + b * b
        d = self.f * (1 + q * (self.k1 + q * (self.k2 + q * self.k3)))
        return d * a + self.cx, d * b + self.cy

    @classmethod
    def of(cls, c) -> "Cam":
        """Adopt any camera object with R, C, f, k1, k2, (k3,) cx, cy and w/h (or width/height)."""
        w = getattr(c, "w", getattr(c, "width", 0))
        h = getattr(c, "h", getattr(c, "height", 0))
        return cls(c.R, c.C, c.f, c.k1, c.k2, c.cx, c.cy, w, h, getattr(c, "k3", 0.0))


def _dev(a, device="cuda") -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)


def rays_world(shape, cam, device="cuda") -> torch.Tensor:
    H, W = shape
    c = Cam.of(cam)
    return dense.rays_world(H, W, c.R, c.cx, c.cy, c.f, c.k1, c.k2, c.k3, device=device)


def box(a: torch.Tensor, r: int) -> torch.Tensor:
    return dense.box_zero_mean(a, r)


def half(g: torch.Tensor) -> torch.Tensor:
    return dense.half(g)


def _geo(cam_r: Cam, cams_s: Sequence[Cam]) -> np.ndarray:
    rows = []
    for cs in cams_s:
        b = cs.R @ (cam_r.C - cs.C)
        rows.append(np.r_[cs.R.ravel(), b, [cs.f, cs.k1, cs.k2, cs.k3, cs.cx, cs.cy]])
    return np.asarray(rows, np.float64)


def sweep(ref, cam_r, srcs, cams_s, hyp: dict, opt: DepthOptions, ray=None, exact: bool = False):
    """densify._sweep with the hypotheses given by a generator description (see dense.sweep):
    winner-take-all over all hypotheses in one fused launch (exact=True, more than 16 sources, a
    window radius above 7 or a top-k above 4: the row/column-streaming kernels in the reference's
    summation order), then the parabola.  -> (depth, best, sd) device float32 (H, W)."""
    r = int(opt.window)
    ref = _dev(ref)
    H, W = ref.shape
    cam_r = Cam.of(cam_r)
    cams_s = [Cam.of(c) for c in cams_s]
    if ray is None:
        ray = rays_world((H, W), cam_r, ref.device)
    mu, sd = dense.ref_stats(ref, r)
    src = [_dev(s) for s in srcs]
    geo = _geo(cam_r, cams_s)
    st = dense.sweep(ref, mu, sd, src, ray, geo, hyp, r, opt.top_k, exact=exact)
    depth = dense.sweep_finalize(ref, mu, sd, src, ray, geo, hyp, r, opt.top_k, st)
    return depth, st.best, sd


def _sweep(ref, cam_r, srcs, cams_s, inv_list, opt: DepthOptions, ray=None, exact: bool = False):
    """The reference signature (densify.py:316): `inv_list` is any sequence of (H, W) inverse-depth
    maps, uniformly spaced (numpy arrays, tensors, or a lazy sequence like densify._Hypotheses)."""
    ref_t = _dev(ref)
    maps = torch.stack([_dev(inv_list[i], ref_t.device) for i in range(len(inv_list))])
    return sweep(ref_t, cam_r, srcs, cams_s, {"mode": "maps", "maps": maps}, opt, ray, exact)


def _upsample2(a: torch.Tensor, H: int, W: int) -> torch.Tensor:
    """np.repeat x2 on both axes cropped to (H, W), edge-padded when the finer level is odd."""
    up = a.repeat_interleave(2, 0).repeat_interleave(2, 1)[:H, :W]
    ph, pw = H - up.shape[0], W - up.shape[1]
    if ph or pw:
        up = torch.nn.functional.pad(up[None, None], (0, pw, 0, ph), mode="replicate")[0, 0]
    return up.contiguous()


def depth_map(k, nbrs, rec, views: Callable, opt: DepthOptions, device="cuda", exact: bool = False):
    """Depth (camera-axis, metres), NCC score and low-texture flag of reference image k at matching
    scale (densify.py:373-437), on the device.  views(j) -> (grey (h, w) float32 NaN = masked,
    rgb, camera) as in the reference.  Returns None when the image has no range or no neighbours.
    exact: every sweep in the reference's summation order (see sweep)."""
    rng = plan.depth_range(rec, k)
    if rng is None or not nbrs:
        return None
    grey, _, cam_r = views(k)
    cam_r = Cam.of(cam_r)
    g = _dev(grey, device)
    ref_mask = ~torch.isfinite(g)
    ref = torch.where(ref_mask, torch.zeros_like(g), g).contiguous()
    src_v = [views(j) for j in nbrs]
    srcs = [_dev(s[0], device) for s in src_v]
    cams_s = [Cam.of(s[2]) for s in src_v]

    L = 2
    pyr_r, pyr_c, pyr_s, pyr_cs = [ref], [cam_r], [srcs], [cams_s]
    for _ in range(L):
        r_ = half(pyr_r[-1])
        pyr_r.append(r_)
        pyr_c.append(pyr_c[-1].scaled(0.5, r_.shape[1], r_.shape[0]))
        s_ = [half(x) for x in pyr_s[-1]]
        pyr_s.append(s_)
        pyr_cs.append([c.scaled(0.5, x.shape[1], x.shape[0]) for c, x in zip(pyr_cs[-1], s_)])

    # rays at full resolution, coarser levels by 2x2 averaging (the pyramid step per channel)
    rays = [rays_world(ref.shape, cam_r, device)]
    for lvl in range(1, L + 1):
        rays.append(torch.stack([half(rays[-1][c]) for c in range(3)]).contiguous())

    nd = plan.num_depths(rec, k, nbrs, rng, pyr_c[L].f, opt.num_depths)
    shape = tuple(pyr_r[L].shape)
    base, _, step = plan.depth_band(rec, k, pyr_c[L], shape, rng, block=max(8, shape[1] // 16), nd=nd,
                                    device=device)
    d, _, _ = sweep(pyr_r[L], pyr_c[L], pyr_s[L], pyr_cs[L], {"mode": "linear", "base": base, "step": step, "n": nd},
                    opt, rays[L], exact)
    dstep = (step / 2).contiguous()
    rs = int(opt.refine_steps)
    score = sd = None
    for lvl in range(L - 1, -1, -1):
        H, W = pyr_r[lvl].shape
        up = _upsample2(d, H, W)
        dstep = _upsample2(dstep, H, W)
        iu = (1.0 / up).contiguous()
        d, score, sd = sweep(pyr_r[lvl], pyr_c[lvl], pyr_s[lvl], pyr_cs[lvl],
                             {"mode": "refine", "iu": iu, "ds": dstep, "steps": rs}, opt, rays[lvl], exact)
        dstep = (dstep / 2).contiguous()
    depth = d
    sd = sd * 64.0
    bad = ref_mask | (depth < rng[0]) | (depth > rng[1])
    depth = torch.where(bad, torch.full_like(depth, float("nan")), depth)
    low_texture = sd < opt.min_texture
    return depth, score, low_texture


def init_normals(D: torch.Tensor, ray: torch.Tensor) -> torch.Tensor:
    """World normals of a depth map from 2-px central differences, facing the camera, (0,0,1) where
    undefined (densify.py:481-496), in NumPy's float32 operation order.  -> (3, H, W) float32."""
    from .batch import init_normals as _k
    return _k(_dev(D, D.device), _dev(ray, ray.device))
