"""Seeded synthetic survey scenes with exact per-view depth, for tests and benchmarks.

Terrain: a base plane at 0 m plus Gaussian bumps (heights summed, clipped to [0, hmax]).
Texture: multi-octave value noise on the ground plane (hashed lattice, smoothstep interpolation),
normalised to [0, 1], plus Gaussian pixel noise.  Views: pinhole cameras (pyOrthomosaic
convention: pixel centres at integers, cx = (W-1)/2, R world->camera, camera-axis depth), rendered
by ray casting the height field (fixed-step march to the first crossing, then bisection).

Patterns from the reference's synthetic survey (tests/synthetic3d.py:106-151, tests/synthetic.py:24-53),
written for torch so that drone-size views render on the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

NADIR = np.diag([1.0, -1.0, -1.0])  # image x = east, image y = south, looking down


@dataclass
class Camera:
    R: np.ndarray       # (3, 3) world -> camera
    C: np.ndarray       # (3,) centre
    f: float
    cx: float
    cy: float
    w: int
    h: int
    k1: float = 0.0
    k2: float = 0.0
    k3: float = 0.0

    def par18(self) -> np.ndarray:
        """PatchMatch source layout (_dense.pyx:430): R (9), C (3), f, k1, k2, k3, cx, cy."""
        return np.r_[self.R.ravel(), self.C, [self.f, self.k1, self.k2, self.k3, self.cx, self.cy]]

    def geo18(self, ref: "Camera") -> np.ndarray:
        """Sweep source layout (_dense.pyx:748): Rs (9), b = Rs (C_ref - C_src) (3), f, k1, k2, k3, cx, cy."""
        b = self.R @ (ref.C - self.C)
        return np.r_[self.R.ravel(), b, [self.f, self.k1, self.k2, self.k3, self.cx, self.cy]]


def pinhole(C, f, w, h, R=NADIR) -> Camera:
    return Camera(np.asarray(R, np.float64), np.asarray(C, np.float64), float(f), (w - 1) / 2.0, (h - 1) / 2.0,
                  int(w), int(h))


@dataclass
class Terrain:
    cx: np.ndarray      # bump centres x (m)
    cy: np.ndarray
    height: np.ndarray  # bump heights (m)
    sigma: np.ndarray   # bump widths (m)
    hmax: float
    tex_seed: int
    octaves: Tuple[Tuple[float, float], ...] = ((8.0, 0.5), (2.0, 0.25), (0.5, 0.15), (0.13, 0.1))

    def height_at(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        """Sum of the bumps at (X, Y), clipped to [0, hmax] (all bumps in one broadcast)."""
        dev, dt = X.device, X.dtype
        cx = torch.as_tensor(self.cx, dtype=dt, device=dev)
        cy = torch.as_tensor(self.cy, dtype=dt, device=dev)
        h = torch.as_tensor(self.height, dtype=dt, device=dev)
        inv2s2 = torch.as_tensor(1.0 / (2.0 * np.asarray(self.sigma) ** 2), dtype=dt, device=dev)
        d2 = (X[..., None] - cx) ** 2 + (Y[..., None] - cy) ** 2
        z = (h * torch.exp(-d2 * inv2s2)).sum(-1)
        return torch.clamp(z, 0.0, self.hmax)


def make_terrain(seed: int, extent: Tuple[float, float, float, float], n_bumps: int = 20, hmax: float = 15.0,
                 sigma: Tuple[float, float] = (5.0, 20.0), tex_seed: Optional[int] = None) -> Terrain:
    """extent = (x0, x1, y0, y1) in metres over which the bump centres are drawn."""
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = extent
    return Terrain(rng.uniform(x0, x1, n_bumps), rng.uniform(y0, y1, n_bumps), rng.uniform(0.0, hmax, n_bumps),
                   rng.uniform(sigma[0], sigma[1], n_bumps), hmax, seed + 1 if tex_seed is None else tex_seed)


def _hash01(i: torch.Tensor, j: torch.Tensor, salt: int) -> torch.Tensor:
    """Deterministic lattice value in [0, 1) from integer coordinates (same on CPU and device)."""
    M = 0xFFFFFFFF
    h = (i * 374761393 + j * 668265263 + salt * 2246822519) & M
    h = ((h ^ (h >> 13)) * 1274126177) & M
    h = h ^ (h >> 16)
    return (h & 0xFFFFFF).to(torch.float64) / 16777216.0


def texture(terr: Terrain, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """Multi-octave value noise at ground positions, normalised by the total amplitude to [0, 1]."""
    out = torch.zeros_like(X)
    tot = 0.0
    for o, (wl, amp) in enumerate(terr.octaves):
        gx, gy = X / wl, Y / wl
        fx, fy = torch.floor(gx), torch.floor(gy)
        tx, ty = gx - fx, gy - fy
        sx, sy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
        i, j = fx.to(torch.int64), fy.to(torch.int64)
        salt = terr.tex_seed * 131 + o
        v00, v10 = _hash01(i, j, salt), _hash01(i + 1, j, salt)
        v01, v11 = _hash01(i, j + 1, salt), _hash01(i + 1, j + 1, salt)
        out = out + amp * ((v00 * (1 - sx) + v10 * sx) * (1 - sy) + (v01 * (1 - sx) + v11 * sx) * sy)
        tot += amp
    return out / tot


def camera_rays(cam: Camera, device, dtype=torch.float64) -> torch.Tensor:
    """(3, h, w) world directions with camera-axis depth 1; a radially distorted camera's pixels
    are undistorted by 20 fixed-point steps of x = n / (1 + r2 (k1 + r2 (k2 + r2 k3)))."""
    v, u = torch.meshgrid(torch.arange(cam.h, device=device, dtype=dtype),
                          torch.arange(cam.w, device=device, dtype=dtype), indexing="ij")
    nx, ny = (u - cam.cx) / cam.f, (v - cam.cy) / cam.f
    x, y = nx, ny
    if cam.k1 or cam.k2 or cam.k3:
        for _ in range(20):
            r2 = x * x + y * y
            d = 1 + r2 * (cam.k1 + r2 * (cam.k2 + r2 * cam.k3))
            x, y = nx / d, ny / d
    xc = torch.stack([x, y, torch.ones_like(u)])
    R = torch.as_tensor(cam.R, dtype=dtype, device=device)
    return torch.einsum("ji,jhw->ihw", R, xc)


def render(terr: Terrain, cam: Camera, noise_seed: int, noise: float = 0.01, device="cpu", march: int = 96,
           bisect: int = 40, rows_per_chunk: int = 1024):
    """-> (image (h, w) float32 in [0,1] + noise, depth (h, w) float32 camera-axis)."""
    img = torch.empty((cam.h, cam.w), dtype=torch.float32, device=device)
    dep = torch.empty((cam.h, cam.w), dtype=torch.float32, device=device)
    C = [float(c) for c in cam.C]
    full = camera_rays(cam, device)
    for r0 in range(0, cam.h, rows_per_chunk):
        d = full[:, r0:r0 + rows_per_chunk]
        dz = d[2]
        # bracket: from the plane z = hmax (above all terrain) to z = 0 (below or on all terrain)
        t_top = torch.clamp((C[2] - terr.hmax) / (-dz), min=0.0)
        t_bot = C[2] / (-dz)
        lo, hi = t_top.clone(), t_bot.clone()
        found = torch.zeros_like(lo, dtype=torch.bool)
        prev_t = t_top
        for s in range(1, march + 1):
            t = t_top + (t_bot - t_top) * (s / march)
            X, Y, Z = C[0] + t * d[0], C[1] + t * d[1], C[2] + t * dz
            below = Z <= terr.height_at(X, Y)
            hit = below & ~found
            lo = torch.where(hit, prev_t, lo)
            hi = torch.where(hit, t, hi)
            found = found | below
            prev_t = t
        for _ in range(bisect):
            m = 0.5 * (lo + hi)
            X, Y, Z = C[0] + m * d[0], C[1] + m * d[1], C[2] + m * dz
            below = Z <= terr.height_at(X, Y)
            hi = torch.where(below, m, hi)
            lo = torch.where(below, lo, m)
        t = 0.5 * (lo + hi)
        X, Y = C[0] + t * d[0], C[1] + t * d[1]
        img[r0:r0 + rows_per_chunk] = texture(terr, X, Y).to(torch.float32)
        dep[r0:r0 + rows_per_chunk] = t.to(torch.float32)
    g = torch.Generator(device="cpu").manual_seed(int(noise_seed))
    if noise > 0:
        img += (torch.randn((cam.h, cam.w), generator=g, dtype=torch.float32) * noise).to(device)
    return img, dep


@dataclass
class View:
    cam: Camera
    img: torch.Tensor
    depth: torch.Tensor


@dataclass
class Scene:
    terrain: Terrain
    ref: View
    srcs: List[View] = field(default_factory=list)

    def geo(self) -> np.ndarray:
        return np.stack([v.cam.geo18(self.ref.cam) for v in self.srcs])

    def par(self) -> np.ndarray:
        return np.stack([v.cam.par18() for v in self.srcs])


def make_scene(seed: int, W: int, H: int, f: float, alt: float, src_xy: Sequence[Tuple[float, float]],
               device="cpu", noise: float = 0.01, ref_xy=(0.0, 0.0), k=(0.0, 0.0, 0.0)) -> Scene:
    """Reference camera at (ref_xy, alt) and one source camera per (x, y), all nadir, all with the
    radial distortion k = (k1, k2, k3)."""
    half_w = (W / 2) / f * alt
    half_h = (H / 2) / f * alt
    xs = [p[0] for p in src_xy] + [ref_xy[0]]
    ys = [p[1] for p in src_xy] + [ref_xy[1]]
    ext = (min(xs) - half_w, max(xs) + half_w, min(ys) - half_h, max(ys) + half_h)
    terr = make_terrain(seed, ext)
    cams = [pinhole((ref_xy[0], ref_xy[1], alt), f, W, H)] + [pinhole((x, y, alt), f, W, H) for x, y in src_xy]
    for c in cams:
        c.k1, c.k2, c.k3 = (float(v) for v in k)
    views = []
    for k, cam in enumerate(cams):
        img, dep = render(terr, cam, noise_seed=seed * 1000 + 2 + k, noise=noise, device=device)
        views.append(View(cam, img, dep))
    return Scene(terr, views[0], views[1:])


def config_c1(device="cpu", scale: float = 1.0, k=(0.0, 0.0, 0.0)) -> Scene:
    """Plain plane sweep: 640x480, f=600, 80 m, 4 sources on a 15 m cross (seed 0); k: radial
    distortion of every camera (0 in the plain configuration)."""
    W, H = int(round(640 * scale)), int(round(480 * scale))
    b = 15.0
    return make_scene(0, W, H, 600.0 * scale, 80.0, [(b, 0.0), (-b, 0.0), (0.0, b), (0.0, -b)], device=device, k=k)


C2_SOURCES = [(x, y) for y in (-10.0, 10.0) for x in (-30.0, -10.0, 10.0, 30.0)]


def config_c2(device="cpu", scale: float = 1.0, k=(0.0, 0.0, 0.0)) -> Scene:
    """Single-view PatchMatch: 2000x1500, f=1800, 100 m, 8 sources on a 2x4 grid, 20 m baseline;
    k: radial distortion of every camera (0 in the configuration)."""
    W, H = int(round(2000 * scale)), int(round(1500 * scale))
    return make_scene(3, W, H, 1800.0 * scale, 100.0, C2_SOURCES, device=device, k=k)


def config_c2_cameras(scale: float = 1.0) -> Scene:
    """The cameras and terrain of config_c2 without rendering (images filled in by the caller)."""
    W, H = int(round(2000 * scale)), int(round(1500 * scale))
    f, alt = 1800.0 * scale, 100.0
    half_w, half_h = (W / 2) / f * alt, (H / 2) / f * alt
    xs = [p[0] for p in C2_SOURCES] + [0.0]
    ys = [p[1] for p in C2_SOURCES] + [0.0]
    terr = make_terrain(3, (min(xs) - half_w, max(xs) + half_w, min(ys) - half_h, max(ys) + half_h))
    cams = [pinhole((0.0, 0.0, alt), f, W, H)] + [pinhole((x, y, alt), f, W, H) for x, y in C2_SOURCES]
    views = [View(c, None, None) for c in cams]
    return Scene(terr, views[0], views[1:])


def pm_init(seed: int, H: int, W: int, ray: torch.Tensor, dmin: float, dmax: float, max_tilt_deg: float = 30.0):
    """Seeded random slanted planes: depth U(dmin, dmax), normals uniform in a cone of
    max_tilt_deg around up, flipped to face the camera.  -> (depth (H,W), normal (3,H,W)) float32."""
    rng = np.random.default_rng(seed)
    depth = rng.uniform(dmin, dmax, (H, W)).astype(np.float32)
    cmin = math.cos(math.radians(max_tilt_deg))
    cz = rng.uniform(cmin, 1.0, (H, W))
    phi = rng.uniform(0, 2 * math.pi, (H, W))
    sz = np.sqrt(1 - cz * cz)
    n = np.stack([sz * np.cos(phi), sz * np.sin(phi), cz])
    r = ray.detach().cpu().numpy().astype(np.float64)
    flip = (n * r).sum(0) > 0
    n[:, flip] *= -1
    return torch.from_numpy(depth), torch.from_numpy(np.ascontiguousarray(n, np.float32))


def rays_world(cam: Camera, device="cpu") -> torch.Tensor:
    """(3, h, w) float32 rays of the reference camera as densify._rays_world builds them for a
    pinhole camera (densify.py:288-297): x = (u - cx) / f, y = (v - cy) / f, then R^T."""
    return camera_rays(cam, device).to(torch.float32).contiguous()


# ---- multi-view survey blocks -------------------------------------------------------------------

C3_SIZES = ((800, 600), (1200, 900), (1600, 1200), (2400, 1800))


@dataclass
class Survey:
    """A block of nadir views over one terrain; nbrs[k] are the source views of view k."""
    terrain: Terrain
    views: List[View]
    nbrs: List[List[int]]
    dmin: float
    dmax: float

    def geo(self, k: int) -> np.ndarray:
        return np.stack([self.views[j].cam.geo18(self.views[k].cam) for j in self.nbrs[k]])

    def par(self, k: int) -> np.ndarray:
        return np.stack([self.views[j].cam.par18() for j in self.nbrs[k]])


def neighbours_by_distance(C: np.ndarray, alt: float, counts: Sequence[int], min_angle_deg: float = 8.0):
    """For each camera the counts[k] nearest others (ties by index) whose baseline subtends at least
    min_angle_deg at the flying height -- the >= 8 degree triangulation preference of
    densify._neighbours (densify.py:198-201) without tie-point counts."""
    out = []
    for k in range(len(C)):
        d = np.linalg.norm(C - C[k], axis=1)
        ang = np.degrees(d / alt)
        cand = [j for j in np.lexsort((np.arange(len(C)), d)) if j != k and ang[j] >= min_angle_deg]
        out.append([int(j) for j in cand[:counts[k]]])
    return out


def survey_cameras(seed: int, nx: int, ny: int, alt: float, sizes: Sequence[Tuple[int, int]], f_per_w: float,
                   forward: float, side: float, s_range: Tuple[int, int]):
    """nx x ny nadir grid with the given forward (x) / side (y) overlaps of the image footprint; each
    view's size drawn uniformly from `sizes`, f = f_per_w * W (same footprint for every size), and
    its neighbour count from U{s_range}.  -> (cameras, counts, terrain extent)."""
    rng = np.random.default_rng(seed)
    W0, H0 = sizes[0]
    foot_w, foot_h = alt / f_per_w, alt / f_per_w * H0 / W0
    dx, dy = (1 - forward) * foot_w, (1 - side) * foot_h
    cams, counts = [], []
    for iy in range(ny):
        for ix in range(nx):
            W, H = sizes[int(rng.integers(len(sizes)))]
            x, y = (ix - (nx - 1) / 2) * dx, (iy - (ny - 1) / 2) * dy
            cams.append(pinhole((x, y, alt), f_per_w * W, W, H))
            counts.append(int(rng.integers(s_range[0], s_range[1] + 1)))
    ext = (-(nx - 1) / 2 * dx - foot_w, (nx - 1) / 2 * dx + foot_w, -(ny - 1) / 2 * dy - foot_h,
           (ny - 1) / 2 * dy + foot_h)
    return cams, counts, ext


def make_survey(seed: int, cams: Sequence[Camera], counts: Sequence[int], ext, alt: float, device="cpu",
                noise: float = 0.01, hmax: float = 15.0, n_bumps: int = 40, march: int = 96,
                bisect: int = 40) -> Survey:
    terr = make_terrain(seed, ext, n_bumps=n_bumps, hmax=hmax)
    views = []
    for k, cam in enumerate(cams):
        img, dep = render(terr, cam, noise_seed=seed * 1000 + 2 + k, noise=noise, device=device, march=march,
                          bisect=bisect)
        views.append(View(cam, img, dep))
    C = np.stack([c.C for c in cams])
    nbrs = neighbours_by_distance(C, alt, counts)
    return Survey(terr, views, nbrs, dmin=alt - hmax - 2.0, dmax=alt + 2.0)


def noisy_depths(survey: Survey, seed: int, rel_noise: float = 2e-3, hole_frac: float = 0.03, n_patches: int = 4):
    """Depth maps as a matcher leaves them, for fusion: each view's rendered camera-axis depth
    x (1 + N(0, rel_noise)), with random single-pixel and rectangular NaN holes (unmatched or masked
    pixels).  -> list of (h, w) float32 numpy arrays."""
    rng = np.random.default_rng(seed)
    out = []
    for v in survey.views:
        d = v.depth.cpu().numpy().astype(np.float64)
        d = (d * (1.0 + rel_noise * rng.standard_normal(d.shape))).astype(np.float32)
        d[rng.random(d.shape) < hole_frac] = np.nan
        h, w = d.shape
        for _ in range(n_patches):
            y0, x0 = int(rng.integers(h)), int(rng.integers(w))
            d[y0:y0 + h // 8, x0:x0 + w // 8] = np.nan
        out.append(d)
    return out


def config_fuse(device="cpu", scale: float = 1.0) -> Survey:
    """Fusion block (seed 11): 8 x 6 nadir grid at 90 m with 70 % forward / 60 % side overlap, 48
    views of 1600x1200 (the reference's default matching size, DepthOptions.max_image_dim), f =
    0.9375 W, 4 nearest views each (DepthOptions.neighbors)."""
    sizes = [(int(round(1600 * scale)), int(round(1200 * scale)))]
    cams, counts, ext = survey_cameras(11, 8, 6, 90.0, sizes, 0.9375, 0.7, 0.6, (4, 4))
    return make_survey(11, cams, counts, ext, 90.0, device=device, march=48, bisect=30, n_bumps=24)


def config_c3(device="cpu", scale: float = 1.0) -> Survey:
    """Batched block: 8 x 6 nadir grid at 90 m with 70 % forward / 60 % side overlap, 48 views whose
    sizes are drawn from 800x600 .. 2400x1800 (f = 0.9375 W), S = U{4..12} nearest views (seed 4)."""
    sizes = [(int(round(w * scale)), int(round(h * scale))) for w, h in C3_SIZES]
    cams, counts, ext = survey_cameras(4, 8, 6, 90.0, sizes, 0.9375, 0.7, 0.6, (4, 12))
    return make_survey(4, cams, counts, ext, 90.0, device=device)


C5_CLASSES = ((1200, 900, 0.6), (2400, 1800, 0.3), (4000, 3000, 0.1))


def config_c5_pools(device="cpu", scale: float = 1.0, nx: int = 4, ny: int = 4) -> List[Survey]:
    """Served depth maps (seed 6): one nadir block per image size class (1200x900, 2400x1800,
    4000x3000; f = 0.9375 W, 90 m, 70/60 % overlap), nx x ny views each, from which requests draw a
    reference and its nearest views.  A lighter terrain (16 bumps, 64 march + 30 bisection steps)
    keeps the render of the 280 MP of pools short."""
    out = []
    for c, (w, h, _) in enumerate(C5_CLASSES):
        size = (int(round(w * scale)), int(round(h * scale)))
        cams, counts, ext = survey_cameras(6 + c, nx, ny, 90.0, [size], 0.9375, 0.7, 0.6, (12, 12))
        out.append(make_survey(6 + c, cams, counts, ext, 90.0, device=device, n_bumps=16, march=64, bisect=30))
    return out


def config_c4(device="cpu", scale: float = 1.0) -> Scene:
    """Large view at reduced precision (seed 5): a 4000x3000 reference (f = 0.9375 W, 90 m) at the
    centre of a 5 x 3 nadir grid (70/60 % overlap) with its 12 nearest views as sources."""
    W, H = int(round(4000 * scale)), int(round(3000 * scale))
    cams, counts, ext = survey_cameras(5, 5, 3, 90.0, [(W, H)], 0.9375, 0.7, 0.6, (12, 12))
    C = np.stack([c.C for c in cams])
    k = 7                                           # the centre of the 5 x 3 grid
    nb = neighbours_by_distance(C, 90.0, [12] * len(cams))[k]
    terr = make_terrain(5, ext, n_bumps=16)
    views = []
    for j in [k] + nb:
        img, dep = render(terr, cams[j], noise_seed=5000 + j, device=device, march=64, bisect=30)
        views.append(View(cams[j], img, dep))
    return Scene(terr, views[0], views[1:])

