"""Host-side planning of one depth map from the sparse reconstruction (NumPy, tie points only).

These are the reference's small NumPy steps that choose what the kernels search; they run on the
host in milliseconds and feed the device sweep:

- `neighbours`   source views per image (densify._neighbours, densify.py:183-205)
- `depth_range`  depth search range of an image (densify._depth_range, densify.py:208-230)
- `depth_band`   per-block inverse-depth band (densify._depth_band, densify.py:233-274)
- `num_depths`   coarse hypothesis count (densify._num_depths, densify.py:277-284)

`rec` is any object with the fields of sfm.Reconstruction that these use (sfm.py:35-61): `used`,
`R` (N,3,3) world->camera, `C` (N,3), `X` (P,3) tie points, `obs_cam`, `obs_pt` (observations),
`intr` (per camera group: f, cx, cy, width, height) and `cam_group`.
"""
from __future__ import annotations

import numpy as np


def neighbours(rec, n_nb: int):
    """For every image: up to n_nb other images sharing >= 15 tie points, ranked by shared points
    weighted by the triangulation-angle preference (>= 8 deg, penalised above 45 deg)."""
    N = len(rec.used)
    shared = np.zeros((N, N), np.int32)
    order = np.argsort(rec.obs_pt, kind="stable")
    pts, cams = rec.obs_pt[order], rec.obs_cam[order]
    cuts = np.flatnonzero(np.diff(pts)) + 1
    for grp in np.split(cams, cuts):
        if len(grp) > 1:
            g = np.unique(grp)
            shared[np.ix_(g, g)] += 1
    np.fill_diagonal(shared, 0)
    med_depth = np.empty(N)
    for k in range(N):
        mine = rec.obs_cam == k
        if np.any(mine):
            med_depth[k] = np.median((rec.R[k] @ (rec.X[rec.obs_pt[mine]] - rec.C[k]).T)[2])
        else:
            med_depth[k] = 1.0
    out = []
    for k in range(N):
        base = np.linalg.norm(rec.C - rec.C[k], axis=1)
        ang = np.degrees(base / max(med_depth[k], 1e-3))
        w = np.clip(ang / 8.0, 0, 1) ** 2 * np.where(ang > 45, 0.2, 1.0)
        score = shared[k] * w
        ranked = [j for j in np.argsort(-score) if score[j] > 0 and shared[k, j] >= 15]
        out.append(ranked[:n_nb])
    return out


def depth_range(rec, k):
    """(near, far) of image k from all tie points inside its view, extended down to the lowest
    ground of the block along the view axis; None with fewer than 10 such points."""
    it = rec.intr[rec.cam_group[k]]
    xc = (rec.X - rec.C[k]) @ rec.R[k].T
    z = xc[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        u = it.f * xc[:, 0] / z + it.cx
        v = it.f * xc[:, 1] / z + it.cy
    seen = (z > 0) & (u >= 0) & (v >= 0) & (u < it.width) & (v < it.height)
    d = z[seen]
    if len(d) < 10:
        return None
    lo, hi = np.percentile(d, [1, 99])
    zlow = float(np.percentile(rec.X[:, 2], 0.5))
    axis_down = max(-float(rec.R[k][2, 2]), 0.2)
    hi = max(hi, (float(rec.C[k][2]) - zlow) / axis_down)
    span = hi - lo
    return max(0.3 * lo, lo - 0.2 * span - 1.0), hi + 0.2 * span + 2.0


def depth_band(rec, k, cam, shape, rng, block: int = 32, margin: float = 1.0):
    """Per-pixel (inv_lo, inv_hi) float32 maps at `shape`: per block of `block` px the depth range of
    image k's own tie points (>= 5 per block), widened over the 3x3 block neighbourhood, by `margin`
    m and by 5 % of the range, clipped to `rng`; other blocks keep `rng`."""
    H, W = shape
    mine = rec.obs_cam == k
    lo_g, hi_g = rng
    inv_lo = np.full((H, W), 1.0 / hi_g, np.float32)
    inv_hi = np.full((H, W), 1.0 / lo_g, np.float32)
    if mine.sum() < 50:
        return inv_lo, inv_hi
    xc = (rec.X[rec.obs_pt[mine]] - cam.C) @ cam.R.T
    front = xc[:, 2] > 0
    u, v = cam.project_cam(xc[front, 0], xc[front, 1], xc[front, 2])
    z = xc[front, 2]
    gh, gw = -(-H // block), -(-W // block)
    bid = (np.clip(v, 0, H - 1) // block).astype(int) * gw + (np.clip(u, 0, W - 1) // block).astype(int)
    inside = (u >= 0) & (v >= 0) & (u < W) & (v < H)
    bid, z = bid[inside], z[inside]
    zlo = np.full(gh * gw, np.inf)
    zhi = np.full(gh * gw, -np.inf)
    cnt = np.bincount(bid, minlength=gh * gw)
    np.minimum.at(zlo, bid, z)
    np.maximum.at(zhi, bid, z)
    ok = (cnt >= 5).reshape(gh, gw)
    zlo, zhi = zlo.reshape(gh, gw), zhi.reshape(gh, gw)
    plo = np.pad(np.where(ok, zlo, np.inf), 1, constant_values=np.inf)
    phi = np.pad(np.where(ok, zhi, -np.inf), 1, constant_values=-np.inf)
    lo3 = np.min([plo[1 + dy:gh + 1 + dy, 1 + dx:gw + 1 + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1)], 0)
    hi3 = np.max([phi[1 + dy:gh + 1 + dy, 1 + dx:gw + 1 + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1)], 0)
    has = ok & np.isfinite(lo3) & np.isfinite(hi3)
    pad = margin + 0.05 * (hi3 - lo3)
    lo_b = np.where(has, np.clip(lo3 - pad, lo_g, hi_g), lo_g)
    hi_b = np.where(has, np.clip(hi3 + pad, lo_g, hi_g), hi_g)
    lo_px = np.repeat(np.repeat(lo_b, block, 0), block, 1)[:H, :W]
    hi_px = np.repeat(np.repeat(hi_b, block, 0), block, 1)[:H, :W]
    return (1.0 / hi_px).astype(np.float32), (1.0 / np.maximum(lo_px, 1e-6)).astype(np.float32)


def num_depths(rec, k, nbrs_k, rng, f_coarse, base: int) -> int:
    """About one hypothesis per coarse-level pixel of disparity for the median baseline, clipped to
    [base, 3 base]."""
    if not nbrs_k:
        return base
    B = float(np.median([np.linalg.norm(rec.C[k] - rec.C[j]) for j in nbrs_k]))
    disp = f_coarse * B * (1.0 / rng[0] - 1.0 / rng[1])
    return int(np.clip(np.ceil(disp), base, 3 * base))
