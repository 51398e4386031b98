"""Seeded random inputs for the depth-map kernels (shared by the CPU and GPU tests).

The sweep cases follow the shape of the reference's own parity cases (test_native_parity.py:36-72):
random reference and sources of slightly different sizes with NaN (masked) source pixels, random
near-nadir rotations, radial distortion, per-pixel inverse-depth maps, plus one infinite inverse
depth (depth 0) and one behind-the-camera hypothesis.
"""
from __future__ import annotations

import numpy as np


def rot(rng, tilt=0.15):
    """Rotation near a 180-degree turn about x (a nadir camera), Rodrigues form."""
    w = rng.normal(0, tilt, 3) + np.array([np.pi, 0, 0])
    th = np.linalg.norm(w)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def box_stats(ref, r):
    """mu, sd of the zero-padded (2r+1)^2 window by direct summation in float32."""
    H, W = ref.shape
    k = 2 * r + 1
    pad = np.pad(ref, r)
    mu = np.zeros_like(ref)
    m2 = np.zeros_like(ref)
    for dy in range(k):
        for dx in range(k):
            mu += pad[dy:dy + H, dx:dx + W]
            m2 += pad[dy:dy + H, dx:dx + W] ** 2
    mu /= k * k
    sd = np.sqrt(np.maximum(m2 / (k * k) - mu * mu, 1e-6)).astype(np.float32)
    return mu.astype(np.float32), sd


def sweep_case(rng, H, W, n, r, nhyp, nan_frac=0.02, distort=True):
    """-> ref, mu, sd, srcs, ray, geo (n, 18), hyps (list of (H, W) inverse-depth maps)."""
    ref = rng.normal(0, 1, (H, W)).astype(np.float32)
    srcs = []
    for _ in range(n):
        h, w = H + int(rng.integers(-3, 4)), W + int(rng.integers(-3, 4))
        s = rng.normal(0, 1, (max(h, 2), max(w, 2))).astype(np.float32)
        s[rng.random(s.shape) < nan_frac] = np.nan
        srcs.append(np.ascontiguousarray(s))
    mu, sd = box_stats(ref, r)
    ray = rng.normal(0, 0.3, (3, H, W)).astype(np.float32)
    ray[2] = 1.0 + rng.normal(0, 0.01, (H, W)).astype(np.float32)
    geo = []
    for _ in range(n):
        Rs = rot(rng, 0.05) @ rot(rng, 0.0).T
        b = rng.normal(0, 0.5, 3)
        f = float(rng.uniform(0.8, 1.2) * W)
        ks = [float(rng.normal(0, 0.05)), float(rng.normal(0, 0.01)), float(rng.normal(0, 0.003))] if distort \
            else [0.0, 0.0, 0.0]
        cxy = [(W - 1) / 2 + float(rng.normal(0, 2)), (H - 1) / 2 + float(rng.normal(0, 2))]
        geo.append(np.r_[Rs.ravel(), b, [f], ks, cxy])
    geo = np.ascontiguousarray(geo, np.float64)
    base = rng.uniform(0.02, 0.2, (H, W)).astype(np.float32)
    step = rng.uniform(1e-4, 5e-3, (H, W)).astype(np.float32)
    hyps = [(base + i * step).astype(np.float32) for i in range(nhyp)]
    if nhyp > 2:
        hyps[1][0, 0] = np.inf        # depth 0
        hyps[2][-1, -1] = -0.01       # behind the camera
    return ref, mu, sd, srcs, ray, geo, hyps


# (H, W, n, r, top_k, n_hyp): the shapes of the reference's parity test plus larger tiles
SWEEP_SHAPES = [(37, 53, 4, 5, 2, 9), (64, 48, 1, 3, 2, 5), (11, 9, 3, 5, 2, 4), (5, 4, 2, 5, 1, 3),
                (30, 31, 5, 2, 3, 6), (2, 40, 2, 1, 2, 3), (40, 3, 4, 4, 4, 3)]


def np_init_normals(D, ray):
    """densify._init_normals (densify.py:481-496) in NumPy's float32 operations, each written out
    in the order NumPy evaluates them (np.cross products then differences, the norm as
    (x0^2 + x1^2) + x2^2); equal bit for bit to the reference's expression on the golden survey
    (tests/test_oracle.py)."""
    D = np.asarray(D, np.float32)
    ray = np.asarray(ray, np.float32)
    P = ray * np.where(np.isfinite(D), D, np.nan)[None]
    tu = np.full_like(P, np.nan)
    tv = np.full_like(P, np.nan)
    tu[:, :, 2:-2] = P[:, :, 4:] - P[:, :, :-4]
    tv[:, 2:-2, :] = P[:, 4:, :] - P[:, :-4, :]
    n0 = tu[1] * tv[2] - tu[2] * tv[1]
    n1 = tu[2] * tv[0] - tu[0] * tv[2]
    n2 = tu[0] * tv[1] - tu[1] * tv[0]
    with np.errstate(invalid="ignore"):
        nn = np.sqrt((n0 * n0 + n1 * n1) + n2 * n2)
        bad = ~np.isfinite(nn) | (nn < 1e-12)
        d = np.where(bad, np.float32(1), nn)
        n = np.stack([n0 / d, n1 / d, n2 / d])
    n[:, bad] = np.array([0.0, 0.0, 1.0], np.float32)[:, None]
    flip = ((n[0] * ray[0] + n[1] * ray[1]) + n[2] * ray[2]) > 0
    n[:, flip] *= -1
    return np.ascontiguousarray(n, np.float32)
