"""Writes tests/golden/pm_refine_*.npz: input/output vectors of the reference's _dense.pm_refine.

pm_refine (orthomosaic/_dense.pyx:566-666) has no test in the reference and no closed form pins
its update order or random stream, so its vectors come from the reference itself: this script
compiles orthomosaic/_dense.pyx with Cython into a temporary directory (gcc -O3 -fno-math-errno,
the flags of setup.py:11), imports it in this process and records pm_refine on small seeded
scenes.  Only the resulting arrays are committed.  Needs the reference tree (default
/root/reference) and Cython; run from the repository root:

    python tests/golden/make_pm_golden.py [/path/to/reference]
"""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys
import sysconfig
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build_reference_dense(ref_root: str, out_dir: str):
    src = os.path.join(ref_root, "orthomosaic", "_dense.pyx")
    c_file = os.path.join(out_dir, "_dense.c")
    subprocess.run([sys.executable, "-m", "cython", "-3", "-o", c_file, src], check=True)
    so = os.path.join(out_dir, "_dense" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["gcc", "-shared", "-fPIC", "-O3", "-fno-math-errno", "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
           f"-I{sysconfig.get_paths()['include']}", f"-I{np.get_include()}",
           f"-I{os.path.join(ref_root, 'orthomosaic')}", "-o", so, c_file, "-lm"]
    subprocess.run(cmd, check=True)
    spec = importlib.util.spec_from_file_location("_dense", so)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def value_noise(rng, X, Y, octaves=((4.0, 0.5), (1.0, 0.3), (0.25, 0.2))):
    out = np.zeros_like(X)
    for wl, amp in octaves:
        n = 64
        lat = rng.random((n + 1, n + 1))
        gx, gy = X / wl + n / 2, Y / wl + n / 2
        i, j = np.floor(gx).astype(int) % n, np.floor(gy).astype(int) % n
        fx, fy = gx - np.floor(gx), gy - np.floor(gy)
        sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)
        v = (lat[j, i] * (1 - sx) + lat[j, i + 1] * sx) * (1 - sy) + (lat[j + 1, i] * (1 - sx) + lat[j + 1, i + 1] * sx) * sy
        out += amp * v
    return out


def nadir(yaw=0.0, tilt=(0.0, 0.0)):
    R = np.diag([1.0, -1.0, -1.0])
    c, s = np.cos(yaw), np.sin(yaw)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    ax, ay = tilt
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    Ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    return Rx @ Ry @ R @ Rz


def render_plane(rng_tex_seed, R, C, f, h, w, cx, cy, plane_z=lambda X, Y: 0.02 * X + 0.01 * Y):
    """Camera (R world->camera, C) looking at the surface z = plane_z(x, y): intensity and depth."""
    v, u = np.mgrid[0:h, 0:w].astype(np.float64)
    d_c = np.stack([(u - cx) / f, (v - cy) / f, np.ones_like(u)])
    d_w = np.einsum("ji,jhw->ihw", R, d_c)
    # intersect C + t d with z = a x + b y (a, b from plane_z at two points)
    a = plane_z(1.0, 0.0) - plane_z(0.0, 0.0)
    b = plane_z(0.0, 1.0) - plane_z(0.0, 0.0)
    c0 = plane_z(0.0, 0.0)
    t = (c0 + a * C[0] + b * C[1] - C[2]) / (d_w[2] - a * d_w[0] - b * d_w[1])
    X, Y = C[0] + t * d_w[0], C[1] + t * d_w[1]
    tex = value_noise(np.random.default_rng(rng_tex_seed), X, Y)
    return tex.astype(np.float32), t.astype(np.float32), d_w


def make_case(seed: int, H: int, W: int, S: int, with_depths: bool):
    rng = np.random.default_rng(seed)
    f = 0.9 * W
    alt = 20.0
    cams = []
    for s in range(S + 1):
        if s == 0:
            C = np.array([0.0, 0.0, alt])
            R = nadir(0.0, (0.0, 0.0))
        else:
            ang = 2 * np.pi * s / S
            C = np.array([3.0 * np.cos(ang), 2.5 * np.sin(ang), alt + rng.normal(0, 0.3)])
            R = nadir(rng.normal(0, 0.05), (rng.normal(0, 0.03), rng.normal(0, 0.03)))
        h, w = H + int(rng.integers(-2, 3)), W + int(rng.integers(-2, 3))
        cams.append((R, C, f * rng.uniform(0.95, 1.05), (w - 1) / 2, (h - 1) / 2, h, w))
    imgs, deps = [], []
    for R, C, fs, cx, cy, h, w in cams:
        img, t, _ = render_plane(7, R, C, fs, h, w, cx, cy)
        img = img + rng.normal(0, 0.01, img.shape).astype(np.float32)
        imgs.append(np.ascontiguousarray(img, np.float32))
        deps.append(np.ascontiguousarray(t, np.float32))
    R0, C0, f0, cx0, cy0, h0, w0 = cams[0]
    ref = imgs[0][:H, :W].copy() if imgs[0].shape[0] >= H and imgs[0].shape[1] >= W else None
    # the reference view uses exactly (H, W)
    ref, _, d_w = render_plane(7, R0, C0, f0, H, W, (W - 1) / 2, (H - 1) / 2)
    ref = ref + rng.normal(0, 0.01, ref.shape).astype(np.float32)
    ref[rng.random((H, W)) < 0.01] = np.nan                     # masked pixels
    ray = np.ascontiguousarray(d_w, np.float32)
    srcs = imgs[1:]
    sdeps = [deps[s] if with_depths and (s % 3 != 2) else None for s in range(1, S + 1)]
    par = np.array([np.r_[R.ravel(), C, fs, 0.0, 0.0, 0.0, cx, cy] for R, C, fs, cx, cy, h, w in cams[1:]], np.float64)
    # initial state: noisy depth (some NaN), normals within ~25 deg of up, facing the camera
    true_d = render_plane(7, R0, C0, f0, H, W, (W - 1) / 2, (H - 1) / 2)[1]
    depth = (true_d * rng.uniform(0.9, 1.1, (H, W))).astype(np.float32)
    depth[rng.random((H, W)) < 0.15] = np.nan
    nrm = rng.normal(0, 0.25, (3, H, W))
    nrm[2] = 1.0
    nrm /= np.linalg.norm(nrm, axis=0, keepdims=True)
    normal = np.ascontiguousarray(nrm, np.float32)
    score = np.full((H, W), -1.0, np.float32)
    active = (rng.random((H, W)) < 0.8).astype(np.uint8)
    dmin, dmax = float(np.nanmin(true_d) * 0.85), float(np.nanmax(true_d) * 1.15)
    return dict(ref=ref.astype(np.float32), ray=ray, Cr=np.ascontiguousarray(C0, np.float64), srcs=srcs,
                src_depths=sdeps, par=par, depth=depth, normal=normal, score=score, active=active, dmin=dmin,
                dmax=dmax)


def main(ref_root: str) -> None:
    with tempfile.TemporaryDirectory() as td:
        dense = build_reference_dense(ref_root, td)
        specs = [  # name, seed, H, W, S, with_depths, r, step, top_k, geo_w, iters
            ("s3", 1, 40, 48, 3, True, 4, 2, 2, 0.5, (1, 2, 3)),
            ("s8", 2, 30, 36, 8, True, 5, 2, 2, 0.5, (1, 2)),
            ("s5_nogeo", 3, 33, 41, 5, False, 3, 1, 3, 0.0, (1, 2)),
        ]
        for name, seed, H, W, S, wd, r, step, top_k, geo_w, iter_list in specs:
            c = make_case(seed, H, W, S, wd)
            out = {}
            for iters in iter_list:
                d, n, s = c["depth"].copy(), c["normal"].copy(), c["score"].copy()
                ch = dense.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], d, n, s,
                                     c["active"], r, step, top_k, geo_w, 0.02, iters, c["dmin"], c["dmax"],
                                     (seed * 2654435761) % 2**32)
                out[f"depth_it{iters}"], out[f"normal_it{iters}"], out[f"score_it{iters}"] = d, n, s
                out[f"changed_it{iters}"] = np.int64(ch)
            arrays = dict(ref=c["ref"], ray=c["ray"], Cr=c["Cr"], par=c["par"], depth0=c["depth"],
                          normal0=c["normal"], score0=c["score"], active=c["active"],
                          params=np.array([r, step, top_k, geo_w, 0.02, c["dmin"], c["dmax"],
                                           (seed * 2654435761) % 2**32], np.float64),
                          iters=np.array(iter_list, np.int64),
                          has_depth=np.array([d is not None for d in c["src_depths"]]))
            for s, img in enumerate(c["srcs"]):
                arrays[f"src{s}"] = img
                if c["src_depths"][s] is not None:
                    arrays[f"dep{s}"] = c["src_depths"][s]
            arrays.update(out)
            path = os.path.join(HERE, f"pm_refine_{name}.npz")
            np.savez_compressed(path, **arrays)
            print(path, {k: int(v) for k, v in out.items() if k.startswith("changed")})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
