"""Writes tests/golden/depthmap_*.npz: vectors of the reference's densify.depth_map and its
planning helpers (_neighbours, _depth_range, _depth_band, _num_depths).

None of these has a reference test or a closed form, so the vectors come from the reference
itself: this script compiles the reference's Cython modules (orthomosaic/*.pyx, the flags of
setup.py:10-34) into a temporary directory, links the reference's .py files beside them, imports
orthomosaic.densify in this process and records its outputs on a small seeded synthetic survey.
Only the arrays are committed.  Needs the reference tree and Cython; run from the repository root:

    python tests/golden/make_depthmap_golden.py [/path/to/reference]
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import sysconfig
import tempfile
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def build_reference_package(ref_root: str, out_dir: str) -> str:
    """Compiled reference package in out_dir/orthomosaic (symlinked sources + built modules)."""
    src = os.path.join(ref_root, "orthomosaic")
    pkg = os.path.join(out_dir, "orthomosaic")
    os.makedirs(pkg)
    for f in glob.glob(os.path.join(src, "*.py")) + glob.glob(os.path.join(src, "*.h")):
        os.symlink(f, os.path.join(pkg, os.path.basename(f)))
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    for pyx in sorted(glob.glob(os.path.join(src, "*.pyx"))):
        name = os.path.basename(pyx)[:-4]
        c_file = os.path.join(out_dir, name + ".c")
        subprocess.run([sys.executable, "-m", "cython", "-3", "-o", c_file, pyx], check=True,
                       stdout=subprocess.DEVNULL)
        flags = ["-O3", "-fno-math-errno"]
        if name == "_fast" or name == "_ba":
            flags.append("-ffp-contract=off")
        cmd = ["gcc", "-shared", "-fPIC", *flags, "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
               f"-I{sysconfig.get_paths()['include']}", f"-I{np.get_include()}", f"-I{src}",
               "-o", os.path.join(pkg, name + suffix), c_file, "-lm", "-w"]
        subprocess.run(cmd, check=True)
    return out_dir


def _rotation(rng, deg: float) -> np.ndarray:
    """Rz(c) Ry(b) Rx(a) with a, b, c uniform in [-deg, deg] degrees."""
    a, b, c = np.radians(rng.uniform(-deg, deg, 3))
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def synthetic_survey(seed: int, W: int, H: int, f: float, alt: float, n_cam_x: int, n_cam_y: int, spacing: float,
                     n_pts: int = 4000, rot_deg: float = 0.0):
    """Grid of cameras over the seeded terrain, nadir or (rot_deg > 0) with each camera turned by a
    seeded small rotation; tie points = terrain points seen by >= 2 views."""
    sys.path.insert(0, ROOT)
    import torch
    from py_ortho_amd import scene
    xs = [(i - (n_cam_x - 1) / 2) * spacing for i in range(n_cam_x)]
    ys = [(j - (n_cam_y - 1) / 2) * spacing for j in range(n_cam_y)]
    centres = [(x, y) for y in ys for x in xs]
    half_w, half_h = (W / 2) / f * alt, (H / 2) / f * alt
    ext = (min(xs) - half_w, max(xs) + half_w, min(ys) - half_h, max(ys) + half_h)
    terr = scene.make_terrain(seed, ext)
    cams = [scene.pinhole((x, y, alt), f, W, H) for x, y in centres]
    if rot_deg > 0:
        rrng = np.random.default_rng(seed + 300)
        for cam in cams:
            cam.R = _rotation(rrng, rot_deg) @ scene.NADIR
    imgs, deps = [], []
    for k, cam in enumerate(cams):
        img, dep = scene.render(terr, cam, noise_seed=seed * 1000 + k, device="cpu")
        imgs.append(img.numpy())
        deps.append(dep.numpy())
    rng = np.random.default_rng(seed + 100)
    px = rng.uniform(ext[0] + half_w * 0.3, ext[1] - half_w * 0.3, n_pts)
    py = rng.uniform(ext[2] + half_h * 0.3, ext[3] - half_h * 0.3, n_pts)
    pz = terr.height_at(torch.from_numpy(px), torch.from_numpy(py)).numpy()
    X = np.stack([px, py, pz], 1)
    obs_cam, obs_pt = [], []
    for k, cam in enumerate(cams):
        xc = (X - cam.C) @ cam.R.T
        u = cam.f * xc[:, 0] / xc[:, 2] + cam.cx
        v = cam.f * xc[:, 1] / xc[:, 2] + cam.cy
        ok = (xc[:, 2] > 0) & (u >= 0) & (v >= 0) & (u < W) & (v < H)
        idx = np.flatnonzero(ok)
        obs_cam.append(np.full(len(idx), k, np.int32))
        obs_pt.append(idx.astype(np.int32))
    obs_cam, obs_pt = np.concatenate(obs_cam), np.concatenate(obs_pt)
    N = len(cams)
    intr = [SimpleNamespace(f=f, cx=(W - 1) / 2, cy=(H - 1) / 2, k1=0.0, k2=0.0, k3=0.0, width=W, height=H)]
    rec = SimpleNamespace(used=list(range(N)), R=np.stack([c.R for c in cams]), C=np.stack([c.C for c in cams]),
                          X=X, obs_cam=obs_cam, obs_pt=obs_pt, intr=intr, cam_group=np.zeros(N, np.int32))
    grey = [((255.0 * im - 128.0) / 64.0).astype(np.float32) for im in imgs]
    mrng = np.random.default_rng(seed + 7)
    for g in grey:                           # a few masked pixels (NaN), as load_mask produces
        g[mrng.random(g.shape) < 0.003] = np.nan
    return rec, grey, imgs, deps, cams


def main(ref_root: str) -> None:
    specs = [  # name, seed, W, H, f, alt, grid, spacing, opt overrides, reference views
        ("grid3x3", 21, 240, 180, 220.0, 60.0, (3, 3), 14.0, dict(num_depths=32, window=3, neighbors=4), (4, 0)),
    ]
    with tempfile.TemporaryDirectory() as td:
        build_reference_package(ref_root, td)
        sys.path.insert(0, td)
        from orthomosaic import densify as D
        for name, seed, W, H, f, alt, (gx, gy), spacing, ov, views in specs:
            rec, grey, imgs, deps, cams = synthetic_survey(seed, W, H, f, alt, gx, gy, spacing)
            opt = D.DepthOptions(**ov)
            dcams = [D._Cam(c.R, c.C, c.f, 0.0, 0.0, c.cx, c.cy, W, H, 0.0) for c in cams]

            def view(j):
                return grey[j], None, dcams[j]

            nbrs = D._neighbours(rec, opt.neighbors)
            out = dict(grey=np.stack(grey), R=rec.R, C=rec.C, X=rec.X, obs_cam=rec.obs_cam, obs_pt=rec.obs_pt,
                       intr=np.array([f, (W - 1) / 2, (H - 1) / 2, W, H], np.float64),
                       opt=np.array([opt.num_depths, opt.window, opt.neighbors, opt.refine_steps, opt.top_k,
                                     opt.min_texture], np.float64),
                       nbrs=np.array([nb + [-1] * (opt.neighbors - len(nb)) for nb in nbrs], np.int32),
                       views=np.array(views, np.int32))
            for k in views:
                rng = D._depth_range(rec, k)
                res = D.depth_map(k, nbrs[k], rec, view, opt)
                assert res is not None
                depth, score, low = res
                Lc = dcams[k].scaled(0.5, W // 2, H // 2).scaled(0.5, W // 4, H // 4)
                nd = D._num_depths(rec, k, nbrs[k], rng, Lc.f, opt.num_depths)
                shape = (H // 4, W // 4)
                inv_lo, inv_hi = D._depth_band(rec, k, Lc, shape, rng, block=max(8, shape[1] // 16))
                ray = D._rays_world((H, W), dcams[k])
                normals = D._init_normals(depth, ray)
                out.update({f"rng_{k}": np.array(rng), f"nd_{k}": np.int64(nd), f"inv_lo_{k}": inv_lo,
                            f"inv_hi_{k}": inv_hi, f"depth_{k}": depth, f"score_{k}": score, f"low_{k}": low,
                            f"ray_{k}": ray, f"normals_{k}": normals})
                fin = np.isfinite(depth)
                rel = np.abs(depth[fin] - deps[k][fin]) / deps[k][fin]
                print(name, k, "nd", nd, "finite", float(fin.mean()), "within 1% of truth", float((rel < 0.01).mean()))
            path = os.path.join(HERE, f"depthmap_{name}.npz")
            np.savez_compressed(path, **out)
            print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
