"""Writes tests/golden/densify_grid3x3.npz: the reference's whole densification (densify.densify,
densify.py:499-662) on a seeded 3 x 3 survey, with PatchMatch repair on (setting "pm") and off
(setting "plain"): the dense cloud and every view's final filtered depth map, and for the PatchMatch run also the
pass-1 maps (p, s, t), read from the kept depth-map files.

densify reads images through imageio.load_rgb and masks through masks.load_mask; here both names
in the reference's densify module are pointed at the seeded inputs below (uint8 RGB views, one view
with a masked rectangle), which are committed with the outputs.  Gains and biases are set for some
views.  Only arrays are committed.  Needs the reference tree and Cython; run from the repository
root:

    python tests/golden/make_densify_golden.py [/path/to/reference]
"""
from __future__ import annotations

import glob
import os
import sys
import tempfile
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_depthmap_golden import build_reference_package, synthetic_survey  # noqa: E402

SEED, W, H, F, ALT, GRID, SPACING = 31, 128, 96, 110.0, 60.0, (3, 3), 13.0
SETTINGS = {
    "pm": dict(num_depths=32, window=3, neighbors=4, patchmatch=True, pm_iters=2, pm_window=3, point_stride=2),
    "plain": dict(num_depths=24, window=2, neighbors=3, patchmatch=False, point_stride=2, min_views=2),
}


def inputs():
    """(rec, rgb list, masks {path: bool array}, gains, biases, frames)."""
    rec, grey, imgs, deps, cams = synthetic_survey(SEED, W, H, F, ALT, GRID[0], GRID[1], SPACING, rot_deg=3.0)
    rng = np.random.default_rng(SEED + 5)
    rgb = []
    for im in imgs:
        ch = [im * 255.0 * a + b for a, b in ((0.95, 4.0), (1.0, 0.0), (0.9, 9.0))]
        x = np.stack(ch, -1) + rng.normal(0, 1.0, im.shape + (3,))
        rgb.append(np.clip(x + 0.5, 0, 255).astype(np.uint8))
    masks = {}
    m = np.ones((H, W), bool)
    m[10:30, 50:80] = False
    masks["view2.jpg"] = m
    gains = {1: np.array([1.05, 0.97, 1.0], np.float32), 4: np.array([0.9, 1.1, 1.02], np.float32)}
    biases = {4: np.array([2.0, -3.0, 0.5], np.float32), 6: np.array([-1.0, 1.0, 0.0], np.float32)}
    frames = [SimpleNamespace(width=W, height=H, path=f"view{i}.jpg", index=i) for i in range(len(rgb))]
    rec.rms_px = 0.55
    return rec, rgb, masks, gains, biases, frames


def main(ref_root: str) -> None:
    rec, rgb, masks, gains, biases, frames = inputs()
    out = dict(R=rec.R, C=rec.C, X=rec.X, obs_cam=rec.obs_cam, obs_pt=rec.obs_pt, rms_px=np.float64(rec.rms_px),
               intr=np.array([rec.intr[0].f, rec.intr[0].cx, rec.intr[0].cy, W, H], np.float64),
               rgb=np.stack(rgb), mask2=masks["view2.jpg"],
               gain_1=gains[1], gain_4=gains[4], bias_4=biases[4], bias_6=biases[6])
    with tempfile.TemporaryDirectory() as td:
        build_reference_package(ref_root, td)
        sys.path.insert(0, td)
        from orthomosaic import densify as D

        def load_rgb(fr, scale):
            assert scale == 1.0
            return rgb[fr.index]

        D.load_rgb = load_rgb
        D.load_mask = lambda path, w, h: masks.get(path)
        ar = SimpleNamespace(frames=frames, workers=2)
        for name, kw in SETTINGS.items():
            opt = D.DepthOptions(keep_depthmaps=True, workers=2, **kw)
            od = os.path.join(td, "out_" + name)
            cloud = D.densify(ar, rec, gains, biases, od, opt)
            out[f"{name}_xyz"] = cloud.xyz
            out[f"{name}_rgb"] = cloud.rgb
            out[f"{name}_views"] = cloud.views
            out[f"{name}_cam"] = cloud.cam
            out[f"{name}_spacing"] = np.float64(cloud.spacing)
            wd = os.path.join(od, "depthmaps")
            for path in sorted(glob.glob(os.path.join(wd, "*.npy"))):
                stem = os.path.basename(path)[:-4]
                if name == "pm" or stem.isdigit():         # pass-1 files of the PatchMatch run only
                    out[f"{name}_{stem}"] = np.load(path)
            print(name, "points", len(cloud.xyz), "files", len(glob.glob(os.path.join(wd, "*.npy"))))
    path = os.path.join(HERE, "densify_grid3x3.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
