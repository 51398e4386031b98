"""Writes tests/golden/sample_view.npz: vectors of the reference's height-sweep sampler
(_mvs.sample_view, _mvs.pyx:13-54) as CPUBackend.sample_view calls it (backend.py:67-77: a (H, W)
height grid or a (D, H, W) stack, one _mvs call per plane).

Cases (seeded):
  tilted    - a camera turned ~6 deg off nadir with k1, k2, k3 != 0 and a non-centred principal
              point, a 3-channel image, fractional sample positions, a (3, H, W) stack of heights
              that reaches past every image edge (invalid cells) and behind the camera;
  barrel    - strong barrel distortion (k1 < 0), 1 channel, one (H, W) grid;
  edges     - a nadir pinhole whose cells land exactly on pixel centres, including the last
              column and row (x1 = x0, y1 = y0 clamps) and one step past them (invalid);
  oblique   - a 35-deg oblique camera, 2 channels, heights spanning 60 m, cells near the horizon.
Only arrays are committed.  Needs the reference tree and Cython; run from the repository root:

    python tests/golden/make_sample_view_golden.py [/path/to/reference]
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_depthmap_golden import _rotation, build_reference_package  # noqa: E402


def cases(rng):
    out = []
    # tilted, distorted, 3 channels, (D, H, W) stack
    R = _rotation(rng, 6.0) @ np.diag([1.0, -1.0, -1.0])
    cam = (R, np.array([2.3, -1.7, 45.0]), 310.0, -0.08, 0.015, 0.004, 121.3, 98.6)
    img = rng.uniform(0, 255, (200, 240, 3)).astype(np.float32)
    Z = rng.uniform(-2.0, 9.0, (3, 57, 73)).astype(np.float32)
    Z[1, 3, 4] = 60.0                               # above the camera: behind it
    out.append(("tilted", img, cam, -31.3, 24.1, 0.8731, Z))
    # strong barrel distortion, one channel, one grid
    R = _rotation(rng, 3.0) @ np.diag([1.0, -1.0, -1.0])
    cam = (R, np.array([-0.4, 0.9, 30.0]), 180.0, -0.21, 0.04, -0.003, 79.25, 61.75)
    img = rng.uniform(-1, 1, (120, 160, 1)).astype(np.float32)
    Z = rng.uniform(0.0, 5.0, (41, 47)).astype(np.float32)
    out.append(("barrel", img, cam, -19.0, 15.0, 0.8125, Z))
    # nadir pinhole, cells on pixel centres incl. the last row/column and one step past them
    f, Hc, h, w = 64.0, 8.0, 30, 40                 # binary fractions: every step below is exact
    img = rng.uniform(0, 1, (h, w, 1)).astype(np.float32)
    gsd = Hc / f                                    # one cell per pixel at Z = 0
    cam = (np.diag([1.0, -1.0, -1.0]), np.array([0.0, 0.0, Hc]), f, 0.0, 0.0, 0.0, (w - 1) / 2, (h - 1) / 2)
    X0 = -((w - 1) / 2 + 0.5) * gsd                 # cell c lands on u = c
    Y0 = ((h - 1) / 2 + 0.5) * gsd                  # cell r lands on v = r
    Z = np.zeros((h + 1, w + 1), np.float32)
    out.append(("edges", img, cam, X0, Y0, gsd, Z))
    # oblique
    a = np.radians(35.0)
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    cam = (Rx @ np.diag([1.0, -1.0, -1.0]), np.array([0.0, -20.0, 50.0]), 250.0, 0.03, -0.01, 0.002, 127.5, 95.5)
    img = rng.uniform(0, 10, (192, 256, 2)).astype(np.float32)
    Z = rng.uniform(-10.0, 50.0, (2, 50, 64)).astype(np.float32)
    out.append(("oblique", img, cam, -55.0, -10.0, 1.7, Z))
    return out


def main(ref_root: str) -> None:
    with tempfile.TemporaryDirectory() as td:
        build_reference_package(ref_root, td)
        sys.path.insert(0, td)
        from orthomosaic.backend import CPUBackend
        be = CPUBackend()
        rng = np.random.default_rng(23)
        rec = {}
        for name, img, cam, X0, Y0, gsd, Z in cases(rng):
            R, C, f, k1, k2, k3, cx, cy = cam
            o, v = be.sample_view(img, (R, C, f, k1, k2, k3, cx, cy), X0, Y0, gsd, Z)
            rec[f"{name}_img"] = img
            rec[f"{name}_R"] = np.asarray(R, np.float64)
            rec[f"{name}_C"] = np.asarray(C, np.float64)
            rec[f"{name}_par"] = np.array([f, k1, k2, k3, cx, cy, X0, Y0, gsd], np.float64)
            rec[f"{name}_Z"] = Z
            rec[f"{name}_out"] = o
            rec[f"{name}_valid"] = v
            print(name, Z.shape, "valid", int(v.sum()), "of", v.size)
        rec["names"] = np.array([c[0] for c in cases(np.random.default_rng(23))])
        path = os.path.join(HERE, "sample_view.npz")
        np.savez_compressed(path, **rec)
        print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
