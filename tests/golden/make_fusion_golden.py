"""Writes tests/golden/fusion_grid3x3.npz: vectors of the reference's geometric consistency check
(densify._consistency, densify.py:441-478) and sequential fusion (densify._fuse_loop, :688-762).

Neither has a reference test or a closed form (the fusion order and the used-pixel marking decide
which points exist), so the vectors come from the reference itself: the compiled reference
package (as in make_depthmap_golden.py) computes the depth maps of a seeded 3 x 3 survey of
cameras turned by up to 4 degrees off nadir (so that no rotation entry is 0 or 1) with
densify.depth_map, filters them as the repair pass does (densify.py:641), then records
_consistency for every view and _fuse_loop over all views for two settings.  Only arrays are
committed; the colour images are regenerated from their seed by the test.  Needs the reference
tree and Cython; run from the repository root:

    python tests/golden/make_fusion_golden.py [/path/to/reference]
"""
from __future__ import annotations

import os
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_depthmap_golden import build_reference_package, synthetic_survey  # noqa: E402

# name, px_tol, min_views, point_stride
SETTINGS = (("a", 1.5, 3, 2), ("b", 1.0, 2, 1))


def colours(seed: int, n: int, H: int, W: int) -> np.ndarray:
    """The colour images the fusion reads (rgb[vv, uu]): seeded noise, regenerated by the test."""
    return np.random.default_rng(seed).integers(0, 256, (n, H, W, 3), dtype=np.uint8)


class _NoCache:
    def clear(self):
        pass


class _Done:
    def __init__(self, v):
        self.v = v

    def result(self):
        return self.v


class _SyncPool:
    """Stands in for the background view decoder of densify._fuse: the same views, synchronously."""

    def submit(self, fn, *a):
        return _Done(fn(*a))


def main(ref_root: str) -> None:
    seed, W, H, f, alt, (gx, gy), spacing, col_seed = 21, 160, 120, 150.0, 60.0, (3, 3), 14.0, 77
    with tempfile.TemporaryDirectory() as td:
        build_reference_package(ref_root, td)
        sys.path.insert(0, td)
        from orthomosaic import densify as D
        rec, grey, imgs, deps, cams = synthetic_survey(seed, W, H, f, alt, gx, gy, spacing, rot_deg=4.0)
        N = len(cams)
        opt = D.DepthOptions(num_depths=32, window=3, neighbors=4)
        dcams = [D._Cam(c.R, c.C, c.f, 0.0, 0.0, c.cx, c.cy, W, H, 0.0) for c in cams]
        rgb = colours(col_seed, N, H, W)

        def view(j):
            return grey[j], rgb[j], dcams[j]

        nbrs = D._neighbours(rec, opt.neighbors)
        dm = {}
        for k in range(N):
            res = D.depth_map(k, nbrs[k], rec, view, opt)
            if res is not None:
                depth, score, low = res
                depth = np.array(depth, np.float32)
                depth[(score < opt.min_ncc) | low] = np.nan          # densify.py:641
                dm[k] = depth
        have = sorted(dm)
        dmed = {k: float(np.median((rec.R[k] @ (rec.X[rec.obs_pt[rec.obs_cam == k]] - rec.C[k]).T)[2]))
                for k in range(N) if np.any(rec.obs_cam == k)}
        out = dict(R=rec.R, C=rec.C, intr=np.array([f, (W - 1) / 2, (H - 1) / 2, W, H], np.float64),
                   nbrs=np.array([nb + [-1] * (opt.neighbors - len(nb)) for nb in nbrs], np.int32),
                   have=np.array(have, np.int32), col_seed=np.int64(col_seed))
        for k in have:
            out[f"depth_{k}"] = dm[k]
        for name, px_tol, min_views, stride in SETTINGS:
            def rel_tols(k, px_tol=px_tol):                         # densify.py:574-582
                o = {}
                for j in nbrs[k]:
                    B = float(np.linalg.norm(rec.C[k] - rec.C[j]))
                    o[j] = float(np.clip(px_tol * dmed.get(k, 1.0) / max(dcams[k].f * B, 1e-9), 0.01, 0.05))
                return o
            out[f"{name}_params"] = np.array([px_tol, min_views, stride], np.float64)
            for k in have:
                tol = rel_tols(k)
                out[f"{name}_reltol_{k}"] = np.array([tol[j] for j in nbrs[k]], np.float64)
                out[f"{name}_cnt_{k}"] = D._consistency(k, dm[k], nbrs[k], dm, dcams, px_tol, tol)
            fopt = D.DepthOptions(num_depths=32, window=3, neighbors=4, point_stride=stride)
            used = {k: np.zeros(dm[k].shape, bool) for k in have}
            pts, cols, nv, owner = [], [], [], []
            pool = ThreadPoolExecutor(2)
            D._fuse_loop(have, dm, used, dcams, nbrs, set(have), rel_tols, fopt, px_tol, min_views, view,
                         _NoCache(), pts, cols, nv, owner, 10 ** 9, pool, _SyncPool())
            pool.shutdown()
            out.update({f"{name}_pts": np.concatenate(pts), f"{name}_cols": np.concatenate(cols),
                        f"{name}_nv": np.concatenate(nv), f"{name}_owner": np.concatenate(owner)})
            for k in have:
                out[f"{name}_used_{k}"] = used[k]
            print(name, "points", len(out[f"{name}_pts"]), "median views/point", np.median(out[f"{name}_nv"]))
        path = os.path.join(HERE, "fusion_grid3x3.npz")
        np.savez_compressed(path, **out)
        print(path, "views", len(have))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
