"""Inputs of the consistency / fusion tests: the golden vectors written by the compiled reference
(tests/golden/fusion_grid3x3.npz) and seeded synthetic surveys with noisy, holed depth maps."""
import os
from types import SimpleNamespace

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden():
    z = np.load(os.path.join(GOLDEN, "fusion_grid3x3.npz"))
    f, cx, cy, W, H = z["intr"]
    W, H = int(W), int(H)
    N = z["R"].shape[0]
    cams = [SimpleNamespace(R=z["R"][k], C=z["C"][k], f=float(f), k1=0.0, k2=0.0, k3=0.0, cx=float(cx),
                            cy=float(cy), w=W, h=H) for k in range(N)]
    have = [int(k) for k in z["have"]]
    depths = {k: z[f"depth_{k}"] for k in have}
    nbrs = [[int(j) for j in row if j >= 0] for row in z["nbrs"]]
    rgb = np.random.default_rng(int(z["col_seed"])).integers(0, 256, (N, H, W, 3), dtype=np.uint8)
    settings = {}
    for name in ("a", "b"):
        px, mv, st = z[f"{name}_params"]
        tols = {k: {j: float(t) for j, t in zip(nbrs[k], z[f"{name}_reltol_{k}"])} for k in have}
        settings[name] = SimpleNamespace(
            px_tol=float(px), min_views=int(mv), stride=int(st), rel_tol=tols,
            cnt={k: z[f"{name}_cnt_{k}"] for k in have}, used={k: z[f"{name}_used_{k}"] for k in have},
            pts=z[f"{name}_pts"], cols=z[f"{name}_cols"], nv=z[f"{name}_nv"], owner=z[f"{name}_owner"])
    return SimpleNamespace(cams=cams, have=have, depths=depths, nbrs=nbrs, rgb=rgb, settings=settings)


def rel_tols_by_baseline(cams, nbrs, px_tol, depth_med):
    """densify.rel_tols (densify.py:574-582): px_tol * median depth / (f * baseline), clipped to
    [0.01, 0.05]."""
    out = {}
    for k, nb in enumerate(nbrs):
        out[k] = {j: float(np.clip(px_tol * depth_med / max(cams[k].f * float(np.linalg.norm(cams[k].C - cams[j].C)),
                                                           1e-9), 0.01, 0.05)) for j in nb}
    return out


def noisy_depths(survey, seed, rel_noise=2e-3, hole_frac=0.03, n_patches=4):
    from py_ortho_amd import scene
    return scene.noisy_depths(survey, seed, rel_noise, hole_frac, n_patches)


def survey_case(seed, nx, ny, sizes, counts=(3, 5), k1=0.0, k2=0.0, alt=60.0, device="cpu"):
    """A small nadir block for fusion: cameras (scene.Camera, radial k1 / k2 set after rendering),
    noisy depth maps, colours, neighbour lists and per-pair tolerances."""
    from py_ortho_amd import scene
    cams, cnt, ext = scene.survey_cameras(seed, nx, ny, alt, sizes, 0.9375, 0.7, 0.6, counts)
    sv = scene.make_survey(seed, cams, cnt, ext, alt, device=device, march=48, bisect=30, n_bumps=12)
    depths = noisy_depths(sv, seed + 1)
    for c in cams:
        c.k1, c.k2 = k1, k2
    rng = np.random.default_rng(seed + 2)
    rgb = [rng.integers(0, 256, (c.h, c.w, 3), dtype=np.uint8) for c in cams]
    return cams, depths, rgb, sv.nbrs


def nbr_table(nbrs, have, rel_tol, n):
    """(n, max_nb) neighbour table (-1 = none or no depth map) and tolerances, as om_fuse takes them."""
    m = max(1, max(len(x) for x in nbrs))
    nb = np.full((n, m), -1, np.int32)
    tol = np.zeros((n, m), np.float64)
    for k in range(n):
        for s, j in enumerate(nbrs[k]):
            if j in have:
                nb[k, s] = j
                tol[k, s] = rel_tol[k][j]
    return nb, tol
