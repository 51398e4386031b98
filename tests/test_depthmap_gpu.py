"""GPU depth_map (py_ortho_amd.mvs.depth_map: pyramid, coarse sweep over the tie-point band,
two refinement levels, range/mask filter) against the reference's own densify.depth_map output
(golden vectors written by the compiled reference), plus its ray map and PatchMatch normals."""
import numpy as np
import pytest
import torch

from depthmap_golden import load

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


def _report(a, b):
    fa, fb = np.isfinite(a), np.isfinite(b)
    both = fa & fb
    return (f"finite-mask equal {np.array_equal(fa, fb)}, bits equal {np.mean(a.view(np.uint32) == b.view(np.uint32)):.6f}, "
            f"max |d| {np.max(np.abs(a[both] - b[both])) if both.any() else 0:.3g}")


def test_depth_map_matches_reference():
    from py_ortho_amd import mvs
    z, rec, opt_kw, nbrs, (W, H, f, cx, cy) = load()
    opt = mvs.DepthOptions(**opt_kw)
    cams = [mvs.Cam(rec.R[j], rec.C[j], f, 0.0, 0.0, cx, cy, W, H) for j in range(len(rec.used))]
    grey = z["grey"]

    def views(j):
        return grey[j], None, cams[j]

    for k in z["views"]:
        k = int(k)
        depth, score, low = mvs.depth_map(k, nbrs[k], rec, views, opt)
        d, s, lo = depth.cpu().numpy(), score.cpu().numpy(), low.cpu().numpy()
        assert _same(d, z[f"depth_{k}"]), (k, _report(d, z[f"depth_{k}"]))
        assert _same(s, z[f"score_{k}"]), (k, _report(s, z[f"score_{k}"]))
        assert np.array_equal(lo, z[f"low_{k}"])


def test_rays_world_matches_reference():
    from py_ortho_amd import mvs
    z, rec, opt_kw, nbrs, (W, H, f, cx, cy) = load()
    for k in z["views"]:
        k = int(k)
        cam = mvs.Cam(rec.R[k], rec.C[k], f, 0.0, 0.0, cx, cy, W, H)
        ray = mvs.rays_world((H, W), cam).cpu().numpy()
        assert _same(ray, z[f"ray_{k}"])


def test_init_normals_match_reference():
    """Device normals == densify._init_normals's recorded output, bit for bit (NumPy's float32
    cross product, norm (x0^2 + x1^2) + x2^2 and division, all correctly rounded on the device)."""
    from py_ortho_amd import mvs
    z, rec, opt_kw, nbrs, (W, H, f, cx, cy) = load()
    for k in z["views"]:
        k = int(k)
        n = mvs.init_normals(torch.from_numpy(z[f"depth_{k}"]).cuda(), torch.from_numpy(z[f"ray_{k}"]).cuda())
        exp = z[f"normals_{k}"]
        got = n.cpu().numpy()
        assert _same(got, exp), (k, float(np.nanmax(np.abs(got - exp))), float(np.mean(got.view(np.uint32) == exp.view(np.uint32))))
