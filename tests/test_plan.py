"""Host planning of a depth map (py_ortho_amd.plan) against the reference's own outputs
(golden vectors written by the compiled reference, tests/golden/make_depthmap_golden.py)."""
import numpy as np

from depthmap_golden import load


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


def test_neighbours_match_reference():
    from py_ortho_amd import plan
    z, rec, opt, nbrs, _ = load()
    assert plan.neighbours(rec, opt["neighbors"]) == nbrs


def test_depth_range_band_and_count_match_reference():
    from py_ortho_amd import mvs, plan
    z, rec, opt, nbrs, (W, H, f, cx, cy) = load()
    for k in z["views"]:
        rng = plan.depth_range(rec, int(k))
        assert np.array_equal(np.array(rng), z[f"rng_{k}"])
        cam = mvs.Cam(rec.R[k], rec.C[k], f, 0.0, 0.0, cx, cy, W, H).scaled(0.5, W // 2, H // 2).scaled(
            0.5, W // 4, H // 4)
        nd = plan.num_depths(rec, int(k), nbrs[k], rng, cam.f, opt["num_depths"])
        assert nd == int(z[f"nd_{k}"])
        shape = (H // 4, W // 4)
        lo, hi = plan.depth_band(rec, int(k), cam, shape, rng, block=max(8, shape[1] // 16))
        assert _same(lo, z[f"inv_lo_{k}"]) and _same(hi, z[f"inv_hi_{k}"])
