"""Host planning of a depth map (py_ortho_amd.plan) against the reference's own outputs
(golden vectors written by the compiled reference, tests/golden/make_depthmap_golden.py)."""
import numpy as np
import pytest

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


def _coarse_cam(rec, k, W, H, f, cx, cy):
    from py_ortho_amd import mvs
    return mvs.Cam(rec.R[k], rec.C[k], f, 0.0, 0.0, cx, cy, W, H).scaled(0.5, W // 2, H // 2).scaled(
        0.5, W // 4, H // 4)


def test_depth_range_and_count_match_reference():
    from py_ortho_amd import plan
    z, rec, opt, nbrs, (W, H, f, cx, cy) = load()
    for k in z["views"]:
        rng = plan.depth_range(rec, int(k))
        assert np.array_equal(np.array(rng), z[f"rng_{k}"])
        cam = _coarse_cam(rec, k, W, H, f, cx, cy)
        nd = plan.num_depths(rec, int(k), nbrs[k], rng, cam.f, opt["num_depths"])
        assert nd == int(z[f"nd_{k}"])


def test_shared_counts_equal_per_point_count():
    """The sparse incidence product counts, for every pair of images, the tie points both observe:
    the same as walking the points (a point observed twice by one image counts once)."""
    from types import SimpleNamespace

    from py_ortho_amd import plan
    rng = np.random.default_rng(3)
    N, P = 7, 60
    cam = rng.integers(0, N, 400)
    pt = rng.integers(0, P, 400)
    rec = SimpleNamespace(used=list(range(N)), obs_cam=cam, obs_pt=pt, X=np.zeros((P, 3)))
    want = np.zeros((N, N), np.int32)
    for p in range(P):
        g = sorted(set(cam[pt == p]))
        for a in g:
            for b in g:
                if a != b:
                    want[a, b] += 1
    assert np.array_equal(plan.TiePoints(rec).shared(), want)


@pytest.mark.gpu
def test_depth_band_on_device_matches_reference():
    """om_depth_band == densify._depth_band's recorded maps bit for bit, and the linear step
    (inv_hi - inv_lo) / max(nd - 1, 1) in float32 (densify.py:411)."""
    from py_ortho_amd import plan
    z, rec, opt, nbrs, (W, H, f, cx, cy) = load()
    for k in z["views"]:
        rng = plan.depth_range(rec, int(k))
        cam = _coarse_cam(rec, k, W, H, f, cx, cy)
        nd = int(z[f"nd_{k}"])
        shape = (H // 4, W // 4)
        lo, hi, step = plan.depth_band(rec, int(k), cam, shape, rng, block=max(8, shape[1] // 16), nd=nd)
        lo0, hi0 = z[f"inv_lo_{k}"], z[f"inv_hi_{k}"]
        assert _same(lo.cpu().numpy(), lo0) and _same(hi.cpu().numpy(), hi0)
        assert _same(step.cpu().numpy(), (hi0 - lo0) / max(nd - 1, 1))
        # fewer than 50 own observations: the image-wide range everywhere
        few = type(rec)(**{**vars(rec), "obs_cam": rec.obs_cam[:0], "obs_pt": rec.obs_pt[:0]})
        lo, hi = plan.depth_band(few, int(k), cam, shape, rng)
        assert _same(lo.cpu().numpy(), np.full(shape, 1.0 / rng[1], np.float32))
        assert _same(hi.cpu().numpy(), np.full(shape, 1.0 / rng[0], np.float32))
