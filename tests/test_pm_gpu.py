"""GPU PatchMatch (libom355.so om_pm_refine) against the reference's golden vectors and the oracle.

Bar: depth, normal and score bit-identical and the same number of accepted changes -- the GPU
replays the reference's red-black order and its xorshift32 stream exactly (jump-ahead per pixel),
and evaluates costs with the reference's FP64 operations.
"""
import numpy as np
import pytest
import torch

import oracle
from golden.make_pm_golden import make_case
from test_oracle import GOLDEN_CASES, load_golden

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _gpu_pm(ref, ray, Cr, srcs, deps, par, depth, normal, score, active, p, iters):
    from py_ortho_amd import dense
    d, n, s = _dev(depth), _dev(normal), _dev(score)
    ch = dense.pm_refine(_dev(ref), _dev(ray), Cr, [_dev(x) for x in srcs], [None if x is None else _dev(x) for x in deps],
                         par, d, n, s, _dev(active), p["r"], p["step"], p["top_k"], p["geo_w"], p["geo_tol"], iters,
                         p["dmin"], p["dmax"], p["seed"])
    return ch, d.cpu().numpy(), n.cpu().numpy(), s.cpu().numpy()


def _report(a, b):
    fin = np.isfinite(a) & np.isfinite(b)
    return (f"finite-mask equal {np.array_equal(np.isfinite(a), np.isfinite(b))}, "
            f"bits equal {np.mean(a.view(np.uint32) == b.view(np.uint32)):.6f}, "
            f"max rel {np.max(np.abs(a[fin] - b[fin]) / np.abs(b[fin])) if fin.any() else 0:.3g}")


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_pm_refine_matches_reference_golden(name):
    z, srcs, deps, p = load_golden(name)
    for iters in z["iters"]:
        ch, d, n, s = _gpu_pm(z["ref"], z["ray"], z["Cr"], srcs, deps, z["par"], z["depth0"], z["normal0"],
                              z["score0"], z["active"], p, int(iters))
        exp_d = z[f"depth_it{iters}"]
        assert _same(d, exp_d), (name, iters, _report(d, exp_d))
        assert _same(n, z[f"normal_it{iters}"]), (name, iters)
        assert _same(s, z[f"score_it{iters}"]), (name, iters)
        assert ch == int(z[f"changed_it{iters}"]), (name, iters, ch)


@pytest.mark.parametrize("S,wd,r,step,top_k,geo_w", [(4, True, 5, 2, 2, 0.5), (12, True, 3, 2, 3, 0.5),
                                                     (16, False, 2, 1, 4, 0.0), (1, True, 4, 2, 2, 0.5)])
def test_pm_refine_matches_oracle(S, wd, r, step, top_k, geo_w):
    c = make_case(40 + S, 26, 34, S, wd)
    p = dict(r=r, step=step, top_k=top_k, geo_w=geo_w, geo_tol=0.02, dmin=c["dmin"], dmax=c["dmax"],
             seed=12345 + S)
    d0, n0, s0 = c["depth"].copy(), c["normal"].copy(), c["score"].copy()
    ch0 = oracle.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], d0, n0, s0,
                           c["active"], r, step, top_k, geo_w, 0.02, 2, c["dmin"], c["dmax"], p["seed"])
    ch, d, n, s = _gpu_pm(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], c["depth"],
                          c["normal"], c["score"], c["active"], p, 2)
    assert _same(d, d0), _report(d, d0)
    assert _same(n, n0) and _same(s, s0) and ch == ch0


@pytest.mark.parametrize("S,wd,r,step,top_k,geo_w", [(4, True, 5, 2, 2, 0.5), (12, False, 3, 1, 3, 0.0)])
def test_pm_refine_direct_path_matches_oracle(monkeypatch, S, wd, r, step, top_k, geo_w):
    """The direct per-pixel propagation kernel (images of 2^27 pixels and more, forced here by
    lowering the memo limit) gives the oracle's bits too."""
    monkeypatch.setenv("OM_PM_MEMO_MAX_PIXELS", "0")
    c = make_case(60 + S, 23, 31, S, wd)
    p = dict(r=r, step=step, top_k=top_k, geo_w=geo_w, geo_tol=0.02, dmin=c["dmin"], dmax=c["dmax"],
             seed=777 + S)
    d0, n0, s0 = c["depth"].copy(), c["normal"].copy(), c["score"].copy()
    ch0 = oracle.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], d0, n0, s0,
                           c["active"], r, step, top_k, geo_w, 0.02, 2, c["dmin"], c["dmax"], p["seed"])
    ch, d, n, s = _gpu_pm(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], c["depth"],
                          c["normal"], c["score"], c["active"], p, 2)
    assert _same(d, d0), _report(d, d0)
    assert _same(n, n0) and _same(s, s0) and ch == ch0


@pytest.mark.parametrize("direct", [False, True])
@pytest.mark.parametrize("H,W", [(1, 1), (1, 6), (4, 1), (2, 3), (3, 7), (6, 11)])
def test_pm_refine_tiny_views_match_oracle(monkeypatch, H, W, direct):
    """Reference images of a few pixels (windows clipped on every side, colour tiles mostly empty),
    on the memoised and the direct path."""
    if direct:
        monkeypatch.setenv("OM_PM_MEMO_MAX_PIXELS", "0")
    c = make_case(90 + H * 10 + W, 9, 11, 4, True)
    # a fresh array per call: a cut of whole rows would otherwise be a view of the case, and the
    # oracle's in-place refinement would become the GPU's input
    cut = lambda a: np.array(a[..., :H, :W], order="C", copy=True)  # noqa: E731
    ref, ray, act = cut(c["ref"]), cut(c["ray"]), cut(c["active"])
    act[:] = 1
    p = dict(r=2, step=1, top_k=2, geo_w=0.5, geo_tol=0.02, dmin=c["dmin"], dmax=c["dmax"], seed=5)
    d0, n0, s0 = cut(c["depth"]), cut(c["normal"]), cut(c["score"])
    ch0 = oracle.pm_refine(ref, ray, c["Cr"], c["srcs"], c["src_depths"], c["par"], d0, n0, s0, act, p["r"],
                           p["step"], p["top_k"], p["geo_w"], p["geo_tol"], 2, p["dmin"], p["dmax"], p["seed"])
    ch, d, n, s = _gpu_pm(ref, ray, c["Cr"], c["srcs"], c["src_depths"], c["par"], cut(c["depth"]),
                          cut(c["normal"]), cut(c["score"]), act, p, 2)
    assert _same(d, d0), _report(d, d0)
    assert _same(n, n0) and _same(s, s0) and ch == ch0


def test_pm_refine_no_sources_returns_zero():
    from py_ortho_amd import dense
    z, srcs, deps, p = load_golden("s3")
    d = _dev(z["depth0"])
    assert dense.pm_refine(_dev(z["ref"]), _dev(z["ray"]), z["Cr"], [], [], np.zeros((0, 18)), d, _dev(z["normal0"]),
                           _dev(z["score0"]), _dev(z["active"]), 4, 2, 2, 0.5, 0.02, 2, 1.0, 100.0, 1) == 0
    assert _same(d.cpu().numpy(), z["depth0"])


def test_pm_refine_drone_scale_memo_equals_direct_and_repeats():
    """At the drone-scale workload's full size (2000x1500 reference, 8 sources, 11x11 window with
    step 2, geometric term, 2 of its 8 iterations): the memoised propagation and the direct one give
    bit-identical depth, normal and score with the same change count, and a second run repeats the
    first bit for bit.  The oracle would need minutes here; these properties do not depend on size."""
    import os
    from py_ortho_amd import dense, scene
    sc = scene.config_c2(device="cuda")
    H, W = sc.ref.img.shape
    ray = scene.rays_world(sc.ref.cam, "cuda")
    depth0, normal0 = scene.pm_init(3, H, W, ray, 80.0, 105.0)
    depth0, normal0 = depth0.cuda(), normal0.cuda()
    ref, srcs = sc.ref.img.contiguous(), [v.img.contiguous() for v in sc.srcs]
    deps, par, Cr = [v.depth.contiguous() for v in sc.srcs], sc.par(), sc.ref.cam.C.copy()
    active = torch.ones((H, W), dtype=torch.uint8, device="cuda")

    def run():
        d, n = depth0.clone(), normal0.clone()
        s = torch.full((H, W), -1.0, dtype=torch.float32, device="cuda")
        ch = dense.pm_refine(ref, ray, Cr, srcs, deps, par, d, n, s, active, 5, 2, 2, 0.5, 0.01, 2, 80.0, 105.0,
                             2654435761)
        return ch, d.cpu().numpy(), n.cpu().numpy(), s.cpu().numpy()

    assert "OM_PM_MEMO_MAX_PIXELS" not in os.environ
    a, b = run(), run()
    os.environ["OM_PM_MEMO_MAX_PIXELS"] = "0"
    try:
        c = run()
    finally:
        del os.environ["OM_PM_MEMO_MAX_PIXELS"]
    assert a[0] > H * W // 2  # most pixels change in the first iterations
    for other in (b, c):
        assert other[0] == a[0]
        for k in (1, 2, 3):
            assert _same(other[k], a[k]), (k, _report(other[k], a[k]))
