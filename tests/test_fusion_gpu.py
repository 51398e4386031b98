"""Consistency and fusion on the device (om_consistency, om_fuse through py_ortho_amd.fusion):
bit for bit against the vectors the compiled reference wrote and against the oracle on seeded
surveys with mixed image sizes, lens distortion, NaN holes and edge settings."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

import fusion_cases
import oracle

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


def _gpu_fuse(order, depths, cams, rgb, nbrs, have, rel_tol, px_tol, min_views, stride, cap=None):
    from py_ortho_amd import fusion
    dm = {k: torch.from_numpy(np.ascontiguousarray(depths[k])).cuda() for k in have}
    rg = None if rgb is None else {k: torch.from_numpy(np.ascontiguousarray(rgb[k])).cuda() for k in have}
    pts, cols, nv, owner = [], [], [], []
    used = fusion.fuse(order, dm, None, cams, nbrs, set(have), lambda k: rel_tol[k],
                       SimpleNamespace(point_stride=stride), px_tol, min_views, rg, pts, cols, nv, owner, cap=cap)
    torch.cuda.synchronize()
    out = [torch.cat(x).cpu().numpy() if x else None for x in (pts, cols, nv, owner)]
    full = {k: fusion.used_full(used[k], cams[k].h, cams[k].w, stride) for k in have}
    return out, full


def _oracle_fuse(order, depths, cams, rgb, nbrs, have, rel_tol, px_tol, min_views, stride):
    n = len(cams)
    nb, tol = fusion_cases.nbr_table(nbrs, set(have), rel_tol, n)
    d = [depths[k] if k in have else None for k in range(n)]
    r = [rgb[k] if k in have else None for k in range(n)]
    pts, cols, nv, owner, used = oracle.fuse(order, d, cams, r, nb, tol, px_tol, min_views, stride)
    return [pts, cols, nv, owner], {k: used[k].astype(bool) for k in have}


@pytest.mark.parametrize("name", ["a", "b"])
def test_consistency_matches_reference(name):
    from py_ortho_amd import fusion
    print(f"consistency golden {name}", flush=True)
    g = fusion_cases.golden()
    s = g.settings[name]
    dm = {k: torch.from_numpy(g.depths[k]).cuda() for k in g.have}
    for k in g.have:
        c = fusion.consistency(k, dm[k], g.nbrs[k], dm, g.cams, s.px_tol, s.rel_tol[k])
        assert _same(c.cpu().numpy(), s.cnt[k]), k


@pytest.mark.parametrize("name", ["a", "b"])
def test_fusion_matches_reference(name):
    print(f"fusion golden {name}", flush=True)
    g = fusion_cases.golden()
    s = g.settings[name]
    (pts, cols, nv, owner), used = _gpu_fuse(g.have, g.depths, g.cams, g.rgb, g.nbrs, g.have, s.rel_tol, s.px_tol,
                                             s.min_views, s.stride)
    assert _same(pts, s.pts) and _same(cols, s.cols) and _same(nv, s.nv) and _same(owner, s.owner)
    for k in g.have:
        assert _same(used[k], s.used[k]), k


SURVEYS = [  # seed, nx, ny, sizes, neighbour counts, k1, k2
    (5, 3, 3, [(320, 240)], (3, 5), 0.0, 0.0),
    (6, 4, 3, [(320, 240), (256, 192), (400, 300)], (2, 6), -0.03, 0.004),
]


@pytest.fixture(scope="module", params=range(len(SURVEYS)))
def survey(request):
    seed, nx, ny, sizes, counts, k1, k2 = SURVEYS[request.param]
    cams, depths, rgb, nbrs = fusion_cases.survey_case(seed, nx, ny, sizes, counts, k1, k2, device="cuda")
    return SimpleNamespace(seed=seed, cams=cams, depths=depths, rgb=rgb, nbrs=nbrs)


def test_consistency_matches_oracle(survey):
    from py_ortho_amd import fusion
    print(f"consistency survey seed {survey.seed}", flush=True)
    n = len(survey.cams)
    tol = fusion_cases.rel_tols_by_baseline(survey.cams, survey.nbrs, 1.5, 60.0)
    dm = {k: torch.from_numpy(survey.depths[k]).cuda() for k in range(n)}
    for k in range(n):
        nb = survey.nbrs[k]
        c = fusion.consistency(k, dm[k], nb, dm, survey.cams, 1.5, tol[k]).cpu().numpy()
        ref = oracle.consistency(survey.depths[k], survey.cams[k], [survey.depths[j] for j in nb],
                                 [survey.cams[j] for j in nb], [tol[k][j] for j in nb], 1.5)
        assert _same(c, ref), k
        assert (c > 0).mean() > 0.3


@pytest.mark.parametrize("px_tol,min_views,stride", [(1.5, 3, 2), (1.0, 2, 1), (2.5, 1, 3), (1.5, 4, 5)])
def test_fusion_matches_oracle(survey, px_tol, min_views, stride):
    print(f"fusion survey seed {survey.seed} px {px_tol} mv {min_views} st {stride}", flush=True)
    n = len(survey.cams)
    tol = fusion_cases.rel_tols_by_baseline(survey.cams, survey.nbrs, px_tol, 60.0)
    have = list(range(n))
    order = have[::-1] if stride == 3 else have                  # any order
    a, ua = _gpu_fuse(order, survey.depths, survey.cams, survey.rgb, survey.nbrs, have, tol, px_tol, min_views,
                      stride)
    b, ub = _oracle_fuse(order, survey.depths, survey.cams, survey.rgb, survey.nbrs, have, tol, px_tol, min_views,
                         stride)
    assert len(b[0]) > 100
    for x, y in zip(a, b):
        assert _same(x, y)
    for k in have:
        assert _same(ua[k], ub[k]), k


def test_fusion_missing_views_and_empty_maps(survey):
    """Views without a depth map are skipped as neighbours (`j in have`); an all-NaN depth map emits
    nothing; colours are optional."""
    print(f"fusion missing views seed {survey.seed}", flush=True)
    n = len(survey.cams)
    tol = fusion_cases.rel_tols_by_baseline(survey.cams, survey.nbrs, 1.5, 60.0)
    have = [k for k in range(n) if k % 3 != 1]
    depths = list(survey.depths)
    depths[have[1]] = np.full_like(depths[have[1]], np.nan)
    a, ua = _gpu_fuse(have, depths, survey.cams, survey.rgb, survey.nbrs, have, tol, 1.5, 2, 2)
    b, ub = _oracle_fuse(have, depths, survey.cams, survey.rgb, survey.nbrs, have, tol, 1.5, 2, 2)
    for x, y in zip(a, b):
        assert _same(x, y)
    assert have[1] not in set(a[3].tolist())
    c, _ = _gpu_fuse(have, depths, survey.cams, None, survey.nbrs, have, tol, 1.5, 2, 2)
    assert c[1] is None and _same(c[0], a[0]) and _same(c[3], a[3])


def test_fusion_capacity_and_errors(survey):
    from py_ortho_amd import _lib, fusion
    print(f"fusion errors seed {survey.seed}", flush=True)
    n = len(survey.cams)
    tol = fusion_cases.rel_tols_by_baseline(survey.cams, survey.nbrs, 1.5, 60.0)
    have = list(range(n))
    with pytest.raises(RuntimeError, match="capacity"):
        _gpu_fuse(have, survey.depths, survey.cams, survey.rgb, survey.nbrs, have, tol, 1.5, 2, 2, cap=10)
    bad = [list(x) for x in survey.nbrs]
    bad[0] = [0] + bad[0]
    with pytest.raises(ValueError, match="neighbour"):
        _gpu_fuse(have, survey.depths, survey.cams, survey.rgb, bad, have, tol | {0: tol[0] | {0: 0.01}}, 1.5, 2, 2)
    dm = {k: torch.from_numpy(survey.depths[k]).cuda() for k in have}
    with pytest.raises(ValueError):
        fusion.consistency(0, dm[0][:, :-1].contiguous(), survey.nbrs[0], dm, survey.cams, 1.5, tol[0])
    assert _lib.OM_EINVAL < 0


def test_fusion_invariants_at_full_size():
    """A 1600x1200 3x2 block: equal to the oracle bit for bit at stride 2 (block offsets summed in
    the emit kernel) and stride 1 (7500 blocks a view: the scan kernel); every point is averaged
    from >= min_views agreeing images, owners follow the fusion order, points lie in the terrain's
    height range, and a second run from fresh used maps repeats the first bit for bit."""
    from py_ortho_amd import fusion, scene
    print("fusion full size 1600x1200", flush=True)
    cams, counts, ext = scene.survey_cameras(9, 3, 2, 90.0, [(1600, 1200)], 0.9375, 0.7, 0.6, (4, 4))
    sv = scene.make_survey(9, cams, counts, ext, 90.0, device="cuda", march=48, bisect=30, n_bumps=12)
    depths = fusion_cases.noisy_depths(sv, 10)
    n = len(cams)
    tol = fusion_cases.rel_tols_by_baseline(cams, sv.nbrs, 1.5, 80.0)
    have = list(range(n))
    for stride in (2, 1):
        print(f"  stride {stride}", flush=True)
        a, ua = _gpu_fuse(have, depths, cams, None, sv.nbrs, have, tol, 1.5, 3, stride)
        b, ub = _oracle_fuse(have, depths, cams, [np.zeros((c.h, c.w, 3), np.uint8) for c in cams], sv.nbrs, have,
                             tol, 1.5, 3, stride)
        assert _same(a[0], b[0]) and _same(a[2], b[2]) and _same(a[3], b[3])
        assert all(_same(ua[k], ub[k]) for k in have)
        P, N, O = a[0], a[2], a[3]
        assert len(P) > 100000 and np.isfinite(P).all()
        assert (N >= 3).all() and (np.diff(O) >= 0).all()
        assert P[:, 2].min() > -3.0 and P[:, 2].max() < 18.0
    dm = {k: torch.from_numpy(depths[k]).cuda() for k in have}
    f = fusion.Fuser(have, dm, cams, sv.nbrs, set(have), lambda k: tol[k], 1.5, 3, 2)
    runs = []
    for _ in range(2):
        f.run(reset_used=True)
        runs.append([x.cpu().numpy().copy() for x in f.result() if x is not None])
    assert all(_same(x, y) for x, y in zip(runs[0], runs[1]))
    assert f.evaluated() > len(runs[0][0])
