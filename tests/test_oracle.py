"""Pins of the CPU oracle (runs without a GPU).

Each check compares the oracle with something other than a retyped copy of itself: the
reference's own compiled code (oracle/_ref, tests/golden), the reference's unfused specification,
NumPy closed forms the reference's tests use, brute force, published values and invariants.
"""
import os

import numpy as np
import pytest

import cases
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


# ------------------------------------------------------------------------ plane sweep
@pytest.mark.parametrize("shape", cases.SWEEP_SHAPES + [(120, 150, 8, 5, 2, 6), (33, 70, 16, 7, 4, 4)])
def test_sweep_matches_reference_row_kernels(shape):
    """Restatement == the reference's _sweep.h compiled from the reference tree, bit for bit.

    The compiled reference is looked up inside the test, so that collecting this module (for
    example in a GPU-only run that deselects it) never loads it."""
    if not oracle.ref_built():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    H, W, n, r, k, nh = shape
    rng = np.random.default_rng(1000 + H + W)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
    a = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    b = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k, fn=oracle.ref_sweep_hypothesis)
    assert _same(a[1], b[1]) and all(_same(x, y) for x, y in zip(a[0], b[0]))


@pytest.mark.parametrize("shape", cases.SWEEP_SHAPES)
def test_fused_equals_unfused_spec(shape):
    """sweep_hypothesis == ncc_warp per source + combine (test_native_parity.py:97-112)."""
    H, W, n, r, k, nh = shape
    rng = np.random.default_rng(7 + H)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
    a_st, a_idx = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st, idx = oracle.new_state(H, W)
    for i, inv in enumerate(hyps):
        S = np.stack([oracle.ncc_warp(ref, mu, sd, s, ray, geo[j], inv, r)[0] for j, s in enumerate(srcs)])
        oracle.combine(S, k, i, *st, idx)
    assert _same(a_idx, idx) and all(_same(x, y) for x, y in zip(a_st, st))


def _brute_ncc(ref, mu, sd, src, ray, g, inv, r):
    """Direct FP64 evaluation of one source's windowed NCC, pixel by pixel (no running sums, no
    float32 rounding of partial sums): the textbook definition with the reference's partial-window
    rules (zero padding, >= 90 % valid, sd floor, clamp)."""
    H, W = ref.shape
    h, w = src.shape
    Rs, b, f, k1, k2, k3, cx, cy = g[:9].reshape(3, 3), g[9:12], *g[12:]
    d = 1.0 / inv.astype(np.float64)
    P = b[:, None, None] + d[None] * np.einsum("ij,jhw->ihw", Rs, ray.astype(np.float64))
    Z = P[2]
    with np.errstate(all="ignore"):
        nx, ny = P[0] / Z, P[1] / Z
        q = nx * nx + ny * ny
        dd = f * (1 + q * (k1 + q * (k2 + q * k3)))
        u, v = dd * nx + cx, dd * ny + cy
    ok = (Z > 1e-6) & (u >= 0) & (v >= 0) & (u <= w - 1.001) & (v <= h - 1.001)
    J = np.zeros((H, W))
    x0 = np.where(ok, u, 0).astype(int)
    y0 = np.where(ok, v, 0).astype(int)
    fx, fy = np.where(ok, u - x0, 0), np.where(ok, v - y0, 0)
    s64 = src.astype(np.float64)
    val = (s64[y0, x0] * (1 - fx) + s64[y0, x0 + 1] * fx) * (1 - fy) + \
          (s64[y0 + 1, x0] * (1 - fx) + s64[y0 + 1, x0 + 1] * fx) * fy
    ok &= np.isfinite(val)
    J = np.where(ok, val.astype(np.float32).astype(np.float64), 0.0)
    V = ok.astype(np.float64)
    out = np.empty((H, W))
    k = 2 * r + 1
    Ip = np.pad(ref.astype(np.float64), r)
    Jp, Vp = np.pad(J, r), np.pad(V, r)
    for y in range(H):
        for x in range(W):
            vw = Vp[y:y + k, x:x + k]
            jw = Jp[y:y + k, x:x + k]
            iw_ = Ip[y:y + k, x:x + k]
            cov = vw.sum() / (k * k)
            if cov <= 0.9:
                out[y, x] = -1
                continue
            mj = jw.sum() / (k * k) / cov
            m2 = (jw * jw).sum() / (k * k) / cov
            mij = (jw * iw_).sum() / (k * k) / cov
            sdj = m2 - mj * mj
            sdj = np.sqrt(sdj) if sdj > 1e-6 else 1e-3
            out[y, x] = np.clip((mij - mu[y, x] * mj) / (sd[y, x] * sdj), -1, 1)
    return out


def test_ncc_matches_brute_force_fp64():
    rng = np.random.default_rng(21)
    for H, W, n, r in ((14, 17, 2, 2), (9, 12, 3, 1), (16, 11, 2, 3)):
        ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, 3)
        for j in range(n):
            for inv in hyps:
                got = oracle.ncc_warp(ref, mu, sd, srcs[j], ray, geo[j], inv, r)[0]
                exp = _brute_ncc(ref, mu, sd, srcs[j], ray, geo[j], inv, r)
                # the reference rounds horizontal window sums to float32: ~1e-7 relative
                assert np.max(np.abs(got - exp)) < 2e-5, (H, W, r, float(np.max(np.abs(got - exp))))


def test_ncc_of_affine_copy_is_one():
    """Identity warp (same camera, b = 0) onto an affine copy of the reference: NCC = 1 on every
    fully valid window, -1 where fewer than 90 % of the window projects."""
    rng = np.random.default_rng(5)
    H, W, r = 20, 24, 2
    ref = rng.uniform(0, 1, (H, W)).astype(np.float32)
    mu, sd = cases.box_stats(ref, r)
    f, cx, cy = 30.0, (W - 1) / 2, (H - 1) / 2
    vv, uu = np.mgrid[0:H, 0:W].astype(np.float64)
    ray = np.stack([(uu - cx) / f, (vv - cy) / f, np.ones_like(uu)]).astype(np.float32)
    g = np.r_[np.eye(3).ravel(), [0, 0, 0], [f, 0, 0, 0, cx, cy]]
    src = (2.0 * ref + 0.25).astype(np.float32)
    inv = np.full((H, W), 0.1, np.float32)
    out = oracle.ncc_warp(ref, mu, sd, src, ray, g, inv, r)[0]
    inner = out[r:H - r - 1, r:W - r - 1]
    assert np.all(np.abs(inner - 1.0) < 1e-3), float(np.abs(inner - 1).max())
    # the last row and column cannot be sampled (u > w - 1.001): windows that reach them have at
    # most 2r/(2r+1) = 80 % valid taps, below 90 %
    assert np.all(out[:, W - 1 - r:] == -1.0) and np.all(out[H - 1 - r:, :] == -1.0)


def test_combine_topk_and_argmax_brute_force():
    rng = np.random.default_rng(3)
    n, H, W, D = 5, 6, 7, 9
    scores = rng.uniform(-1, 1, (D, n, H, W)).astype(np.float32)
    scores[3] = scores[2]  # exact ties between hypotheses 2 and 3
    for k in (1, 2, 3, 5):
        st, idx = oracle.new_state(H, W)
        for i in range(D):
            oracle.combine(scores[i], k, i, *st, idx)
        s = np.sort(scores, axis=1)[:, ::-1, :, :][:, :k].astype(np.float64).sum(1) / k
        best_i = np.argmax(s, axis=0)  # first index on ties
        assert np.array_equal(idx, best_i)
        assert np.allclose(st[0], s.max(0), atol=1e-6)
        ii, yy, xx = best_i, *np.mgrid[0:H, 0:W]
        prev_exp = np.where(ii > 0, s[np.maximum(ii - 1, 0), yy, xx], -2)
        next_exp = np.where(ii < D - 1, s[np.minimum(ii + 1, D - 1), yy, xx], -2)
        assert np.allclose(st[2], prev_exp, atol=1e-6) and np.allclose(st[3], next_exp, atol=1e-6)


def test_parabola_recovers_vertex():
    """Scores sampled from a parabola: the sub-step refinement returns its vertex exactly."""
    H, W, n = 4, 5, 11
    rng = np.random.default_rng(0)
    vert = rng.uniform(2.0, 8.0, (H, W))
    inv0 = np.full((H, W), 0.01, np.float32)
    inv1 = np.full((H, W), 0.0125, np.float32)
    i = np.arange(n)[:, None, None]
    s = (0.9 - 0.01 * (i - vert[None]) ** 2).astype(np.float32)
    idx = np.argmax(s, 0).astype(np.int32)
    yy, xx = np.mgrid[0:H, 0:W]
    best, spb, snb = s[idx, yy, xx], s[idx - 1, yy, xx], s[idx + 1, yy, xx]
    depth = oracle.sweep_finalize(inv0, inv1, n, best, spb, snb, idx)
    exp = 1.0 / (0.01 + vert * (np.float32(0.0125) - np.float32(0.01)))
    assert np.allclose(depth, exp, rtol=2e-4)


# ------------------------------------------------------------------------ helpers
def test_hyp_and_box_match_numpy_closed_forms():
    """The NumPy expressions the reference's kernels replace (test_native_parity.py:401-422)."""
    rng = np.random.default_rng(14)
    for H, W in ((37, 53), (3, 2), (120, 90)):
        base = rng.uniform(0.01, 0.05, (H, W)).astype(np.float32)
        step = rng.uniform(1e-5, 1e-3, (H, W)).astype(np.float32)
        for i in (0, 1, 7, 191):
            assert _same((base + i * step).astype(np.float32), oracle.hyp_linear(base, step, i))
        iu = rng.uniform(-1e-3, 0.05, (H, W)).astype(np.float32)
        iu[0, 0] = np.nan
        for t in (-3, -1, 0, 2, 3):
            assert _same(np.maximum(iu + t * step, 1e-6).astype(np.float32), oracle.hyp_refine(iu, step, t))
        a = rng.normal(0, 1, (H, W)).astype(np.float32)
        a[a > 1.5] = -0.0
        for x in (a, a * a):
            for r in (0, 1, 5, 7):
                k = 2 * r + 1
                c = np.cumsum(np.pad(x, ((r + 1, r), (0, 0))), 0)
                a1 = c[k:] - c[:-k]
                c = np.cumsum(np.pad(a1, ((0, 0), (r + 1, r))), 1)
                ref = (c[:, k:] - c[:, :-k]) * np.float32(1.0 / (k * k))
                assert _same(ref, oracle.box_zero_mean(x, r)), (H, W, r)


def test_half_matches_numpy():
    rng = np.random.default_rng(2)
    for H, W in ((10, 12), (11, 13), (2, 3)):
        g = rng.normal(0, 1, (H, W)).astype(np.float32)
        h, w = H // 2 * 2, W // 2 * 2
        gg = g[:h, :w]
        exp = ((gg[0::2, 0::2] + gg[1::2, 0::2] + gg[0::2, 1::2] + gg[1::2, 1::2]) * 0.25).astype(np.float32)
        assert _same(exp, oracle.half(g))


def test_rays_grid_pinhole_and_round_trip():
    H, W, f, cx, cy = 30, 40, 55.0, 19.5, 14.5
    x, y = oracle.rays_grid(H, W, cx, cy, f, 0.0, 0.0, 0.0)
    vv, uu = np.mgrid[0:H, 0:W].astype(np.float64)
    assert _same(x, (uu - cx) / f) and _same(y, (vv - cy) / f)
    k1, k2, k3 = -0.05, 0.01, 0.002
    x, y = oracle.rays_grid(H, W, cx, cy, f, k1, k2, k3)
    q = x * x + y * y
    d = f * (1 + q * (k1 + q * (k2 + q * k3)))
    assert np.max(np.abs(d * x + cx - uu)) < 1e-6 and np.max(np.abs(d * y + cy - vv)) < 1e-6


# ------------------------------------------------------------------------ PatchMatch
def test_xorshift32_published_value():
    """Marsaglia (2003), 'Xorshift RNGs': y = 2463534242 -> first output 723471715."""
    assert oracle.xorshift32_step(2463534242) == 723471715
    # the draw is the low 24 bits of the new state over 2^24, from the state seed | 1
    s1 = oracle.xorshift32_step(2463534243)
    assert oracle.xorshift_draws(2463534242, 1)[0] == (s1 & 0xFFFFFF) / 16777216.0


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, f"pm_refine_{name}.npz"))
    S = z["par"].shape[0]
    srcs = [z[f"src{s}"] for s in range(S)]
    deps = [z[f"dep{s}"] if z["has_depth"][s] else None for s in range(S)]
    r, step, top_k, geo_w, geo_tol, dmin, dmax, seed = z["params"]
    params = dict(r=int(r), step=int(step), top_k=int(top_k), geo_w=float(geo_w), geo_tol=float(geo_tol),
                  dmin=float(dmin), dmax=float(dmax), seed=int(seed))
    return z, srcs, deps, params


GOLDEN_CASES = ["s3", "s8", "s5_nogeo"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_pm_refine_matches_reference_golden(name):
    z, srcs, deps, p = load_golden(name)
    for iters in z["iters"]:
        d, n, s = z["depth0"].copy(), z["normal0"].copy(), z["score0"].copy()
        ch = oracle.pm_refine(z["ref"], z["ray"], z["Cr"], srcs, deps, z["par"], d, n, s, z["active"], p["r"],
                              p["step"], p["top_k"], p["geo_w"], p["geo_tol"], int(iters), p["dmin"], p["dmax"],
                              p["seed"])
        assert ch == int(z[f"changed_it{iters}"]), (name, iters)
        assert _same(d, z[f"depth_it{iters}"]) and _same(n, z[f"normal_it{iters}"]), (name, iters)
        assert _same(s, z[f"score_it{iters}"]), (name, iters)


def test_pm_refine_planar_fixed_point():
    """On a textured plane seen by exact cameras with exact source depth maps, the true plane has
    zero geometric error and near-perfect NCC, so PatchMatch keeps it: every pixel that starts at
    the true plane stays there."""
    z, srcs, deps, p = load_golden("s3")
    H, W = z["ref"].shape
    # true depth and normal of the plane z = 0.02 x + 0.01 y seen from the reference camera
    Cr = z["Cr"]
    ray = z["ray"].astype(np.float64)
    a, b = 0.02, 0.01
    t = (a * Cr[0] + b * Cr[1] - Cr[2]) / (ray[2] - a * ray[0] - b * ray[1])
    nrm = np.array([-a, -b, 1.0]) / np.linalg.norm([-a, -b, 1.0])
    if (nrm[:, None, None] * ray).sum(0).mean() > 0:
        nrm = -nrm
    depth = t.astype(np.float32)
    normal = np.ascontiguousarray(np.broadcast_to(nrm[:, None, None], (3, H, W)), np.float32)
    score = np.full((H, W), -1.0, np.float32)
    active = np.ones((H, W), np.uint8)
    d0 = depth.copy()
    oracle.pm_refine(z["ref"], z["ray"], Cr, srcs, deps, z["par"], depth, normal, score, active, p["r"], p["step"],
                     p["top_k"], p["geo_w"], p["geo_tol"], 2, p["dmin"], p["dmax"], p["seed"])
    inner = np.s_[6:-6, 6:-6]
    rel = np.abs(depth[inner] - d0[inner]) / d0[inner]
    assert np.nanmedian(rel) < 1e-3 and np.nanmean(rel < 5e-3) > 0.95
    assert np.nanmedian(score[inner]) > 0.8


# ------------------------------------------------------------------ height-sweep sampler
SV_GOLDEN = os.path.join(GOLDEN, "sample_view.npz")


def sample_view_cases():
    """(name, img, cam, X0, Y0, gsd, Z, out, valid) from the compiled reference's CPUBackend.sample_view
    (tests/golden/make_sample_view_golden.py)."""
    z = np.load(SV_GOLDEN)
    for name in z["names"]:
        p = z[f"{name}_par"]
        cam = (z[f"{name}_R"], z[f"{name}_C"], *p[:6])
        yield (str(name), z[f"{name}_img"], cam, p[6], p[7], p[8], z[f"{name}_Z"], z[f"{name}_out"],
               z[f"{name}_valid"])


@pytest.mark.parametrize("case", list(sample_view_cases()), ids=lambda c: c[0])
def test_sample_view_matches_reference_golden(case):
    """orc_sample_view == _mvs.sample_view (_mvs.pyx:13-54) bit for bit: distortion k1..k3 != 0,
    tilted and oblique cameras, fractional positions, (D, H, W) stacks, points behind the camera,
    and cells exactly on the last pixel row and column (the x1 = x0, y1 = y0 clamps)."""
    name, img, cam, X0, Y0, gsd, Z, want, wvalid = case
    out, valid = oracle.sample_view(img, cam, X0, Y0, gsd, Z)
    assert np.array_equal(valid, wvalid)
    assert _same(out, want)
    assert valid.any() and not valid.all()


def test_np_init_normals_matches_reference_golden():
    """cases.np_init_normals (the checker of the device normals) == densify._init_normals's output
    recorded from the reference (tests/golden/depthmap_grid3x3.npz), bit for bit."""
    z = np.load(os.path.join(GOLDEN, "depthmap_grid3x3.npz"))
    for k in z["views"]:
        got = cases.np_init_normals(z[f"depth_{k}"], z[f"ray_{k}"])
        assert _same(got, z[f"normals_{k}"])


def test_combine_many_sources_sums_top_k_in_descending_order():
    """orc_combine with more than 16 sources: the mean of the top k scores summed largest first in
    float32 (the loop of _dense.pyx:124-133 for any n; the reference's own buffer holds 16)."""
    rng = np.random.default_rng(8)
    n, H, W, D = 23, 5, 6, 7
    scores = rng.uniform(-1, 1, (D, n, H, W)).astype(np.float32)
    for k in (1, 6, 23):
        st, idx = oracle.new_state(H, W)
        means = np.empty((D, H, W), np.float32)
        for i in range(D):
            oracle.combine(scores[i], k, i, *st, idx)
            srt = -np.sort(-scores[i], axis=0)
            acc = np.zeros((H, W), np.float32)
            for a in range(k):
                acc = (acc + srt[a]).astype(np.float32)
            means[i] = (acc / np.float32(k)).astype(np.float32)
            assert _same(st[1], means[i])                       # prev = this hypothesis's mean
        assert np.array_equal(idx, np.argmax(means, axis=0))
        assert _same(st[0], means.max(0))
