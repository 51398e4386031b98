"""GPU plane sweep (libom355.so) against the oracle on the reference's parity shapes, larger tiled
/ chunked cases, the 640x480 four-source scene at its full size, every hypothesis generator, and
the parabola.

Bar (tests/sweep_bar.py): every score within 1e-4 of the oracle's, the winner identical except
where the oracle's two best scores differ by less than 1e-4, the validity mask identical, NaNs at
the same pixels.  Bit-identity is reported separately (it holds whenever no FP64 window sum
rounds, i.e. in practice on these inputs).
"""
import numpy as np
import pytest
import torch

import cases
import oracle
import sweep_bar

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


def _state_np(st):
    return [t.cpu().numpy() for t in (st.best, st.prev, st.s_prev_best, st.s_next_best)], st.idx.cpu().numpy()


def _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k):
    return sweep_bar.oracle_scores(ref, mu, sd, srcs, ray, geo, hyps, r, k)


def _check_state(orc, st, idx, label):
    st0, idx0, S = orc
    stats = sweep_bar.check(S, st0, idx0, st, idx, label)
    print("sweep bar", stats)
    return stats


# (H, W, sources, r, top_k, hypotheses); the last five are one pixel, one row or one column wide
# (r = 0 gives them scores other than -1; with r >= 1 every window there is under 90 % covered)
@pytest.mark.parametrize("shape", cases.SWEEP_SHAPES + [(120, 150, 8, 5, 2, 6), (200, 333, 4, 2, 2, 12),
                                                       (97, 140, 12, 7, 3, 5), (64, 64, 16, 1, 4, 4),
                                                       (1, 1, 2, 0, 1, 3), (1, 70, 3, 0, 2, 4), (70, 1, 3, 0, 2, 4),
                                                       (3, 70, 2, 1, 2, 5), (1, 70, 3, 2, 2, 4)])
def test_fused_sweep_matches_oracle(shape):
    H, W, n, r, k, nh = shape
    rng = np.random.default_rng(100 + H * 7 + W)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
    orc = _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    from py_ortho_amd import dense
    dsrc = [_dev(s) for s in srcs]
    maps = _dev(np.stack(hyps))
    for chunks, tile_h in ((1, 0), (0, 0), (3, 0), (2, 8), (1, 16)):  # 0: the default 32x32 tiles
        st = dense.sweep(_dev(ref), _dev(mu), _dev(sd), dsrc, _dev(ray), geo, {"mode": "maps", "maps": maps}, r, k,
                         chunks=chunks, tile_h=tile_h)
        torch.cuda.synchronize()
        a, ia = _state_np(st)
        _check_state(orc, a, ia, (shape, chunks, tile_h))


def test_sweep_hypothesis_per_call_matches_oracle():
    rng = np.random.default_rng(5)
    from py_ortho_amd import dense
    for H, W, n, r, k, nh in cases.SWEEP_SHAPES:
        ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
        orc = _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k)
        st = dense.SweepState(H, W)
        args = [_dev(a) for a in (ref, mu, sd)]
        dsrc = [_dev(s) for s in srcs]
        dray = _dev(ray)
        for i, inv in enumerate(hyps):
            dense.sweep_hypothesis(*args, dsrc, dray, geo, _dev(inv), r, k, i, *st.tensors())
        torch.cuda.synchronize()
        a, ia = _state_np(st)
        _check_state(orc, a, ia, (H, W, n, r))


def test_sweep_flat_sources_keep_first_winner():
    # equal scores everywhere: ties keep the first hypothesis (test_native_parity.py:107-112)
    rng = np.random.default_rng(9)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, 20, 24, 3, 2, 5, 0.0)
    flat = [np.ones_like(s) for s in srcs]
    orc = _oracle(ref, mu, sd, flat, ray, geo, hyps, 2, 2)
    from py_ortho_amd import dense
    st = dense.sweep(_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in flat], _dev(ray), geo,
                     {"mode": "maps", "maps": _dev(np.stack(hyps))}, 2, 2)
    a, ia = _state_np(st)
    # ties keep the first hypothesis: exact, not up to near-ties
    assert _same(ia, orc[1]) and all(_same(x, y) for x, y in zip(a, orc[0]))


def test_generators_and_finalize():
    from py_ortho_amd import dense
    rng = np.random.default_rng(11)
    H, W, n, r, k = 70, 90, 4, 2, 2
    ref, mu, sd, srcs, ray, geo, _ = cases.sweep_case(rng, H, W, n, r, 1)
    base = rng.uniform(0.02, 0.05, (H, W)).astype(np.float32)
    step = rng.uniform(1e-4, 1e-3, (H, W)).astype(np.float32)
    nd = 24
    hyps = [oracle.hyp_linear(base, step, i) for i in range(nd)]
    orc = _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st0, idx0 = orc[0], orc[1]
    d0 = oracle.sweep_finalize(hyps[0], hyps[1], nd, st0[0], st0[2], st0[3], idx0)
    args = (_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo)
    lin = {"mode": "linear", "base": _dev(base), "step": _dev(step), "n": nd}
    st = dense.sweep(*args, lin, r, k)
    a, ia = _state_np(st)
    if _check_state(orc, a, ia, "linear")["bit_identical_state"]:
        d1 = dense.sweep_finalize(*args, lin, r, k, st).cpu().numpy()
        assert _same(d0, d1)
    # refine generator: max(iu + f32(q - rs) * ds, 1e-6), q = 0 .. 2 rs
    iu = rng.uniform(-1e-3, 0.05, (H, W)).astype(np.float32)
    iu[0, 0] = np.nan
    rs = 3
    hyps = [oracle.hyp_refine(iu, step, q - rs) for q in range(2 * rs + 1)]
    orc = _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st0, idx0 = orc[0], orc[1]
    d0 = oracle.sweep_finalize(hyps[0], hyps[1], len(hyps), st0[0], st0[2], st0[3], idx0)
    rf = {"mode": "refine", "iu": _dev(iu), "ds": _dev(step), "steps": rs}
    st = dense.sweep(*args, rf, r, k)
    a, ia = _state_np(st)
    if _check_state(orc, a, ia, "refine")["bit_identical_state"]:
        d1 = dense.sweep_finalize(*args, rf, r, k, st).cpu().numpy()
        assert _same(d0, d1)
    # constant planes (fronto-parallel sweep)
    inv = (1.0 / np.linspace(2.0, 30.0, 17)).astype(np.float32)
    hyps = [np.full((H, W), v, np.float32) for v in inv]
    orc = _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    for tile_h in (0, 16, 8):  # 0: 32x32 tiles
        st = dense.sweep(*args, {"mode": "const", "inv": _dev(inv)}, r, k, tile_h=tile_h)
        a, ia = _state_np(st)
        _check_state(orc, a, ia, ("const", tile_h))


@pytest.mark.parametrize("k", [1, 3, 4])
def test_const_planes_r5_topk_variants(k):
    """Constant planes at r = 5 with top-1 (the four-workgroup kernels, halo indices recomputed)
    and top-3 / top-4 (halo indices in LDS, three workgroups), 32 x 32 and 64 x 16 tiles, on an
    image with interior and border tiles, against the oracle."""
    from py_ortho_amd import dense
    rng = np.random.default_rng(40 + k)
    H, W, n, r = 100, 130, 6, 5
    ref, mu, sd, srcs, ray, geo, _ = cases.sweep_case(rng, H, W, n, r, 1)
    inv = (1.0 / np.linspace(2.0, 30.0, 13)).astype(np.float32)
    hyps = [np.full((H, W), v, np.float32) for v in inv]
    orc = _oracle(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    args = (_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo)
    for tile_h, chunks in ((0, 0), (16, 1), (0, 3)):
        st = dense.sweep(*args, {"mode": "const", "inv": _dev(inv)}, r, k, tile_h=tile_h, chunks=chunks)
        a, ia = _state_np(st)
        _check_state(orc, a, ia, ("const r5", k, tile_h, chunks))


def test_depth_helpers_bit_exact():
    from py_ortho_amd import dense
    rng = np.random.default_rng(14)
    for H, W in ((37, 53), (3, 2), (120, 90), (481, 643)):
        base = rng.uniform(0.01, 0.05, (H, W)).astype(np.float32)
        step = rng.uniform(1e-5, 1e-3, (H, W)).astype(np.float32)
        for i in (0, 1, 7, 191):
            assert _same(oracle.hyp_linear(base, step, i), dense.hyp_linear(_dev(base), _dev(step), i).cpu().numpy())
        iu = rng.uniform(-1e-3, 0.05, (H, W)).astype(np.float32)
        iu[0, 0] = np.nan
        for t in (-3, -1, 0, 2, 3):
            assert _same(oracle.hyp_refine(iu, step, t), dense.hyp_refine(_dev(iu), _dev(step), t).cpu().numpy())
        a = rng.normal(0, 1, (H, W)).astype(np.float32)
        a[a > 1.5] = -0.0
        for r in (0, 1, 5, 7):
            assert _same(oracle.box_zero_mean(a, r), dense.box_zero_mean(_dev(a), r).cpu().numpy()), (H, W, r)
            mu0, sd0 = oracle.ref_stats(a, r)
            mu1, sd1 = dense.ref_stats(_dev(a), r)
            assert _same(mu0, mu1.cpu().numpy()) and _same(sd0, sd1.cpu().numpy()), (H, W, r)
        assert _same(oracle.half(a), dense.half(_dev(a)).cpu().numpy())


def test_c1_full_size_sweep_meets_bar():
    """The plain plane sweep at its full size: the seeded terrain scene rendered into 5 nadir
    cameras (640x480, f = 600, 80 m, 15 m cross baseline), the centre view against 4 sources, 128
    fronto-parallel planes uniform in depth over 60-85 m, 5x5 window, top-2, against the oracle
    under the bar of tests/sweep_bar.py; then the parabola depth."""
    from py_ortho_amd import dense, scene
    sc = scene.config_c1(device="cuda")
    r, k, D = 2, 2, 128
    ref = sc.ref.img.contiguous()
    H, W = ref.shape
    assert (H, W) == (480, 640) and len(sc.srcs) == 4
    ray = scene.rays_world(sc.ref.cam, "cuda")
    mu, sd = dense.ref_stats(ref, r)
    inv = (1.0 / torch.linspace(60.0, 85.0, D, dtype=torch.float64)).to(torch.float32)
    srcs = [v.img.contiguous() for v in sc.srcs]
    geo = sc.geo()
    hyp = {"mode": "const", "inv": inv.cuda()}
    st = dense.sweep(ref, mu, sd, srcs, ray, geo, hyp, r, k)
    depth = dense.sweep_finalize(ref, mu, sd, srcs, ray, geo, hyp, r, k, st).cpu().numpy()
    torch.cuda.synchronize()
    a, ia = _state_np(st)
    c = lambda t: t.cpu().numpy()  # noqa: E731
    hyps = [np.full((H, W), v, np.float32) for v in inv.numpy()]
    orc = _oracle(c(ref), c(mu), c(sd), [c(s) for s in srcs], c(ray), geo, hyps, r, k)
    stats = _check_state(orc, a, ia, "c1 640x480 S=4 D=128 r=2")
    st0, idx0 = orc[0], orc[1]
    d0 = oracle.sweep_finalize(hyps[0], hyps[1], D, st0[0], st0[2], st0[3], idx0)
    if stats["bit_identical_state"]:
        assert _same(depth, d0)
    else:
        rel = np.abs(depth.astype(np.float64) - d0) / d0
        assert np.mean(rel <= 2e-3) >= 0.99, float(np.mean(rel <= 2e-3))
    # the sweep finds the terrain: the winning plane (planes uniform in depth, so the parabola's
    # linear-in-inverse-depth step does not apply) within one plane spacing of the rendered depth
    truth = c(sc.ref.depth)
    planes = np.linspace(60.0, 85.0, D)
    ok = (st0[0] >= sweep_bar.MIN_NCC) & np.isfinite(truth)
    err = np.abs(planes[ia[ok]] - truth[ok])
    assert np.median(err) < 25.0 / 127, float(np.median(err))
