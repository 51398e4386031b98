"""GPU plane sweep (libom355.so) against the oracle: bit-exact winner state on the reference's
parity shapes and larger tiled / chunked cases, every hypothesis generator, and the parabola.

Bar: identical float32 bits and indices; where FP64 window sums round differently (values whose
magnitudes span more than 2^29 inside one window) the score may differ by an FP64 rounding, so
a case that is not bit-identical must still satisfy |score - oracle| <= 1e-4 with identical
winners wherever the two best scores are more than 1e-4 apart.
"""
import numpy as np
import pytest
import torch

import cases
import oracle

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


def _check_state(ref_st, ref_idx, st, idx, label):
    if _same(ref_idx, idx) and all(_same(a, b) for a, b in zip(ref_st, st)):
        return
    # tolerance bar (see module docstring)
    d = np.abs(ref_st[0] - st[0])
    assert np.nanmax(d) <= 1e-4, (label, float(np.nanmax(d)))
    raise AssertionError(f"{label}: not bit-identical (max |d best| {float(np.nanmax(d)):.3g}, "
                         f"idx mismatches {int((ref_idx != idx).sum())})")


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
    (st0, idx0) = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    from py_ortho_amd import dense
    dsrc = [_dev(s) for s in srcs]
    maps = _dev(np.stack(hyps))
    for chunks, tile_h in ((1, 0), (0, 0), (3, 0), (2, 8)):  # 0: the default 16-row tiles
        st = dense.sweep(_dev(ref), _dev(mu), _dev(sd), dsrc, _dev(ray), geo, {"mode": "maps", "maps": maps}, r, k,
                         chunks=chunks, tile_h=tile_h)
        torch.cuda.synchronize()
        a, ia = _state_np(st)
        _check_state(st0, idx0, a, ia, (shape, chunks, tile_h))


def test_sweep_hypothesis_per_call_matches_oracle():
    rng = np.random.default_rng(5)
    from py_ortho_amd import dense
    for H, W, n, r, k, nh in cases.SWEEP_SHAPES:
        ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
        st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
        st = dense.SweepState(H, W)
        args = [_dev(a) for a in (ref, mu, sd)]
        dsrc = [_dev(s) for s in srcs]
        dray = _dev(ray)
        for i, inv in enumerate(hyps):
            dense.sweep_hypothesis(*args, dsrc, dray, geo, _dev(inv), r, k, i, *st.tensors())
        torch.cuda.synchronize()
        a, ia = _state_np(st)
        _check_state(st0, idx0, a, ia, (H, W, n, r))


def test_sweep_flat_sources_keep_first_winner():
    # equal scores everywhere: ties keep the first hypothesis (test_native_parity.py:107-112)
    rng = np.random.default_rng(9)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, 20, 24, 3, 2, 5, 0.0)
    flat = [np.ones_like(s) for s in srcs]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, flat, ray, geo, hyps, 2, 2)
    from py_ortho_amd import dense
    st = dense.sweep(_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in flat], _dev(ray), geo,
                     {"mode": "maps", "maps": _dev(np.stack(hyps))}, 2, 2)
    a, ia = _state_np(st)
    _check_state(st0, idx0, a, ia, "flat")


def test_generators_and_finalize():
    from py_ortho_amd import dense
    rng = np.random.default_rng(11)
    H, W, n, r, k = 70, 90, 4, 2, 2
    ref, mu, sd, srcs, ray, geo, _ = cases.sweep_case(rng, H, W, n, r, 1)
    base = rng.uniform(0.02, 0.05, (H, W)).astype(np.float32)
    step = rng.uniform(1e-4, 1e-3, (H, W)).astype(np.float32)
    nd = 24
    hyps = [oracle.hyp_linear(base, step, i) for i in range(nd)]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    d0 = oracle.sweep_finalize(hyps[0], hyps[1], nd, st0[0], st0[2], st0[3], idx0)
    args = (_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo)
    lin = {"mode": "linear", "base": _dev(base), "step": _dev(step), "n": nd}
    st = dense.sweep(*args, lin, r, k)
    a, ia = _state_np(st)
    _check_state(st0, idx0, a, ia, "linear")
    d1 = dense.sweep_finalize(*args, lin, r, k, st).cpu().numpy()
    assert _same(d0, d1)
    # refine generator: max(iu + f32(q - rs) * ds, 1e-6), q = 0 .. 2 rs
    iu = rng.uniform(-1e-3, 0.05, (H, W)).astype(np.float32)
    iu[0, 0] = np.nan
    rs = 3
    hyps = [oracle.hyp_refine(iu, step, q - rs) for q in range(2 * rs + 1)]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    d0 = oracle.sweep_finalize(hyps[0], hyps[1], len(hyps), st0[0], st0[2], st0[3], idx0)
    rf = {"mode": "refine", "iu": _dev(iu), "ds": _dev(step), "steps": rs}
    st = dense.sweep(*args, rf, r, k)
    a, ia = _state_np(st)
    _check_state(st0, idx0, a, ia, "refine")
    d1 = dense.sweep_finalize(*args, rf, r, k, st).cpu().numpy()
    assert _same(d0, d1)
    # constant planes (fronto-parallel sweep)
    inv = (1.0 / np.linspace(2.0, 30.0, 17)).astype(np.float32)
    hyps = [np.full((H, W), v, np.float32) for v in inv]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st = dense.sweep(*args, {"mode": "const", "inv": _dev(inv)}, r, k)
    a, ia = _state_np(st)
    _check_state(st0, idx0, a, ia, "const")


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
