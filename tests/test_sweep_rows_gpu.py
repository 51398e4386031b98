"""The row/column-streaming sweep (om_sweep_rows): the reference's ncc_warp + combine order, so its
winner state equals the oracle's bit for bit on every input -- also where the fused kernel's short
window sums round differently -- and it takes what the fused kernels do not: more than 16 sources,
window radii above 7 and top-k above 4 (densify.py:343-353, _dense.pyx:13-144, :759-771)."""
import numpy as np
import pytest
import torch

import cases
import oracle

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _state(st):
    return [t.cpu().numpy() for t in (st.best, st.prev, st.s_prev_best, st.s_next_best)], st.idx.cpu().numpy()


def _assert_same(a, ia, st0, idx0, label):
    assert np.array_equal(ia, idx0), (label, int((ia != idx0).sum()))
    for x, y, nm in zip(a, st0, ("best", "prev", "s_prev_best", "s_next_best")):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (label, nm)


def _oracle_unfused(ref, mu, sd, srcs, ray, geo, hyps, r, k):
    """ncc_warp per source + combine, hypothesis by hypothesis (any source count)."""
    H, W = ref.shape
    st, idx = oracle.new_state(H, W)
    for i, inv in enumerate(hyps):
        S = np.stack([oracle.ncc_warp(ref, mu, sd, s, ray, geo[j], inv, r)[0] for j, s in enumerate(srcs)])
        oracle.combine(S, k, i, *st, idx)
    return st, idx


# (H, W, sources, r, top_k, hypotheses, chunks)
@pytest.mark.parametrize("shape", [(37, 53, 4, 5, 2, 9, 0), (64, 48, 1, 3, 2, 5, 2), (30, 31, 5, 2, 3, 6, 4),
                                   (70, 90, 4, 9, 2, 5, 0), (40, 57, 3, 12, 1, 4, 3), (33, 41, 20, 2, 6, 4, 0),
                                   (25, 30, 18, 4, 18, 3, 2), (12, 9, 7, 0, 5, 3, 1), (1, 40, 3, 1, 2, 3, 0)])
def test_rows_sweep_matches_oracle(shape):
    from py_ortho_amd import dense
    H, W, n, r, k, nh, chunks = shape
    rng = np.random.default_rng(500 + H * 3 + W + n)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
    if n <= 16:
        st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    else:
        st0, idx0 = _oracle_unfused(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st = dense.sweep(_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo,
                     {"mode": "maps", "maps": _dev(np.stack(hyps))}, r, k, chunks=chunks, exact=True)
    torch.cuda.synchronize()
    _assert_same(*_state(st), st0, idx0, shape)
    if not dense.fused_ok(n, r, k):
        # the same call without exact=True also takes this path (the fused kernels do not)
        st2 = dense.sweep(_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo,
                          {"mode": "maps", "maps": _dev(np.stack(hyps))}, r, k)
        _assert_same(*_state(st2), st0, idx0, (shape, "auto"))


def test_rows_generators_and_single_hypothesis():
    """linear, refine and constant generators, and sweep_hypothesis with r = 8 (outside the fused
    kernels), against the oracle."""
    from py_ortho_amd import dense
    rng = np.random.default_rng(21)
    H, W, n, r, k = 45, 61, 4, 8, 2
    ref, mu, sd, srcs, ray, geo, _ = cases.sweep_case(rng, H, W, n, r, 1)
    base = rng.uniform(0.02, 0.05, (H, W)).astype(np.float32)
    step = rng.uniform(1e-4, 1e-3, (H, W)).astype(np.float32)
    args = (_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo)
    nd = 11
    hyps = [oracle.hyp_linear(base, step, i) for i in range(nd)]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st = dense.sweep(*args, {"mode": "linear", "base": _dev(base), "step": _dev(step), "n": nd}, r, k)
    _assert_same(*_state(st), st0, idx0, "linear")
    d0 = oracle.sweep_finalize(hyps[0], hyps[1], nd, st0[0], st0[2], st0[3], idx0)
    d1 = dense.sweep_finalize(*args, {"mode": "linear", "base": _dev(base), "step": _dev(step), "n": nd}, r, k, st)
    assert np.array_equal(d1.cpu().numpy().view(np.uint32), d0.view(np.uint32))
    iu = rng.uniform(0.01, 0.05, (H, W)).astype(np.float32)
    rs = 2
    hyps = [oracle.hyp_refine(iu, step, q - rs) for q in range(2 * rs + 1)]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st = dense.sweep(*args, {"mode": "refine", "iu": _dev(iu), "ds": _dev(step), "steps": rs}, r, k)
    _assert_same(*_state(st), st0, idx0, "refine")
    inv = (1.0 / np.linspace(2.0, 30.0, 7)).astype(np.float32)
    hyps = [np.full((H, W), v, np.float32) for v in inv]
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st = dense.sweep(*args, {"mode": "const", "inv": _dev(inv)}, r, k)
    _assert_same(*_state(st), st0, idx0, "const")
    # one hypothesis at a time (the reference's sweep_hypothesis takes any r)
    st1 = dense.SweepState(H, W)
    for i, h in enumerate(hyps):
        dense.sweep_hypothesis(*args[:3], args[3], args[4], geo, _dev(h), r, k, i, *st1.tensors())
    _assert_same(*_state(st1), st0, idx0, "per hypothesis")


def test_rows_sweep_exact_where_fused_sums_round():
    """Grey values of both signs near zero (the reference's matching image, (v - 128) / 64) make
    FP64 partial window sums round; the row/column path still equals the oracle bit for bit."""
    from py_ortho_amd import dense
    rng = np.random.default_rng(77)
    H, W, n, r, k, nh = 96, 128, 4, 3, 2, 24
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
    # quantised grey around 128 with tiny bilinear results near zero crossings
    ref = ((np.round(rng.normal(128, 6, (H, W))) - 128) / 64).astype(np.float32)
    srcs = [((np.round(rng.normal(128, 6, s.shape)) - 128) / 64).astype(np.float32) for s in srcs]
    mu, sd = cases.box_stats(ref, r)
    st0, idx0 = oracle.sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, k)
    st = dense.sweep(_dev(ref), _dev(mu), _dev(sd), [_dev(s) for s in srcs], _dev(ray), geo,
                     {"mode": "maps", "maps": _dev(np.stack(hyps))}, r, k, exact=True)
    _assert_same(*_state(st), st0, idx0, "signed grey")
