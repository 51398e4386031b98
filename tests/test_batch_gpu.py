"""Batched views (om_batch): the persistent multi-view sweep and the batched PatchMatch equal the
per-view calls bit for bit, and a HIP-graph replay equals the eager run on every replay."""
import numpy as np
import pytest
import torch

import cases
import oracle
from golden.make_pm_golden import make_case

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


def _sweep_views(seed=0, r=3, k=2):
    rng = np.random.default_rng(seed)
    out = []
    for (H, W, n, nh, mode) in [(37, 53, 4, 9, "maps"), (120, 150, 8, 12, "linear"), (64, 200, 3, 7, "refine"),
                                (200, 90, 12, 5, "const"), (16, 16, 2, 3, "maps")]:
        ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, H, W, n, r, nh)
        if mode == "maps":
            hyp = {"mode": "maps", "maps": _dev(np.stack(hyps))}
            ho = hyps
        elif mode == "linear":
            base = rng.uniform(0.02, 0.05, (H, W)).astype(np.float32)
            step = rng.uniform(1e-4, 1e-3, (H, W)).astype(np.float32)
            hyp = {"mode": "linear", "base": _dev(base), "step": _dev(step), "n": nh}
            ho = [oracle.hyp_linear(base, step, i) for i in range(nh)]
        elif mode == "refine":
            iu = rng.uniform(0.01, 0.05, (H, W)).astype(np.float32)
            ds = rng.uniform(1e-4, 1e-3, (H, W)).astype(np.float32)
            rs = (nh - 1) // 2
            hyp = {"mode": "refine", "iu": _dev(iu), "ds": _dev(ds), "steps": rs}
            ho = [oracle.hyp_refine(iu, ds, q - rs) for q in range(2 * rs + 1)]
        else:
            inv = (1.0 / np.linspace(2.0, 30.0, nh)).astype(np.float32)
            hyp = {"mode": "const", "inv": _dev(inv)}
            ho = [np.full((H, W), v, np.float32) for v in inv]
        out.append(dict(ref=ref, mu=mu, sd=sd, srcs=srcs, ray=ray, geo=geo, hyp=hyp, hyps=ho))
    return out


@pytest.mark.parametrize("chunks,tile_h", [(0, 0), (1, 16), (3, 8), (2, 16), (2, 32), ("mixed", "mixed")])
def test_batch_sweep_equals_oracle_per_view(chunks, tile_h):
    """Every view equals the oracle's sweep + parabola; "mixed" gives each view its own tile height
    and chunk count (two persistent launches, one per tile height)."""
    from py_ortho_amd import batch
    r, k = 3, 2
    vs = _sweep_views(1, r, k)
    B = batch.Batch()
    bviews = []
    for j, v in enumerate(vs):
        H, W = v["ref"].shape
        bviews.append(batch.SweepView(_dev(v["ref"]), _dev(v["mu"]), _dev(v["sd"]), [_dev(s) for s in v["srcs"]],
                                      _dev(v["ray"]), v["geo"], v["hyp"],
                                      depth=torch.empty((H, W), dtype=torch.float32, device="cuda"),
                                      tile_h=(8, 16, 32)[j % 3] if chunks == "mixed" else 0,
                                      chunks=(1, 2, 3)[j % 3] if chunks == "mixed" else 0))
    if chunks == "mixed":
        B.set_sweep(bviews, r, k)
    else:
        B.set_sweep(bviews, r, k, chunks=chunks, tile_h=tile_h)
    B.run()
    torch.cuda.synchronize()
    for v, bv in zip(vs, bviews):
        st0, idx0 = oracle.sweep_all(v["ref"], v["mu"], v["sd"], v["srcs"], v["ray"], v["geo"], v["hyps"], r, k)
        assert _same(bv.state.idx.cpu().numpy(), idx0)
        for a, b in zip((bv.state.best, bv.state.prev, bv.state.s_prev_best, bv.state.s_next_best), st0):
            assert _same(a.cpu().numpy(), b)
        n = len(v["hyps"])
        d0 = oracle.sweep_finalize(v["hyps"][0], v["hyps"][1] if n > 1 else v["hyps"][0], n, st0[0], st0[2],
                                   st0[3], idx0)
        assert _same(bv.depth.cpu().numpy(), d0)
    B.close()


def _pm_views(specs):
    from py_ortho_amd import batch
    out, ref_args = [], []
    for seed, H, W, S, wd in specs:
        c = make_case(seed, H, W, S, wd)
        d, n, s = _dev(c["depth"]), _dev(c["normal"]), _dev(c["score"])
        out.append(batch.PMView(_dev(c["ref"]), _dev(c["ray"]), c["Cr"], [_dev(x) for x in c["srcs"]],
                                [None if x is None else _dev(x) for x in c["src_depths"]], c["par"], d, n, s,
                                _dev(c["active"]), 0.02, c["dmin"], c["dmax"], 1000 + seed))
        ref_args.append(c)
    return out, ref_args


def test_batch_pm_equals_per_view():
    from py_ortho_amd import batch
    r, step, k, gw, iters = 4, 2, 2, 0.5, 2
    specs = [(51, 30, 36, 3, True), (52, 25, 41, 8, True), (53, 33, 29, 5, False)]
    views, cs = _pm_views(specs)
    B = batch.Batch()
    B.set_pm(views, r, step, k, gw, iters)
    B.run()
    changed = B.pm_changed()
    for V, c, ch in zip(views, cs, changed):
        d0, n0, s0 = c["depth"].copy(), c["normal"].copy(), c["score"].copy()
        ch0 = oracle.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], d0, n0, s0,
                               c["active"], r, step, k, gw, 0.02, iters, c["dmin"], c["dmax"], V.seed)
        assert ch == ch0
        assert _same(V.depth.cpu().numpy(), d0) and _same(V.normal.cpu().numpy(), n0)
        assert _same(V.score.cpu().numpy(), s0)
    B.close()


def test_batch_pm_direct_path_equals_oracle(monkeypatch):
    """Batched PatchMatch on the direct per-pixel kernels (forced by lowering the memo limit)."""
    from py_ortho_amd import batch
    monkeypatch.setenv("OM_PM_MEMO_MAX_PIXELS", "0")
    r, step, k, gw, iters = 3, 2, 2, 0.5, 2
    specs = [(61, 27, 33, 4, True), (62, 31, 26, 7, False)]
    views, cs = _pm_views(specs)
    B = batch.Batch()
    B.set_pm(views, r, step, k, gw, iters)
    B.run()
    changed = B.pm_changed()
    for V, c, ch in zip(views, cs, changed):
        d0, n0, s0 = c["depth"].copy(), c["normal"].copy(), c["score"].copy()
        ch0 = oracle.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], d0, n0, s0,
                               c["active"], r, step, k, gw, 0.02, iters, c["dmin"], c["dmax"], V.seed)
        assert ch == ch0
        assert _same(V.depth.cpu().numpy(), d0) and _same(V.normal.cpu().numpy(), n0)
        assert _same(V.score.cpu().numpy(), s0)
    B.close()


def test_batch_graph_replays_match_eager():
    """Capture sweep + PatchMatch of a batch in a HIP graph; every replay (with eager device work
    and synchronisation between replays) reproduces the eager result exactly."""
    from py_ortho_amd import batch
    r, k = 3, 2
    vs = _sweep_views(2, r, k)
    specs = [(61, 30, 36, 3, True), (62, 26, 40, 4, False)]
    views, cs = _pm_views(specs)
    B = batch.Batch()
    sv = []
    for v in vs:
        H, W = v["ref"].shape
        sv.append(batch.SweepView(_dev(v["ref"]), _dev(v["mu"]), _dev(v["sd"]), [_dev(s) for s in v["srcs"]],
                                  _dev(v["ray"]), v["geo"], v["hyp"],
                                  depth=torch.empty((H, W), dtype=torch.float32, device="cuda")))
    B.set_sweep(sv, r, k)
    B.set_pm(views, 4, 2, 2, 0.5, 2)
    init = [(V.depth.clone(), V.normal.clone(), V.score.clone()) for V in views]

    def reset_pm():
        for V, (d, n, s) in zip(views, init):
            V.depth.copy_(d)
            V.normal.copy_(n)
            V.score.copy_(s)

    B.run()
    torch.cuda.synchronize()
    eager = [x.cpu().numpy() for bv in sv for x in (bv.state.best, bv.state.idx, bv.depth)]
    eager += [x.cpu().numpy() for V in views for x in (V.depth, V.normal, V.score)]
    B.capture()
    for rep in range(3):
        reset_pm()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        B.replay()
        torch.cuda.synchronize()
        got = [x.cpu().numpy() for bv in sv for x in (bv.state.best, bv.state.idx, bv.depth)]
        got += [x.cpu().numpy() for V in views for x in (V.depth, V.normal, V.score)]
        assert all(_same(a, b) for a, b in zip(got, eager)), rep
        del junk
    B.close()


@pytest.mark.parametrize("direct", [False, True])
def test_batch_pm_with_tiny_views_equals_oracle(monkeypatch, direct):
    """A batch that mixes reference images of a few pixels (1x1, 1x6, 4x1, 6x11, with windows
    clipped on every side and mostly empty colour tiles) with an ordinary view: the per-view block,
    row and pixel offsets of the batched kernels, on the memoised and the direct path."""
    from py_ortho_amd import batch
    if direct:
        monkeypatch.setenv("OM_PM_MEMO_MAX_PIXELS", "0")
    r, step, k, gw, iters = 2, 1, 2, 0.5, 2
    views, cs = [], []
    for j, (H, W) in enumerate([(1, 1), (1, 6), (30, 36), (4, 1), (6, 11)]):
        c = make_case(70 + j, 9, 11, 4, True) if H < 9 else make_case(70 + j, H, W, 4, True)
        cut = lambda a: np.array(a[..., :H, :W], order="C", copy=True)  # noqa: E731
        c = dict(c, ref=cut(c["ref"]), ray=cut(c["ray"]), depth=cut(c["depth"]), normal=cut(c["normal"]),
                 score=cut(c["score"]), active=np.ones((H, W), np.uint8))
        views.append(batch.PMView(_dev(c["ref"]), _dev(c["ray"]), c["Cr"], [_dev(x) for x in c["srcs"]],
                                  [None if x is None else _dev(x) for x in c["src_depths"]], c["par"],
                                  _dev(c["depth"]), _dev(c["normal"]), _dev(c["score"]), _dev(c["active"]), 0.02,
                                  c["dmin"], c["dmax"], 500 + j))
        cs.append(c)
    B = batch.Batch()
    B.set_pm(views, r, step, k, gw, iters)
    B.run()
    changed = B.pm_changed()
    for V, c, ch in zip(views, cs, changed):
        d0, n0, s0 = c["depth"].copy(), c["normal"].copy(), c["score"].copy()
        ch0 = oracle.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["src_depths"], c["par"], d0, n0, s0,
                               c["active"], r, step, k, gw, 0.02, iters, c["dmin"], c["dmax"], V.seed)
        assert ch == ch0, (c["ref"].shape, ch, ch0)
        assert _same(V.depth.cpu().numpy(), d0) and _same(V.normal.cpu().numpy(), n0), c["ref"].shape
        assert _same(V.score.cpu().numpy(), s0), c["ref"].shape
    B.close()
