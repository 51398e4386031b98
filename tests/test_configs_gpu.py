"""The batched, precision and served workloads at their full sizes.

- Batched block: 48 views of 800x600 .. 2400x1800 from a seeded 8 x 6 grid at 90 m (70 % / 60 %
  overlap), 4-12 sources each, 192 planes + parabola + 6 PatchMatch iterations in one batch:
  batched = the per-view calls bit for bit, a HIP-graph replay = the eager run, and on a centre
  crop of one view of every size class the batched sweep and PatchMatch against the oracle.
- Precision: a 4000x3000 reference with 12 sources, 384 planes + 8 PatchMatch iterations, images
  in FP16, BF16 and FP32 storage against the FP64 path (which is checked against the oracle on a
  crop), FP16 and BF16 also against FP32 storage.
- Served: requests of the served size classes (1200x900, 2400x1800, 4000x3000), 4-12 sources and
  128-384 planes, served = the single-view calls bit for bit and alone = in the mix.
"""
import dataclasses

import numpy as np
import pytest
import torch

import cases
import oracle
import sweep_bar

pytestmark = pytest.mark.gpu


def _bits(ts):
    return [t.detach().contiguous().view(torch.int32).cpu().numpy().copy() for t in ts]


def _np(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------------------ batched block
@pytest.fixture(scope="module")
def c3():
    from py_ortho_amd import block, scene
    sv = scene.config_c3(device="cuda")
    assert len(sv.views) == 48
    opt = block.BlockOptions(num_depths=192, pm_iters=6)
    blk = block.BlockDepth([v.img for v in sv.views], [v.cam for v in sv.views], sv.nbrs,
                           [(sv.dmin, sv.dmax)] * len(sv.views), opt)
    blk.plan()
    blk.run()
    torch.cuda.synchronize()
    yield sv, blk
    blk.close()


def test_c3_block_batched_equals_per_view_and_graph(c3):
    sv, blk = c3
    sizes = {tuple(v.img.shape) for v in sv.views}
    assert sizes == {(600, 800), (900, 1200), (1200, 1600), (1800, 2400)}
    assert {len(nb) for nb in sv.nbrs} <= set(range(4, 13))
    outs = (blk.raw, blk.depth, blk.normal, blk.score)
    batched = [_bits(o) for o in outs]
    print("c3 pixel-hypotheses (sweep, PatchMatch):", blk.pixel_hypotheses())
    blk.run_single()
    torch.cuda.synchronize()
    for name, a, o in zip(("raw", "depth", "normal", "score"), batched, outs):
        assert all(np.array_equal(x, y) for x, y in zip(a, _bits(o))), name
    blk.capture()
    for rep in range(2):
        for t in blk.depth + blk.raw:
            t.fill_(0.0)
        torch.cuda.synchronize()
        blk.replay()
        torch.cuda.synchronize()
        for name, a, o in zip(("raw", "depth", "normal", "score"), batched, outs):
            assert all(np.array_equal(x, y) for x, y in zip(a, _bits(o))), (name, rep)


def _crop_cam(cam, y0, x0, h, w):
    return dataclasses.replace(cam, cx=cam.cx - x0, cy=cam.cy - y0, w=w, h=h)


def test_c3_crops_batched_sweep_and_patchmatch_match_oracle(c3):
    """A 72x96 centre crop of one view per size class, as its own reference view against the
    block's full source images and pass-1 depth maps: the batched sweep (192 planes, 11x11 window)
    and parabola, then batched PatchMatch (6 iterations) from that pass-1 state, against the oracle."""
    from py_ortho_amd import batch, block, dense
    sv, blk = c3
    o = blk.opt
    picks, seen = [], set()
    for k, v in enumerate(sv.views):
        if tuple(v.img.shape) not in seen:
            seen.add(tuple(v.img.shape))
            picks.append(k)
    ch, cw = 72, 96
    sweeps, pms, cases_ = [], [], []
    for k in picks:
        v = sv.views[k]
        H, W = v.img.shape
        y0, x0 = (H - ch) // 2, (W - cw) // 2
        cam = _crop_cam(v.cam, y0, x0, ch, cw)
        img = v.img[y0:y0 + ch, x0:x0 + cw].contiguous()
        ref = torch.where(torch.isfinite(img), img, torch.zeros_like(img)).contiguous()
        mu, sd = dense.ref_stats(ref, o.window)
        ray = dense.rays_world(ch, cw, cam.R, cam.cx, cam.cy, cam.f, cam.k1, cam.k2, cam.k3)
        nb = sv.nbrs[k]
        srcs = [sv.views[j].img for j in nb]
        geo = np.stack([block._geo18(sv.views[j].cam, cam) for j in nb])
        par = np.stack([block._par18(sv.views[j].cam) for j in nb])
        planes = blk.planes[k]
        st = dense.SweepState(ch, cw)
        raw = torch.empty((ch, cw), dtype=torch.float32, device="cuda")
        sweeps.append(batch.SweepView(ref, mu, sd, srcs, ray, geo, {"mode": "const", "inv": planes}, state=st,
                                      depth=raw))
        dep = torch.empty_like(raw)
        nrm = torch.empty((3, ch, cw), dtype=torch.float32, device="cuda")
        sc = torch.empty_like(raw)
        act = torch.ones((ch, cw), dtype=torch.uint8, device="cuda")
        sdeps = [blk.raw[j] for j in nb]
        pms.append(batch.PMView(img, ray, np.asarray(cam.C, np.float64), srcs, sdeps, par, dep, nrm, sc, act,
                                o.geo_tol, sv.dmin, sv.dmax, blk.seed(k), depth_from=raw, score_from=st.best))
        cases_.append(dict(k=k, ref=ref, mu=mu, sd=sd, ray=ray, img=img, srcs=srcs, geo=geo, par=par, planes=planes,
                           sdeps=sdeps, C=np.asarray(cam.C, np.float64)))
    B = batch.Batch()
    B.set_sweep(sweeps, o.window, o.top_k)
    B.set_pm(pms, o.pm_window, o.pm_step, o.top_k, o.pm_geo_weight, o.pm_iters, init_normals=True)
    B.run()
    changed = B.pm_changed()
    torch.cuda.synchronize()
    for c, swv, pmv, ch_dev in zip(cases_, sweeps, pms, changed):
        hyps = [np.full((ch, cw), x, np.float32) for x in _np(c["planes"])]
        srcs = [_np(s) for s in c["srcs"]]
        st0, idx0, S = sweep_bar.oracle_scores(_np(c["ref"]), _np(c["mu"]), _np(c["sd"]), srcs, _np(c["ray"]),
                                                c["geo"], hyps, o.window, o.top_k)
        st = swv.state
        stats = sweep_bar.check(S, st0, idx0, [_np(t) for t in (st.best, st.prev, st.s_prev_best, st.s_next_best)],
                                _np(st.idx), ("c3 crop", c["k"]))
        print("c3 crop sweep", stats)
        d0 = oracle.sweep_finalize(hyps[0], hyps[1], len(hyps), st0[0], st0[2], st0[3], idx0)
        if not stats["bit_identical_state"]:
            pytest.fail(f"view {c['k']}: crop sweep passes the bar but is not bit-identical; PatchMatch start differs")
        assert np.array_equal(_np(swv.depth).view(np.uint32), d0.view(np.uint32))
        nrm0 = cases.np_init_normals(d0, _np(c["ray"]))
        dep, nrm, sco = d0.copy(), nrm0, st0[0].copy()
        ch0 = oracle.pm_refine(_np(c["img"]), _np(c["ray"]), c["C"], srcs, [_np(x) for x in c["sdeps"]], c["par"],
                               dep, nrm, sco, np.ones((ch, cw), np.uint8), o.pm_window, o.pm_step, o.top_k,
                               o.pm_geo_weight, o.geo_tol, o.pm_iters, sv.dmin, sv.dmax, blk.seed(c["k"]))
        assert ch_dev == ch0, (c["k"], ch_dev, ch0)
        for a, b in ((pmv.depth, dep), (pmv.normal, nrm), (pmv.score, sco)):
            assert np.array_equal(_np(a).view(np.uint32), b.view(np.uint32)), c["k"]
    B.close()


# ------------------------------------------------------------------------------ precision
def test_c4_fp32_fp16_bf16_against_fp64_full_size():
    """4000x3000, 12 sources, 384 planes (11x11 window, top-2) + parabola + 8 PatchMatch iterations.
    The FP64 path is checked against the oracle on the interior of a 160x160 centre crop; FP32,
    FP16 and BF16 image storage are checked against the FP64 path: same winning plane on >= 98 %
    (FP16), 99.5 % (FP32) / 85 % (BF16) of the valid pixels, |best score - FP64| at the 99th
    percentile <= 1e-3 / 0.1 / 0.5 (FP32 / FP16 / BF16), and PatchMatch depth within 0.2 % of FP64 on >= 99 % of the pixels valid in
    both.  FP16 and BF16 are also reported against FP32 storage (their storage error alone)."""
    from py_ortho_amd import batch, block, dense, scene
    sc = scene.config_c4(device="cuda")
    ref = sc.ref.img.contiguous()
    H, W = ref.shape
    assert (H, W) == (3000, 4000) and len(sc.srcs) == 12
    r, k, D, iters, pr, pstep = 5, 2, 384, 8, 5, 2
    dmin, dmax = 73.0, 92.0
    ray = scene.rays_world(sc.ref.cam, "cuda")
    mu, sd = dense.ref_stats(ref, r)
    geo = sc.geo()
    inv = torch.from_numpy(block.inverse_depth_planes(dmin, dmax, D)).cuda()
    hyp = {"mode": "const", "inv": inv}
    srcs = [v.img.contiguous() for v in sc.srcs]
    deps = [v.depth.contiguous() for v in sc.srcs]
    active = torch.ones((H, W), dtype=torch.uint8, device="cuda")

    def run(fmt):
        st = dense.SweepState(H, W)
        if fmt:
            dense.sweep_lp(dense.to_lp(ref, fmt), mu, sd, [dense.to_lp(x, fmt) for x in srcs], ray, geo, hyp, r, k,
                           fmt, state=st)
        else:
            dense.sweep(ref, mu, sd, srcs, ray, geo, hyp, r, k, state=st)
        d = dense.sweep_finalize(ref, mu, sd, srcs, ray, geo, hyp, r, k, st)
        pm = d.clone()
        nrm = batch.init_normals(d, ray)
        score = st.best.clone()
        lp_srcs = [dense.to_lp(x, fmt) for x in srcs] if fmt else srcs
        dense.pm_refine(ref, ray, sc.ref.cam.C, lp_srcs, deps, sc.par(), pm, nrm, score, active, pr, pstep, k, 0.5,
                        0.01, iters, dmin, dmax, 2654435761, fmt=fmt)
        torch.cuda.synchronize()
        return st, d, pm

    st64, d64, pm64 = run(None)
    # FP64 against the oracle on a crop interior (the oracle's windows see zero padding at the crop edge)
    cs = 160
    y0, x0 = (H - cs) // 2, (W - cs) // 2
    sl = np.s_[y0:y0 + cs, x0:x0 + cs]
    hyps = [np.full((cs, cs), x, np.float32) for x in _np(inv)]
    c = lambda t: np.ascontiguousarray(_np(t)[sl])  # noqa: E731
    st0, idx0, S = sweep_bar.oracle_scores(c(ref), c(mu), c(sd), [_np(x) for x in srcs],
                                            np.ascontiguousarray(_np(ray)[:, y0:y0 + cs, x0:x0 + cs]), geo, hyps, r, k)
    inner = np.s_[r:-r, r:-r]
    dev = [np.ascontiguousarray(c(t)[inner]) for t in (st64.best, st64.prev, st64.s_prev_best, st64.s_next_best)]
    stats = sweep_bar.check(np.ascontiguousarray(S[(slice(None),) + inner]),
                            [np.ascontiguousarray(a[inner]) for a in st0], np.ascontiguousarray(idx0[inner]), dev,
                            np.ascontiguousarray(c(st64.idx)[inner]), "c4 fp64 crop")
    print("c4 fp64 vs oracle", stats)
    valid = (st64.best > 0.3) & (64.0 * sd > 0.5)

    def errors(a, b):
        (st, d, pm), (st_b, d_b, pm_b) = a, b
        same = float((st.idx == st_b.idx)[valid].float().mean())
        db = (st.best - st_b.best).abs()[valid].float()
        db99 = float(torch.quantile(db[:: max(1, db.numel() // 1000000)], 0.99))
        rel_sw = ((d - d_b).abs() / d_b)[valid].float()
        both = torch.isfinite(pm) & torch.isfinite(pm_b)
        within = float((((pm - pm_b).abs() / pm_b)[both] < 2e-3).float().mean())
        return same, db99, float(torch.quantile(rel_sw[:: max(1, rel_sw.numel() // 1000000)], 0.99)), within

    run64 = (st64, d64, pm64)
    run32 = run("f32")  # the variants' FP32 arithmetic with the images as they are
    for fmt, win_min, db_max in (("f32", 0.995, 1e-3), ("f16", 0.98, 0.1), ("bf16", 0.85, 0.5)):
        out = run32 if fmt == "f32" else run(fmt)
        same, db99, rel99, within = errors(out, run64)
        print(f"c4 {fmt} against FP64: same winner {same:.4f}, |d best| p99 {db99:.4g}, sweep depth rel p99 "
              f"{rel99:.3g}, PatchMatch within 0.2 % {within:.5f}")
        if fmt != "f32":
            print(f"c4 {fmt} against FP32 storage: same winner %.4f, |d best| p99 %.4g, sweep depth rel p99 %.3g, "
                  "PatchMatch within 0.2 %% %.5f" % errors(out, run32))
        assert same >= win_min and db99 <= db_max and within >= 0.99, fmt


# ------------------------------------------------------------------------------ served
def test_c5_served_full_sizes_equal_single_view_and_alone():
    import test_serve_gpu as tsg
    from py_ortho_amd import scene
    from py_ortho_amd.serve import DepthRequest
    pools = scene.config_c5_pools(device="cuda", nx=4, ny=4)
    images = {}
    for cl, pool in enumerate(pools):
        for j, v in enumerate(pool.views):
            images[(cl, j)] = (v.img.cpu().numpy(), v.cam)
    rng = np.random.default_rng(6)
    reqs = []
    for i in range(12):
        cl = (0, 0, 1, 0, 2, 1, 0, 0, 1, 0, 2, 0)[i]            # 60 / 30 / 10 % of the sizes, each class present
        pool = pools[cl]
        k = int(rng.integers(len(pool.views)))
        S = int(rng.integers(4, 13))
        reqs.append(DepthRequest((cl, k), [(cl, j) for j in pool.nbrs[k][:S]], pool.dmin, pool.dmax,
                                 int(rng.integers(128, 385)), tag=i))
    assert {images[q.ref][0].shape for q in reqs} == {(900, 1200), (1800, 2400), (3000, 4000)}
    res, srv, store = tsg._serve(images, reqs, 5, 2, step_planes=32, budget=200e9)
    assert srv.stats["finished"] == len(reqs) and srv.stats["max_active"] > 1
    for r, q in zip(res, reqs):
        d, s = tsg._plain(images, q, 5, 2)
        assert tsg._same(r.depth, d), q.tag
        assert tsg._same(r.score, s), q.tag
    for i in (4, 7):
        alone, _, _ = tsg._serve(images, [reqs[i]], 5, 2, step_planes=32, budget=200e9)
        assert tsg._same(alone[0].depth, res[i].depth) and tsg._same(alone[0].score, res[i].score)
