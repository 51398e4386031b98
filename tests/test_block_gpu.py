"""A block of views of different sizes and source counts in one batch (sweep, parabola, PatchMatch
with the neighbours' pass-1 maps) equals the per-view calls bit for bit, with the autotuner's
per-size-class tile shapes, eagerly and as a replayed HIP graph."""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _block(scale=0.1, planes=24, iters=2, **kw):
    from py_ortho_amd import block, scene
    sizes = [(int(round(w * scale)), int(round(h * scale))) for w, h in scene.C3_SIZES]
    cams, counts, ext = scene.survey_cameras(4, 4, 3, 90.0, sizes, 0.9375, 0.7, 0.6, (2, 6))
    sv = scene.make_survey(4, cams, counts, ext, 90.0, device="cuda")
    opt = block.BlockOptions(num_depths=planes, pm_iters=iters, **kw)
    blk = block.BlockDepth([v.img for v in sv.views], [v.cam for v in sv.views], sv.nbrs,
                           [(sv.dmin, sv.dmax)] * len(sv.views), opt)
    return sv, blk


def _bits(ts):
    return [t.detach().view(torch.int32).cpu().numpy().copy() for t in ts]


def test_block_batch_equals_single_view_and_graph(tmp_path):
    from py_ortho_amd import autotune
    sv, blk = _block()
    assert len({blk.size_class(k) for k in range(blk.n)}) >= 3
    cache = str(tmp_path / "tune.json")
    tune, table = autotune.tune_block(blk, planes=8, views_per_class=2, reps=1, cache_path=cache)
    assert set(tune) == {blk.size_class(k) for k in range(blk.n)}
    assert all(v in autotune.CANDIDATES for v in tune.values())
    # a second call reads the choices from the cache file without timing anything
    tune2, table2 = autotune.tune_block(blk, planes=8, views_per_class=2, reps=1, cache_path=cache)
    assert tune2 == tune and table2 == {} and json.load(open(cache))
    # mixed tile heights: force both groups to be present
    classes = sorted(tune)
    tune = {c: ((16, 1) if i % 2 else (8, 2)) for i, c in enumerate(classes)}
    blk.plan(tune)
    blk.run()
    torch.cuda.synchronize()
    raw_b, dep_b, nrm_b, sc_b = (_bits(blk.raw), _bits(blk.depth), _bits(blk.normal), _bits(blk.score))
    blk.run_single()
    torch.cuda.synchronize()
    for a, b in zip(raw_b, _bits(blk.raw)):
        assert np.array_equal(a, b)
    for a, b in zip(dep_b, _bits(blk.depth)):
        assert np.array_equal(a, b)
    for a, b in zip(nrm_b, _bits(blk.normal)):
        assert np.array_equal(a, b)
    for a, b in zip(sc_b, _bits(blk.score)):
        assert np.array_equal(a, b)
    # graph replays reproduce the eager batch (state is reset inside the graph)
    blk.capture()
    for rep in range(2):
        for t in blk.depth + blk.raw:
            t.fill_(0.0)
        torch.cuda.synchronize()
        blk.replay()
        torch.cuda.synchronize()
        assert all(np.array_equal(a, b) for a, b in zip(raw_b, _bits(blk.raw))), rep
        assert all(np.array_equal(a, b) for a, b in zip(dep_b, _bits(blk.depth))), rep
    blk.close()


def test_block_depth_accuracy():
    """The batched pass-1 depth and the PatchMatch depth land on the rendered terrain."""
    sv, blk = _block(scale=0.15, planes=48, iters=2)
    blk.run()
    torch.cuda.synchronize()
    for maps in (blk.raw, blk.depth):
        rel = torch.cat([((d - v.depth).abs() / v.depth).flatten() for d, v in zip(maps, sv.views)])
        rel = rel[torch.isfinite(rel)]
        assert float(rel.median()) < 0.01
    blk.close()


def test_block_reference_repair_batch_equals_single_view():
    """repair="inconsistent": PatchMatch over the reference's pass-2 inputs (densify.py:611-634)
    built from the sweep's maps.  The batched run equals the per-view calls bit for bit; the inputs
    are the reference's expressions over the consistency count (om_consistency, pinned to the
    reference's _consistency by golden vectors) and the float16-rounded score."""
    from py_ortho_amd import formats, fusion
    sv, blk = _block(repair="inconsistent", min_ncc=0.5)
    blk.plan()
    blk.run()
    torch.cuda.synchronize()
    raw_b, dep_b, nrm_b, sc_b, act_b = (_bits(blk.raw), _bits(blk.depth), _bits(blk.normal), _bits(blk.score),
                                        [a.cpu().numpy().copy() for a in blk.active])
    inputs = blk.repair_inputs()
    mv = blk.min_views()
    n_act = n_px = 0
    for k, (active, geo_tol, S) in enumerate(inputs):
        # the inputs, recomputed here from their parts
        D = blk.raw[k]
        fin = D[torch.isfinite(D)]
        tol = blk.rel_tols(k, float(fin.median()))
        assert set(tol) == set(blk.nbrs[k]) and all(0.01 <= t <= 0.05 for t in tol.values())
        assert geo_tol == float(np.median(list(tol.values())))
        cnt = fusion.consistency(k, D, blk.nbrs[k], {j: blk.raw[j] for j in range(blk.n)}, blk.cams, blk.opt.px_tol,
                                 tol).cpu().numpy()
        s16 = formats.to_f16(blk.states[k].best).cpu().numpy().astype(np.float32)
        d, img = D.cpu().numpy(), blk.imgs[k].cpu().numpy()
        want = np.isfinite(img) & ((cnt < mv - 1) | (s16 < blk.opt.min_ncc) | ~np.isfinite(d))
        assert np.array_equal(active.cpu().numpy().astype(bool), want), k
        assert np.array_equal(act_b[k].astype(bool), want), k
        n_act += int(want.sum())
        n_px += want.size
    assert 0 < n_act < n_px, (n_act, n_px)       # some pixels repaired, others kept
    blk.run_single()
    torch.cuda.synchronize()
    for got, want in ((raw_b, blk.raw), (dep_b, blk.depth), (nrm_b, blk.normal), (sc_b, blk.score)):
        assert all(np.array_equal(a, b) for a, b in zip(got, _bits(want)))
    with pytest.raises(ValueError):
        blk.capture()
    blk.close()
