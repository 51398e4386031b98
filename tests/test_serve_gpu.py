"""The served path on the device: om_sweep_multi steps over requests of different sizes, source
counts and plane counts equal the single-view sweep + parabola of each request bit for bit, and a
request's result is the same alone and in the mix."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pool():
    from py_ortho_amd import scene
    pools = scene.config_c5_pools(device="cuda", scale=0.08, nx=4, ny=3)   # 96x72, 192x144, 320x240
    images = {}
    for c, sv in enumerate(pools):
        for k, v in enumerate(sv.views):
            images[(c, k)] = (v.img.cpu().numpy(), v.cam)
    return pools, images


def _requests(pools, n, seed):
    from py_ortho_amd.serve import DepthRequest
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = int(rng.integers(len(pools)))
        sv = pools[c]
        k = int(rng.integers(len(sv.views)))
        S = int(rng.integers(2, 7))
        out.append(DepthRequest((c, k), [(c, j) for j in sv.nbrs[k][:S]], sv.dmin, sv.dmax,
                                int(rng.integers(20, 61)), tag=i))
    return out


def _plain(images, req, window, top_k):
    from py_ortho_amd import block, dense
    img, cam = images[req.ref]
    ref = torch.from_numpy(img).cuda()
    ref = torch.where(torch.isfinite(ref), ref, torch.zeros_like(ref)).contiguous()
    H, W = ref.shape
    mu, sd = dense.ref_stats(ref, window)
    ray = dense.rays_world(H, W, cam.R, cam.cx, cam.cy, cam.f, cam.k1, cam.k2, cam.k3)
    srcs = [torch.from_numpy(images[k][0]).cuda() for k in req.srcs]
    geo = np.stack([block._geo18(images[k][1], cam) for k in req.srcs])
    hyp = {"mode": "const", "inv": torch.from_numpy(block.inverse_depth_planes(req.dmin, req.dmax, req.planes)).cuda()}
    st = dense.sweep(ref, mu, sd, srcs, ray, geo, hyp, window, top_k)
    d = dense.sweep_finalize(ref, mu, sd, srcs, ray, geo, hyp, window, top_k, st)
    return d.cpu().numpy(), st.best.cpu().numpy()


def _serve(images, reqs, window, top_k, step_planes, budget=4e9):
    from py_ortho_amd import serve
    ex = serve.GpuExecutor(window=window, top_k=top_k)
    store = serve.ImageStore(lambda key: images[key], 2e9, upload=ex.upload)
    srv = serve.DepthServer(store, ex, budget_bytes=budget, step_planes=step_planes).start()
    futs = [srv.submit(r) for r in reqs]
    res = [f.result(timeout=120) for f in futs]
    srv.stop()
    return res, srv, store


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_served_equals_single_view_and_alone():
    pools, images = _pool()
    reqs = _requests(pools, 9, seed=3)
    res, srv, store = _serve(images, reqs, 3, 2, step_planes=16)
    assert srv.stats["finished"] == len(reqs) and srv.stats["max_active"] > 1
    assert store.uploads <= len(images)                       # each image uploaded once while resident
    for r, q in zip(res, reqs):
        d, s = _plain(images, q, 3, 2)
        assert _same(r.depth, d), q.tag
        assert _same(r.score, s), q.tag
        assert r.latency > 0
    alone, _, _ = _serve(images, [reqs[4]], 3, 2, step_planes=7)
    assert _same(alone[0].depth, res[4].depth) and _same(alone[0].score, res[4].score)


def test_sweep_multi_rejects_mixed_windows():
    from py_ortho_amd import _lib, block, dense
    pools, images = _pool()
    reqs = _requests(pools, 2, seed=5)
    descs = (_lib.OmSweepDesc * 2)()
    keep = []
    for j, (q, r) in enumerate(zip(reqs, (2, 3))):
        img, cam = images[q.ref]
        ref = torch.from_numpy(np.nan_to_num(img)).cuda()
        H, W = ref.shape
        mu, sd = dense.ref_stats(ref, r)
        ray = dense.rays_world(H, W, cam.R, cam.cx, cam.cy, cam.f)
        srcs = [torch.from_numpy(images[k][0]).cuda() for k in q.srcs]
        geo = np.stack([block._geo18(images[k][1], cam) for k in q.srcs])
        inv = torch.from_numpy(block.inverse_depth_planes(q.dmin, q.dmax, q.planes)).cuda()
        st = dense.SweepState(H, W)
        d, k = dense._sweep_desc(ref, mu, sd, srcs, ray, geo, r, 2, st, _lib.OM_HYP_CONST, inv, None, 0, 0, q.planes)
        descs[j] = d
        keep += [k, ref, mu, sd, ray, srcs, inv, st]
    L = _lib.lib()
    nb = L.om_sweep_multi_bytes(2)
    table = torch.empty(nb, dtype=torch.uint8, device="cuda")
    staging = torch.empty(nb, dtype=torch.uint8, pin_memory=True)
    rc = L.om_sweep_multi(descs, 2, None, 0, table.data_ptr(), nb, staging.data_ptr(), None)
    assert rc == _lib.OM_EINVAL and "window radius" in _lib.last_error()
