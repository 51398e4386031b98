"""om_sweep_noalloc: the plane sweep with its chunk partials in the library's static scratch array.

Checked against om_sweep with a caller workspace (the same kernels), bit for bit, through the C ABI,
on two calls in a row, the second on a smaller view (the scratch keeps the first call's contents).
"""
import ctypes
import os

import numpy as np
import pytest
import torch

import cases
from py_ortho_amd import _lib, dense


def _desc(case, r, k, state, chunks):
    ref, mu, sd, srcs, ray, geo, inv = case
    d, keep = dense._sweep_desc(ref, mu, sd, srcs, ray, geo, r, k, state, _lib.OM_HYP_CONST, inv, None, 0, 0,
                                int(inv.shape[0]), chunks, 0)
    return d, keep


def _case(seed, H, W, n, D, dev):
    rng = np.random.default_rng(seed)
    ref, mu, sd, srcs, ray, geo, _ = cases.sweep_case(rng, H, W, n, 2, 1)
    inv = np.linspace(0.02, 0.2, D).astype(np.float32)    # the inverse depths of sweep_case
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    return t(ref), t(mu), t(sd), [t(s) for s in srcs], t(ray), geo, t(inv)


def _run(case, r, k, chunks, noalloc):
    H, W = case[0].shape
    st = dense.SweepState(H, W, case[0].device)
    d, keep = _desc(case, r, k, st, chunks)
    L = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        _lib.check(L.om_sweep_noalloc(ctypes.byref(d), stream), "sweep_noalloc")
    else:
        nb = L.om_sweep_workspace_bytes(ctypes.byref(d))
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=case[0].device)
        d.workspace, d.workspace_bytes = ws.data_ptr(), nb
        _lib.check(L.om_sweep(ctypes.byref(d), stream), "sweep")
    torch.cuda.synchronize()
    del keep
    return [x.cpu().numpy() for x in st.tensors()]


def _bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


@pytest.mark.gpu
@pytest.mark.parametrize("chunks", [0, 4])
def test_noalloc_sweep_equals_sweep_twice(chunks):
    dev = torch.device("cuda:0")
    big = _case(11, 180, 250, 4, 40, dev)     # 4 x 12 tiles: chunked by default, partials in the scratch
    small = _case(12, 70, 90, 3, 24, dev)     # second call: a smaller view over the same scratch
    for case in (big, small):
        want = _run(case, 2, 2, chunks, noalloc=False)
        got = _run(case, 2, 2, chunks, noalloc=True)
        for a, b in zip(want, got):
            assert np.array_equal(_bits(a), _bits(b))


def test_noalloc_rejects_partials_beyond_scratch():
    """A chunk count whose partials exceed the scratch is refused before any device work."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libom355.so not built")
    L = _lib.lib()
    cap = L.om_sweep_noalloc_max_bytes()
    assert cap == 64 << 20
    H, W, chunks = 1500, 2000, 8           # 8 x 6 x 4 B x 3 MP = 576 MB of partials
    d = _lib.OmSweepDesc()
    srcs = (_lib.OmSrc * 1)()
    srcs[0].img, srcs[0].h, srcs[0].w = 1, H, W
    d.ref = d.mu = d.sd = d.ray = 1
    d.H, d.W, d.n_src, d.src = H, W, 1, ctypes.cast(srcs, ctypes.POINTER(_lib.OmSrc))
    d.r, d.top_k, d.hyp_mode, d.hyp_a, d.nh, d.chunks = 2, 1, _lib.OM_HYP_CONST, 1, 64, chunks
    assert L.om_sweep_noalloc(ctypes.byref(d), None) == _lib.OM_ESHAPE
    assert b"scratch" in L.om_last_error()
