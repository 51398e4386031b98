"""Argument checks of the C ABI, run without a GPU: each case is rejected (or returns without
work) before the library makes any HIP call, so no device memory is touched -- the pointers below
are never dereferenced.

The reference's own contract (_dense.pyx): pm_refine returns 0 without work when the source count
is outside 1..16 (:587-588); sweep_hypothesis raises ValueError for a source count outside 1..16
or a negative radius (:759-771) and returns early for an empty image (:772-773).  The library adds
the limits of its own kernels: images of fewer than 2^30 pixels with sides below 2^24, window radius <= 7 and
min(top_k, n) <= 4 in the fused sweep (OM_EUNSUP), tile heights of 8 or 16 rows.
"""
import ctypes
import os

import pytest

from py_ortho_amd import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libom355.so not built")

FAKE = 0x1000  # a non-null address the checks see but never read


def _pm(H=4, W=4, S=1, step=1, top_k=1, ih=4, iw=4, workspace=FAKE, nbytes=None, iters=1, null_sizes=False):
    L = _lib.lib()
    if nbytes is None:
        nbytes = L.om_pm_workspace_bytes(max(H, 1), max(W, 1))
    n = max(S, 1)
    imgs = (ctypes.c_void_p * n)(*([FAKE] * n))
    deps = (ctypes.c_void_p * n)(*([None] * n))
    ihs = (ctypes.c_int32 * n)(*([ih] * n))
    iws = (ctypes.c_int32 * n)(*([iw] * n))
    par = (ctypes.c_double * (18 * n))(*([0.0] * (18 * n)))
    cr = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    changed = ctypes.c_int64(-7)
    rc = L.om_pm_refine(FAKE, H, W, FAKE, cr, S, imgs, deps, None if null_sizes else ihs, iws, par, FAKE, FAKE,
                        FAKE, FAKE, 2, step, top_k, 0.5, 0.01, iters, 1.0, 100.0, 5, workspace, nbytes,
                        ctypes.byref(changed), None)
    return rc, changed.value


@pytest.mark.parametrize("S", [0, 17])
def test_pm_refine_source_count_outside_1_16_returns_zero_changes(S):
    assert _pm(S=S) == (_lib.OM_OK, 0)


def test_pm_refine_empty_image_or_no_iterations_is_a_no_op():
    assert _pm(H=0, W=5) == (_lib.OM_OK, 0)
    assert _pm(iters=0) == (_lib.OM_OK, 0)


@pytest.mark.parametrize("kw,code,text", [
    (dict(H=-1), _lib.OM_ESHAPE, "negative"),
    (dict(step=0), _lib.OM_EINVAL, "step"),
    (dict(top_k=0), _lib.OM_EINVAL, "top_k"),
    (dict(null_sizes=True), _lib.OM_EINVAL, "null"),
    (dict(H=1 << 15, W=1 << 15), _lib.OM_ESHAPE, "2^30"),
    (dict(workspace=None), _lib.OM_EINVAL, "workspace"),
    (dict(nbytes=255), _lib.OM_EINVAL, "workspace"),
    (dict(ih=0), _lib.OM_ESHAPE, "empty"),
    (dict(ih=1 << 15, iw=1 << 15), _lib.OM_ESHAPE, "2^30"),
    (dict(H=1, W=1 << 24), _lib.OM_ESHAPE, "2^24"),          # a side past the 24-bit row multiply
    (dict(ih=1, iw=1 << 24), _lib.OM_ESHAPE, "2^24"),
])
def test_pm_refine_rejects_bad_arguments(kw, code, text):
    rc, changed = _pm(**kw)
    assert rc == code and changed == 0
    assert text in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, "om_pm_refine")


def _sweep(H=4, W=4, n=2, r=1, top_k=1, mode=0, nh=1, src_hw=(4, 4), tile_h=0, null_src=False):
    src = (_lib.OmSrc * n)()
    for s in src:
        s.img, s.h, s.w = FAKE, src_hw[0], src_hw[1]
    d = _lib.OmSweepDesc(ref=FAKE, mu=FAKE, sd=FAKE, ray=FAKE, H=H, W=W, n_src=n,
                         src=None if null_src else ctypes.cast(src, ctypes.POINTER(_lib.OmSrc)), r=r, top_k=top_k,
                         hyp_mode=mode, hyp_a=FAKE, hyp_b=FAKE, hyp_offset=0, i0=0, nh=nh, best=FAKE, prev=FAKE,
                         s_prev_best=FAKE, s_next_best=FAKE, idx=FAKE, workspace=None, workspace_bytes=0, chunks=1,
                         tile_h=tile_h)
    return _lib.lib().om_sweep(ctypes.byref(d), None)


def test_sweep_empty_image_or_no_hypotheses_is_a_no_op():
    assert _sweep(H=0) == _lib.OM_OK
    assert _sweep(W=0) == _lib.OM_OK
    assert _sweep(nh=0) == _lib.OM_OK


@pytest.mark.parametrize("kw,code", [
    (dict(n=0), _lib.OM_ENSRC),
    (dict(n=17), _lib.OM_ENSRC),
    (dict(r=-1), _lib.OM_EINVAL),
    (dict(H=-1), _lib.OM_ESHAPE),
    (dict(H=1 << 15, W=1 << 15), _lib.OM_ESHAPE),
    (dict(src_hw=(1 << 15, 1 << 15)), _lib.OM_ESHAPE),
    (dict(H=1 << 24, W=1), _lib.OM_ESHAPE),
    (dict(src_hw=(1 << 24, 1)), _lib.OM_ESHAPE),
    (dict(top_k=0), _lib.OM_EINVAL),
    (dict(mode=99), _lib.OM_EINVAL),
    (dict(null_src=True), _lib.OM_EINVAL),
    (dict(tile_h=12), _lib.OM_EINVAL),
    (dict(r=8), _lib.OM_EUNSUP),
    (dict(n=5, top_k=5), _lib.OM_EUNSUP),
])
def test_sweep_rejects_bad_arguments(kw, code):
    rc = _sweep(**kw)
    assert rc == code
    assert _lib.last_error()
    with pytest.raises(NotImplementedError if code == _lib.OM_EUNSUP else ValueError):
        _lib.check(rc, "om_sweep")
