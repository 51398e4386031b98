"""CPU oracle for the depth-map path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this package, as the
checker and as the same-run CPU baseline.  The product (py_ortho_amd, libom355.so) never imports it.

- liborc.so (om_oracle.c): C restatement of _dense.sweep_hypothesis / ncc_warp / combine /
  pm_refine and the _fast depth helpers, bit-for-bit in the reference's arithmetic.
- _ref/libref.so (ref_sweep_driver.c + the reference's own _sweep.h): the reference's compiled
  sweep row kernels, used to pin the restatement.  Built only where the reference tree exists.
- NumPy restatements of the Python glue: the parabola of densify._sweep (densify.py:354-361).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lock = threading.Lock()
_orc = None
_ref = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vpp = ctypes.POINTER(ctypes.c_void_p)


def _ptr(a: np.ndarray, kind):
    return a.ctypes.data_as(kind)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def orc():
    """liborc.so (built on demand with gcc)."""
    global _orc
    if _orc is None:
        with _lock:
            if _orc is None:
                path = os.path.join(HERE, "liborc.so")
                if not os.path.exists(path):
                    build()
                L = ctypes.CDLL(path)
                L.orc_sweep_hypothesis.restype = ctypes.c_int
                L.orc_ncc_warp.restype = ctypes.c_int
                L.orc_combine.restype = ctypes.c_int
                L.orc_box_zero_mean.restype = ctypes.c_int
                L.orc_pm_refine.restype = ctypes.c_int64
                L.orc_bench_sweep.restype = ctypes.c_int
                L.orc_bench_pm.restype = ctypes.c_int
                _orc = L
    return _orc


def ref_built() -> bool:
    """Whether oracle/_ref/libref.so exists (checked without loading it)."""
    return os.path.exists(os.path.join(HERE, "_ref", "libref.so"))


def ref_lib():
    """oracle/_ref/libref.so (the reference's _sweep.h), or None where it is not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_sweep_hypothesis.restype = ctypes.c_int
        _ref = L
    return _ref


def _c32(a):
    return np.ascontiguousarray(a, np.float32)


def _src_args(srcs):
    srcs = [_c32(s) for s in srcs]
    ptrs = (ctypes.c_void_p * len(srcs))(*[s.ctypes.data for s in srcs])
    hw = np.ascontiguousarray([v for s in srcs for v in s.shape], np.int64)
    return srcs, ptrs, hw


def _sweep_call(fn, ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, best, prev, spb, snb, idx):
    ref, mu, sd, ray, inv = (_c32(a) for a in (ref, mu, sd, ray, inv))
    geo = np.ascontiguousarray(geo, np.float64)
    keep, ptrs, hw = _src_args(srcs)
    H, W = ref.shape
    for a in (best, prev, spb, snb):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    assert idx.dtype == np.int32 and idx.flags.c_contiguous
    rc = fn(_ptr(ref, _f32p), _ptr(mu, _f32p), _ptr(sd, _f32p), ctypes.c_int64(H), ctypes.c_int64(W),
            ctypes.c_int(len(keep)), ptrs, _ptr(hw, _i64p), _ptr(ray, _f32p), _ptr(geo, _f64p), _ptr(inv, _f32p),
            ctypes.c_int(r), ctypes.c_int(top_k), ctypes.c_int(i), _ptr(best, _f32p), _ptr(prev, _f32p),
            _ptr(spb, _f32p), _ptr(snb, _f32p), _ptr(idx, _i32p))
    if rc == -1:
        raise ValueError("sweep_hypothesis: invalid arguments")
    if rc:
        raise MemoryError()


def sweep_hypothesis(ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, best, prev, spb, snb, idx):
    """Restatement of _dense.sweep_hypothesis (_dense.pyx:740-838); state arrays updated in place."""
    _sweep_call(orc().orc_sweep_hypothesis, ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, best, prev, spb, snb, idx)


def ref_sweep_hypothesis(ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, best, prev, spb, snb, idx):
    """The reference's own _sweep.h row kernels (oracle/_ref); raises if not built."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError("oracle/_ref/libref.so is not built (reference tree absent)")
    _sweep_call(L.ref_sweep_hypothesis, ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, best, prev, spb, snb, idx)


def ncc_warp(ref, mu, sd, src, ray, g18, inv, r):
    """Restatement of _dense.ncc_warp (_dense.pyx:13-110) -> (out, J, V, hs)."""
    ref, mu, sd, src, ray, inv = (_c32(a) for a in (ref, mu, sd, src, ray, inv))
    g18 = np.ascontiguousarray(g18, np.float64)
    H, W = ref.shape
    out = np.empty((H, W), np.float32)
    J = np.empty((H, W), np.float32)
    V = np.empty((H, W), np.float32)
    hs = np.empty((4, H, W), np.float32)
    rc = orc().orc_ncc_warp(_ptr(ref, _f32p), _ptr(mu, _f32p), _ptr(sd, _f32p), ctypes.c_int64(H), ctypes.c_int64(W),
                            _ptr(src, _f32p), ctypes.c_int64(src.shape[0]), ctypes.c_int64(src.shape[1]),
                            _ptr(ray, _f32p), _ptr(g18, _f64p), _ptr(inv, _f32p), ctypes.c_int(r), _ptr(out, _f32p),
                            _ptr(J, _f32p), _ptr(V, _f32p), _ptr(hs, _f32p))
    if rc:
        raise ValueError("ncc_warp failed")
    return out, J, V, hs


def combine(S, top_k, i, best, prev, spb, snb, idx):
    """Restatement of _dense.combine (_dense.pyx:113-144)."""
    S = _c32(S)
    n, H, W = S.shape
    rc = orc().orc_combine(_ptr(S, _f32p), ctypes.c_int(n), ctypes.c_int64(H), ctypes.c_int64(W), ctypes.c_int(top_k),
                           ctypes.c_int(i), _ptr(best, _f32p), _ptr(prev, _f32p), _ptr(spb, _f32p), _ptr(snb, _f32p),
                           _ptr(idx, _i32p))
    if rc:
        raise ValueError("combine failed")


def new_state(H, W):
    """Initial winner state of densify._sweep (densify.py:331-335)."""
    return ([np.full((H, W), -2.0, np.float32) for _ in range(4)], np.full((H, W), -1, np.int32))


def sweep_all(ref, mu, sd, srcs, ray, geo, hyps, r, top_k, fn=None):
    """Runs every hypothesis map of `hyps` in order -> ((best, prev, spb, snb), idx)."""
    fn = fn or sweep_hypothesis
    H, W = ref.shape
    st, idx = new_state(H, W)
    for i, inv in enumerate(hyps):
        fn(ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, st[0], st[1], st[2], st[3], idx)
    return st, idx


def sweep_finalize(inv0, inv1, n, best, spb, snb, idx):
    """The parabola of densify._sweep, densify.py:354-361, in NumPy's own arithmetic."""
    inv0 = np.asarray(inv0, np.float32)
    dinv = (np.asarray(inv1, np.float32) - inv0) if n > 1 else np.zeros_like(inv0)
    den = spb - 2 * best + snb
    interior = (idx > 0) & (idx < n - 1) & (den < -1e-6)
    off = np.where(interior, 0.5 * (spb - snb) / np.where(interior, den, -1.0), 0.0)
    inv = inv0 + (idx + np.clip(off, -0.5, 0.5)) * dinv
    return (1.0 / np.maximum(inv, 1e-9)).astype(np.float32)


def hyp_linear(base, step, i):
    base, step = _c32(base), _c32(step)
    out = np.empty_like(base)
    orc().orc_hyp_linear(_ptr(base, _f32p), _ptr(step, _f32p), ctypes.c_int64(base.size), ctypes.c_int64(i),
                         _ptr(out, _f32p))
    return out


def hyp_refine(iu, ds, t):
    iu, ds = _c32(iu), _c32(ds)
    out = np.empty_like(iu)
    orc().orc_hyp_refine(_ptr(iu, _f32p), _ptr(ds, _f32p), ctypes.c_int64(iu.size), ctypes.c_int64(t),
                         _ptr(out, _f32p))
    return out


def box_zero_mean(a, r):
    a = _c32(a)
    H, W = a.shape
    out = np.empty((H, W), np.float32)
    if orc().orc_box_zero_mean(_ptr(a, _f32p), ctypes.c_int64(H), ctypes.c_int64(W), ctypes.c_int(r),
                               _ptr(out, _f32p)):
        raise ValueError("box_zero_mean")
    return out


def ref_stats(ref, r):
    """mu, sd of densify._sweep (densify.py:327-328)."""
    ref = _c32(ref)
    mu = box_zero_mean(ref, r)
    sd = np.sqrt(np.maximum(box_zero_mean(ref * ref, r) - mu * mu, 1e-6)).astype(np.float32)
    return mu, sd


def half(g):
    g = _c32(g)
    H, W = g.shape
    out = np.empty((H // 2, W // 2), np.float32)
    orc().orc_half(_ptr(g, _f32p), ctypes.c_int64(H), ctypes.c_int64(W), _ptr(out, _f32p))
    return out


def rays_grid(H, W, cx, cy, f, k1, k2, k3):
    x = np.empty((H, W), np.float64)
    y = np.empty((H, W), np.float64)
    orc().orc_rays_grid(ctypes.c_int64(H), ctypes.c_int64(W), *(ctypes.c_double(v) for v in (cx, cy, f, k1, k2, k3)),
                        _ptr(x, _f64p), _ptr(y, _f64p))
    return x, y


def rays_world(H, W, R, cx, cy, f, k1=0.0, k2=0.0, k3=0.0):
    """densify._rays_world (densify.py:288-297): (3, H, W) float32 world rays."""
    x, y = rays_grid(H, W, cx, cy, f, k1, k2, k3)
    ray_c = np.stack([x, y, np.ones_like(x)])
    return np.ascontiguousarray(np.einsum("ji,jhw->ihw", np.asarray(R, np.float64), ray_c), np.float32)


def xorshift32_step(y):
    L = orc()
    L.orc_xorshift32_step.restype = ctypes.c_uint32
    return int(L.orc_xorshift32_step(ctypes.c_uint32(y & 0xFFFFFFFF)))


def xorshift_draws(seed, n):
    out = np.empty(n, np.float64)
    orc().orc_xorshift_draws(ctypes.c_uint32(seed & 0xFFFFFFFF), ctypes.c_int64(n), _ptr(out, _f64p))
    return out


def _pm_args(ref, ray, Cr, srcs, src_depths, par):
    ref, ray = _c32(ref), _c32(ray)
    Cr = np.ascontiguousarray(Cr, np.float64)
    par = np.ascontiguousarray(par, np.float64)
    srcs = [_c32(s) for s in srcs]
    deps = [None if d is None else _c32(d) for d in src_depths]
    S = len(srcs)
    imgs = (ctypes.c_void_p * max(S, 1))(*[s.ctypes.data for s in srcs])
    dps = (ctypes.c_void_p * max(S, 1))(*[None if d is None else d.ctypes.data for d in deps])
    ih = np.ascontiguousarray([s.shape[0] for s in srcs], np.int32)
    iw = np.ascontiguousarray([s.shape[1] for s in srcs], np.int32)
    return (ref, ray, Cr, par, srcs, deps, imgs, dps, ih, iw)


def pm_refine(ref, ray, Cr, srcs, src_depths, par, depth, normal, score, active, r, step, top_k, geo_w, geo_tol,
              iters, dmin, dmax, seed):
    """Restatement of _dense.pm_refine (_dense.pyx:566-666); depth/normal/score in place."""
    keep = _pm_args(ref, ray, Cr, srcs, src_depths, par)
    ref, ray, Cr, par, srcs, deps, imgs, dps, ih, iw = keep
    H, W = ref.shape
    for a in (depth, normal, score):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    active = np.ascontiguousarray(active, np.uint8)
    return int(orc().orc_pm_refine(
        _ptr(ref, _f32p), ctypes.c_int(H), ctypes.c_int(W), _ptr(ray, _f32p), _ptr(Cr, _f64p), ctypes.c_int(len(srcs)),
        imgs, dps, _ptr(ih, _i32p), _ptr(iw, _i32p), _ptr(par, _f64p), _ptr(depth, _f32p), _ptr(normal, _f32p),
        _ptr(score, _f32p), _ptr(active, _u8p), ctypes.c_int(r), ctypes.c_int(step), ctypes.c_int(top_k),
        ctypes.c_double(geo_w), ctypes.c_double(geo_tol), ctypes.c_int(iters), ctypes.c_double(dmin),
        ctypes.c_double(dmax), ctypes.c_uint32(int(seed) & 0xFFFFFFFF)))


def bench_sweep(nthreads, ref, mu, sd, srcs, ray, geo, base, step, nd, r, top_k):
    """nthreads concurrent copies of a linear-hypothesis sweep (one view per thread)."""
    ref, mu, sd, ray, base, step = (_c32(a) for a in (ref, mu, sd, ray, base, step))
    geo = np.ascontiguousarray(geo, np.float64)
    keep, ptrs, hw = _src_args(srcs)
    H, W = ref.shape
    rc = orc().orc_bench_sweep(ctypes.c_int(nthreads), _ptr(ref, _f32p), _ptr(mu, _f32p), _ptr(sd, _f32p),
                               ctypes.c_int64(H), ctypes.c_int64(W), ctypes.c_int(len(keep)), ptrs, _ptr(hw, _i64p),
                               _ptr(ray, _f32p), _ptr(geo, _f64p), _ptr(base, _f32p), _ptr(step, _f32p),
                               ctypes.c_int(nd), ctypes.c_int(r), ctypes.c_int(top_k))
    if rc:
        raise RuntimeError(f"bench_sweep rc={rc}")


def bench_pm(nthreads, ref, ray, Cr, srcs, src_depths, par, depth0, normal0, active, r, step, top_k, geo_w, geo_tol,
             iters, dmin, dmax, seed):
    """nthreads concurrent copies of one PatchMatch refinement."""
    keep = _pm_args(ref, ray, Cr, srcs, src_depths, par)
    ref, ray, Cr, par, srcs, deps, imgs, dps, ih, iw = keep
    H, W = ref.shape
    depth0, normal0 = _c32(depth0), _c32(normal0)
    active = np.ascontiguousarray(active, np.uint8)
    rc = orc().orc_bench_pm(ctypes.c_int(nthreads), _ptr(ref, _f32p), ctypes.c_int(H), ctypes.c_int(W),
                            _ptr(ray, _f32p), _ptr(Cr, _f64p), ctypes.c_int(len(srcs)), imgs, dps, _ptr(ih, _i32p),
                            _ptr(iw, _i32p), _ptr(par, _f64p), _ptr(depth0, _f32p), _ptr(normal0, _f32p),
                            _ptr(active, _u8p), ctypes.c_int(r), ctypes.c_int(step), ctypes.c_int(top_k),
                            ctypes.c_double(geo_w), ctypes.c_double(geo_tol), ctypes.c_int(iters),
                            ctypes.c_double(dmin), ctypes.c_double(dmax), ctypes.c_uint32(int(seed) & 0xFFFFFFFF))
    if rc:
        raise RuntimeError(f"bench_pm rc={rc}")


def sample_view(img, cam, X0, Y0, gsd, Z):
    """Restatement of _mvs.sample_view (_mvs.pyx:13-54) over a (..., H, W) stack of heights, as
    CPUBackend.sample_view calls it (backend.py:67-77) -> (out (..., H, W, ch), valid uint8)."""
    R, C, f, k1, k2, k3, cx, cy = cam
    img = _c32(img)
    if img.ndim == 2:
        img = img[:, :, None]
    Z = _c32(Z)
    R = np.ascontiguousarray(R, np.float64)
    C = np.ascontiguousarray(C, np.float64)
    h, w, nch = img.shape
    out = np.zeros(Z.shape + (nch,), np.float32)
    valid = np.zeros(Z.shape, np.uint8)
    Zs = Z.reshape((-1,) + Z.shape[-2:])
    os_ = out.reshape((-1,) + out.shape[-3:])
    vs = valid.reshape((-1,) + Z.shape[-2:])
    H, W = Z.shape[-2:]
    for d in range(Zs.shape[0]):
        zd = np.ascontiguousarray(Zs[d])
        od = np.zeros((H, W, nch), np.float32)
        vd = np.zeros((H, W), np.uint8)
        orc().orc_sample_view(_ptr(img, _f32p), ctypes.c_int64(h), ctypes.c_int64(w), ctypes.c_int64(nch),
                              _ptr(R, _f64p), _ptr(C, _f64p), *(ctypes.c_double(float(v)) for v in
                                                                 (f, k1, k2, k3, cx, cy, X0, Y0, gsd)),
                              _ptr(zd, _f32p), ctypes.c_int64(H), ctypes.c_int64(W), _ptr(od, _f32p), _ptr(vd, _u8p))
        os_[d], vs[d] = od, vd
    return out, valid


# ------------------------------------------------------------------ consistency + fusion (om_fuse.c)
class OrcCam(ctypes.Structure):
    _fields_ = [("R", ctypes.c_double * 9), ("C", ctypes.c_double * 3), ("f", ctypes.c_double),
                ("k1", ctypes.c_double), ("k2", ctypes.c_double), ("k3", ctypes.c_double),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("w", ctypes.c_int32), ("h", ctypes.c_int32)]


def orc_cam(cam) -> OrcCam:
    """OrcCam of any object with R, C, f, k1, k2, k3, cx, cy, w, h (densify._Cam's fields)."""
    c = OrcCam()
    c.R[:] = [float(v) for v in np.asarray(cam.R, np.float64).ravel()]
    c.C[:] = [float(v) for v in np.asarray(cam.C, np.float64).ravel()]
    for k in ("f", "k1", "k2", "k3", "cx", "cy"):
        setattr(c, k, float(getattr(cam, k)))
    c.w, c.h = int(cam.w), int(cam.h)
    return c


def consistency(D, cam, nb_depths, nb_cams, rel_tol, px_tol):
    """Restatement of densify._consistency (densify.py:441-478) -> int16 (H, W) counts."""
    D = _c32(D)
    cr = orc_cam(cam)
    nd = [_c32(x) for x in nb_depths]
    cams = (OrcCam * max(1, len(nd)))(*[orc_cam(c) for c in nb_cams])
    ptrs = (ctypes.c_void_p * max(1, len(nd)))(*[x.ctypes.data for x in nd])
    tol = np.ascontiguousarray(rel_tol, np.float64)
    cnt = np.zeros(D.shape, np.int16)
    rc = orc().orc_consistency(_ptr(D, _f32p), ctypes.byref(cr), ctypes.c_int32(len(nd)), ptrs, cams,
                               _ptr(tol, _f64p), ctypes.c_double(px_tol), cnt.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise RuntimeError(f"orc_consistency rc={rc}")
    return cnt


def fuse(order, depths, cams, rgbs, nbrs, rel_tol, px_tol, min_views, stride, used=None, cap=None):
    """Restatement of densify._fuse_loop (densify.py:688-762).  depths / cams / rgbs: per view
    (None where a view has no depth map); nbrs (n_views, max_nb) int32 with -1 for none (and for
    neighbours without a depth map); rel_tol (n_views, max_nb).  Returns (pts, cols, nv, owner,
    used maps)."""
    n = len(depths)
    order = np.ascontiguousarray(order, np.int32)
    dd = [None if d is None else _c32(d) for d in depths]
    rg = [None if r is None else np.ascontiguousarray(r, np.uint8) for r in rgbs]
    if used is None:
        used = [None if d is None else np.zeros(d.shape, np.uint8) for d in dd]
    camv = (OrcCam * n)(*[orc_cam(c) if c is not None else OrcCam() for c in cams])
    vp = ctypes.c_void_p * n
    dp = vp(*[0 if d is None else d.ctypes.data for d in dd])
    rp = vp(*[0 if r is None else r.ctypes.data for r in rg])
    up = vp(*[0 if u is None else u.ctypes.data for u in used])
    nbrs = np.ascontiguousarray(nbrs, np.int32)
    tol = np.ascontiguousarray(rel_tol, np.float64)
    if cap is None:
        cap = sum(int(np.ceil(d.shape[0] / stride) * np.ceil(d.shape[1] / stride)) for d in dd if d is not None)
    pts = np.zeros((cap, 3), np.float64)
    cols = np.zeros((cap, 3), np.uint8)
    nv = np.zeros(cap, np.uint8)
    owner = np.zeros(cap, np.int32)
    L = orc()
    L.orc_fuse.restype = ctypes.c_int64
    tot = L.orc_fuse(ctypes.c_int32(len(order)), _ptr(order, _i32p), ctypes.c_int32(n), dp, camv, rp, up,
                     _ptr(nbrs, _i32p), ctypes.c_int32(nbrs.shape[1]), _ptr(tol, _f64p), ctypes.c_double(px_tol),
                     ctypes.c_int32(min_views), ctypes.c_int32(stride), _ptr(pts, _f64p), _ptr(cols, _u8p),
                     _ptr(nv, _u8p), _ptr(owner, _i32p), ctypes.c_int64(cap))
    if tot < 0:
        raise RuntimeError(f"orc_fuse rc={tot}")
    m = min(tot, cap)
    return pts[:m], cols[:m], nv[:m], owner[:m], used


# ------------------------------------------------------------------ semi-global matching (om_sgm.c)
def sgm(cost, P1, P2):
    """Restatement of _mvs.sgm (_mvs.pyx:98-250): (D, H, W) float32 -> aggregated (D, H, W)."""
    cost = _c32(cost)
    D, H, W = cost.shape
    agg = np.zeros((D, H, W), np.float32)
    rc = orc().orc_sgm(_ptr(cost, _f32p), ctypes.c_int64(D), ctypes.c_int64(H), ctypes.c_int64(W),
                       ctypes.c_float(P1), ctypes.c_float(P2), _ptr(agg, _f32p))
    if rc:
        raise RuntimeError(f"orc_sgm rc={rc}")
    return agg


# ------------------------------------------------------------------ matching-image preparation
def depth_view(rgb, gain=None, bias=None, mask=None):
    """Restatement of densify.load_view's NumPy path (densify.py:537-547), which the reference's
    test pairs bit for bit with _fast.depth_view (test_native_parity.py:380-398):
    -> (grey (h, w) float32, rgb8 (h, w, 3) uint8)."""
    v = np.asarray(rgb, np.uint8).astype(np.float32)
    if gain is not None:
        v *= np.asarray(gain, np.float32)
    if bias is not None:
        v += np.asarray(bias, np.float32)
    s = v[..., 0] + v[..., 1]
    s = s + v[..., 2]
    s = (s / np.float32(3.0)).astype(np.float32)
    grey = np.ascontiguousarray(((s - np.float32(128.0)) / np.float32(64.0)).astype(np.float32))
    if mask is not None:
        grey[~np.asarray(mask, bool)] = np.nan
    rgb8 = np.clip(v + np.float32(0.5), 0, 255).astype(np.uint8)
    return grey, rgb8
