"""Device-side mirror of the reference's `_dense` / `_fast` depth-map kernels on torch tensors.

Every function keeps the reference's name, argument order and meaning, and raises the exception
type the reference raises for the same bad input (ValueError for sizes and shapes, MemoryError
for allocation failures).  Arrays are CUDA (HIP) torch tensors, C-contiguous, float32 unless the
reference uses another type (idx: int32, active: uint8); outputs are written in place like the
reference's typed memoryviews.  Work is enqueued on the current torch stream.

Reference: orthomosaic/_dense.pyx:566-666 (pm_refine), :740-838 (sweep_hypothesis);
orthomosaic/_fast.pyx:628-713 (hyp_linear, hyp_refine, box_zero_mean).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import OmSrc, OmSweepDesc, check

__all__ = ["sweep_hypothesis", "sweep", "fused_ok", "sweep_finalize", "pm_refine", "box_zero_mean", "hyp_linear",
           "hyp_refine", "half", "ref_stats", "SweepState"]


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _need(t: torch.Tensor, name: str, dtype=torch.float32, shape=None) -> int:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device (cuda) tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t.data_ptr()


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


class SweepState:
    """The streaming winner state of _sweep (densify.py:331-335): best/prev/s_prev_best/s_next_best
    start at -2, idx at -1."""

    def __init__(self, H: int, W: int, device="cuda"):
        # the four score planes share one allocation, so a fresh or reset state costs two fills
        self._scores = torch.full((4, H, W), -2.0, dtype=torch.float32, device=device)
        self.best, self.prev, self.s_prev_best, self.s_next_best = self._scores.unbind(0)
        self.idx = torch.full((H, W), -1, dtype=torch.int32, device=device)

    def tensors(self):
        return self.best, self.prev, self.s_prev_best, self.s_next_best, self.idx

    def reset(self):
        planes = (self.best, self.prev, self.s_prev_best, self.s_next_best)
        shared = getattr(self, "_scores", None)
        if shared is not None and all(t.data_ptr() == shared[k].data_ptr() for k, t in enumerate(planes)):
            shared.fill_(-2.0)
        else:
            for t in planes:
                t.fill_(-2.0)
        self.idx.fill_(-1)


def _src_array(srcs: Sequence[torch.Tensor], geo: np.ndarray, dtype=torch.float32):
    n = len(srcs)
    arr = (OmSrc * max(n, 1))()
    g = np.ascontiguousarray(geo, dtype=np.float64)
    for s, img in enumerate(srcs):
        if img.dim() != 2:
            raise ValueError("sources must be 2-D images")
        arr[s].img = _need(img, f"srcs[{s}]", dtype)
        arr[s].h, arr[s].w = int(img.shape[0]), int(img.shape[1])
        ctypes.memmove(ctypes.addressof(arr[s].geo), g[s].ctypes.data, 18 * 8)
    return arr


def _geo(geo, n: int) -> np.ndarray:
    g = np.ascontiguousarray(geo.detach().cpu().numpy() if isinstance(geo, torch.Tensor) else geo, np.float64)
    if g.ndim != 2 or g.shape[0] != n or g.shape[1] != 18:
        raise ValueError("geo must be (n_sources, 18)")
    return g


def sweep_hypothesis(ref, mu_i, sd_i, srcs, ray, geo, inv, r, top_k, i, best, prev, s_prev_best, s_next_best,
                     idx) -> None:
    """One plane-sweep hypothesis for all sources (_dense.pyx:740-838), one kernel launch.

    geo: (n, 18) float64 per source = Rs (9, row-major), b = Rs (C_ref - C_src) (3), f, k1, k2, k3,
    cx, cy (host array or tensor).  inv: (H, W) inverse depth of hypothesis i."""
    n = len(srcs)
    if n < 1 or n > 16:
        raise ValueError("sweep_hypothesis needs 1..16 sources")
    g = _geo(geo, n)
    H, W = int(ref.shape[0]), int(ref.shape[1])
    for name, t in (("mu_i", mu_i), ("sd_i", sd_i), ("inv", inv), ("best", best), ("prev", prev),
                    ("s_prev_best", s_prev_best), ("s_next_best", s_next_best), ("idx", idx)):
        if tuple(t.shape) != (H, W):
            raise ValueError("sweep_hypothesis: array shapes do not match the reference image")
    if tuple(ray.shape) != (3, H, W):
        raise ValueError("sweep_hypothesis: array shapes do not match the reference image")
    if r < 0:
        raise ValueError("sweep_hypothesis: negative window radius")
    if H == 0 or W == 0:
        return
    if not fused_ok(n, r, top_k):
        # the reference takes any r and top_k here: the row/column-streaming kernels, map i0 = i
        st = SweepState.__new__(SweepState)
        st.best, st.prev, st.s_prev_best, st.s_next_best, st.idx = best, prev, s_prev_best, s_next_best, idx
        _need(inv, "inv")
        d, keep = _sweep_desc(ref, mu_i, sd_i, srcs, ray, g, r, top_k, st, _lib.OM_HYP_MAPS, inv, None, 0, int(i), 1)
        L = _lib.lib()
        nbytes = L.om_sweep_rows_workspace_bytes(ctypes.byref(d))
        ws = _workspace(nbytes, ref.device)
        d.workspace, d.workspace_bytes = ws.data_ptr(), nbytes
        check(L.om_sweep_rows(ctypes.byref(d), _stream()), "sweep_hypothesis")
        del keep, ws
        return
    ptrs = (ctypes.c_void_p * n)(*[_need(s, f"srcs[{k}]") for k, s in enumerate(srcs)])
    hw = (ctypes.c_int32 * (2 * n))(*[v for s in srcs for v in (int(s.shape[0]), int(s.shape[1]))])
    gd = (ctypes.c_double * (18 * n))(*g.ravel().tolist())
    rc = _lib.lib().om_sweep_hypothesis(
        _need(ref, "ref"), _need(mu_i, "mu_i"), _need(sd_i, "sd_i"), H, W, n, ptrs, hw, _need(ray, "ray"), gd,
        _need(inv, "inv"), int(r), int(top_k), int(i), _need(best, "best"), _need(prev, "prev"),
        _need(s_prev_best, "s_prev_best"), _need(s_next_best, "s_next_best"),
        _need(idx, "idx", torch.int32), _stream())
    check(rc, "sweep_hypothesis")


def _sweep_desc(ref, mu, sd, srcs, ray, geo, r, top_k, state: SweepState, mode, hyp_a, hyp_b=None,
                hyp_offset=0, i0=0, nh=0, chunks=0, tile_h=0, src_dtype=torch.float32):
    n = len(srcs)
    if n < 1 or n > MAX_SOURCES:
        raise ValueError(f"sweep needs 1..{MAX_SOURCES} sources")
    g = _geo(geo, n)
    H, W = int(ref.shape[0]), int(ref.shape[1])
    d = OmSweepDesc()
    d.ref, d.mu, d.sd = _need(ref, "ref", shape=(H, W)), _need(mu, "mu", shape=(H, W)), _need(sd, "sd", shape=(H, W))
    d.ray = _need(ray, "ray", shape=(3, H, W))
    d.H, d.W, d.n_src = H, W, n
    arr = _src_array(srcs, g, src_dtype)
    d.src = ctypes.cast(arr, ctypes.POINTER(OmSrc))
    d.r, d.top_k = int(r), int(top_k)
    d.hyp_mode = int(mode)
    d.hyp_a = hyp_a.data_ptr() if hyp_a is not None else None
    d.hyp_b = hyp_b.data_ptr() if hyp_b is not None else None
    d.hyp_offset = int(hyp_offset)
    d.i0, d.nh = int(i0), int(nh)
    d.best = _need(state.best, "best", shape=(H, W))
    d.prev = _need(state.prev, "prev", shape=(H, W))
    d.s_prev_best = _need(state.s_prev_best, "s_prev_best", shape=(H, W))
    d.s_next_best = _need(state.s_next_best, "s_next_best", shape=(H, W))
    d.idx = _need(state.idx, "idx", torch.int32, shape=(H, W))
    d.chunks, d.tile_h = int(chunks), int(tile_h)
    return d, arr


def _hyp_inputs(hyp: dict, H: int, W: int):
    """hyp: {"mode": "maps", "maps": (D,H,W)} | {"mode": "linear", "base": (H,W), "step": (H,W), "n": D}
    | {"mode": "refine", "iu": (H,W), "ds": (H,W), "steps": rs} (2 rs + 1 hypotheses, t = i - rs)
    | {"mode": "const", "inv": (D,)}"""
    m = hyp["mode"]
    if m == "maps":
        maps = hyp["maps"]
        _need(maps, "maps", shape=(maps.shape[0], H, W))
        return _lib.OM_HYP_MAPS, maps, None, 0, int(maps.shape[0])
    if m == "linear":
        _need(hyp["base"], "base", shape=(H, W))
        _need(hyp["step"], "step", shape=(H, W))
        return _lib.OM_HYP_LINEAR, hyp["base"], hyp["step"], 0, int(hyp["n"])
    if m == "refine":
        _need(hyp["iu"], "iu", shape=(H, W))
        _need(hyp["ds"], "ds", shape=(H, W))
        rs = int(hyp["steps"])
        return _lib.OM_HYP_REFINE, hyp["iu"], hyp["ds"], rs, 2 * rs + 1
    if m == "const":
        inv = hyp["inv"]
        _need(inv, "inv", shape=(inv.shape[0],))
        return _lib.OM_HYP_CONST, inv, None, 0, int(inv.shape[0])
    raise ValueError(f"unknown hypothesis mode {m!r}")


MAX_SOURCES = 256      # sources of om_sweep_rows; the fused kernels take 16
FUSED_MAX_R = 7        # window radius of the fused kernels
FUSED_MAX_K = 4        # min(top_k, sources) of the fused kernels


def fused_ok(n: int, r: int, top_k: int) -> bool:
    """Whether the fused kernels take this sweep (else om_sweep_rows does)."""
    return 1 <= n <= 16 and 0 <= r <= FUSED_MAX_R and 1 <= min(top_k, n) <= FUSED_MAX_K


def sweep(ref, mu, sd, srcs, ray, geo, hyp: dict, r: int, top_k: int, state: Optional[SweepState] = None,
          i0: int = 0, nh: Optional[int] = None, chunks: int = 0, tile_h: int = 0, exact: bool = False) -> SweepState:
    """Hypotheses [i0, i0+nh) of one view: equal to calling sweep_hypothesis for i = i0 .. i0+nh-1
    in order (densify.py:339-342) on the same state.

    The fused kernel (one launch) takes 1..16 sources, r <= 7 and min(top_k, n) <= 4; its window sums
    equal the reference's whenever no FP64 partial sum rounds.  Otherwise, or with exact=True, the
    row/column-streaming kernels (om_sweep_rows) run the reference's ncc_warp + combine order, bit
    for bit for any input (densify.py:343-353); `chunks` is then the hypotheses per pass."""
    H, W = int(ref.shape[0]), int(ref.shape[1])
    mode, a, b, off, n_all = _hyp_inputs(hyp, H, W)
    if state is None:
        state = SweepState(H, W, ref.device)
    if nh is None:
        nh = n_all - i0
    if r < 0:
        raise ValueError("sweep: negative window radius")
    rows = exact or not fused_ok(len(srcs), r, top_k)
    d, keep = _sweep_desc(ref, mu, sd, srcs, ray, geo, r, top_k, state, mode, a, b, off, i0, nh, chunks,
                          0 if rows else tile_h)
    if mode in (_lib.OM_HYP_MAPS, _lib.OM_HYP_CONST) and i0 != 0:
        raise ValueError("maps / const hypotheses are indexed from i0 = 0 in this call")
    L = _lib.lib()
    if rows:
        nbytes = L.om_sweep_rows_workspace_bytes(ctypes.byref(d))
        if nbytes == 0 and H * W and nh:
            check(L.om_sweep_rows(ctypes.byref(d), _stream()), "sweep_rows")   # raises with the reason
        ws = _workspace(nbytes, ref.device)
        d.workspace, d.workspace_bytes = ws.data_ptr(), nbytes
        check(L.om_sweep_rows(ctypes.byref(d), _stream()), "sweep_rows")
        del keep, ws
        return state
    nbytes = L.om_sweep_workspace_bytes(ctypes.byref(d))
    ws = None
    if nbytes:
        ws = _workspace(nbytes, ref.device)
        d.workspace, d.workspace_bytes = ws.data_ptr(), nbytes
    check(L.om_sweep(ctypes.byref(d), _stream()), "sweep")
    del keep, ws  # torch's caching allocator keeps the workspace alive for the enqueued kernels
    return state


def sweep_finalize(ref, mu, sd, srcs, ray, geo, hyp: dict, r: int, top_k: int, state: SweepState) -> torch.Tensor:
    """Sub-step parabola and depth of _sweep (densify.py:354-361) for the full hypothesis set."""
    H, W = int(ref.shape[0]), int(ref.shape[1])
    mode, a, b, off, n_all = _hyp_inputs(hyp, H, W)
    d, keep = _sweep_desc(ref, mu, sd, srcs, ray, geo, r, top_k, state, mode, a, b, off, 0, n_all)
    depth = torch.empty((H, W), dtype=torch.float32, device=ref.device)
    check(_lib.lib().om_sweep_finalize(ctypes.byref(d), depth.data_ptr(), _stream()), "sweep_finalize")
    del keep
    return depth


def hyp_linear(base, step, i) -> torch.Tensor:
    """(base + i * step).astype(float32) (_fast.pyx:628-641)."""
    if tuple(step.shape) != tuple(base.shape):
        raise ValueError("hyp_linear: shapes differ")
    out = torch.empty_like(base)
    check(_lib.lib().om_hyp_linear(_need(base, "base"), _need(step, "step"), base.numel(), int(i), out.data_ptr(),
                                   _stream()), "hyp_linear")
    return out


def hyp_refine(iu, ds, t) -> torch.Tensor:
    """np.maximum(iu + t * ds, 1e-6) in float32 (_fast.pyx:644-658)."""
    if tuple(ds.shape) != tuple(iu.shape):
        raise ValueError("hyp_refine: shapes differ")
    out = torch.empty_like(iu)
    check(_lib.lib().om_hyp_refine(_need(iu, "iu"), _need(ds, "ds"), iu.numel(), int(t), out.data_ptr(), _stream()),
          "hyp_refine")
    return out


def box_zero_mean(a, r: int) -> torch.Tensor:
    """densify._box: zero-padded (2r+1)^2 window mean from float32 cumulative sums (_fast.pyx:661-713)."""
    if r < 0:
        raise ValueError("box_zero_mean needs r >= 0")
    H, W = int(a.shape[0]), int(a.shape[1])
    out = torch.empty((H, W), dtype=torch.float32, device=a.device)
    if H == 0 or W == 0:
        return out
    L = _lib.lib()
    nb = L.om_box_workspace_bytes(H, W, int(r))
    ws = _workspace(nb, a.device)
    check(L.om_box_zero_mean(_need(a, "a"), H, W, int(r), out.data_ptr(), ws.data_ptr(), nb, _stream()),
          "box_zero_mean")
    return out


def ref_stats(ref, r: int):
    """(mu, sd) of densify._sweep: mu = _box(ref), sd = sqrt(max(_box(ref*ref) - mu*mu, 1e-6))
    (densify.py:327-328)."""
    H, W = int(ref.shape[0]), int(ref.shape[1])
    mu = torch.empty((H, W), dtype=torch.float32, device=ref.device)
    sd = torch.empty((H, W), dtype=torch.float32, device=ref.device)
    if H == 0 or W == 0:
        return mu, sd
    L = _lib.lib()
    nb = 2 * L.om_box_workspace_bytes(H, W, int(r))
    ws = _workspace(nb, ref.device)
    check(L.om_ref_stats(_need(ref, "ref"), H, W, int(r), mu.data_ptr(), sd.data_ptr(), ws.data_ptr(), nb,
                         _stream()), "ref_stats")
    return mu, sd


def half(g) -> torch.Tensor:
    """densify._half: 2x2 box downsample in float32 (densify.py:127-130)."""
    H, W = int(g.shape[0]), int(g.shape[1])
    out = torch.empty((H // 2, W // 2), dtype=torch.float32, device=g.device)
    check(_lib.lib().om_half(_need(g, "g"), H, W, out.data_ptr(), _stream()), "half")
    return out


def rays_world(H: int, W: int, R, cx: float, cy: float, f: float, k1: float = 0.0, k2: float = 0.0,
               k3: float = 0.0, device="cuda") -> torch.Tensor:
    """densify._rays_world (densify.py:288-297): (3, H, W) float32 world rays (camera-axis depth 1)."""
    Rm = np.ascontiguousarray(np.asarray(R, np.float64).reshape(3, 3))
    out = torch.empty((3, H, W), dtype=torch.float32, device=device)
    rd = (ctypes.c_double * 9)(*Rm.ravel().tolist())
    check(_lib.lib().om_rays_world(int(H), int(W), rd, float(cx), float(cy), float(f), float(k1), float(k2),
                                   float(k3), out.data_ptr(), _stream()), "rays_world")
    return out


def pm_refine(ref, ray, Cr, srcs, src_depths, par, depth, normal, score, active, r, step, top_k, geo_w, geo_tol,
              iters, dmin, dmax, seed, fmt: Optional[str] = None, stats: Optional[dict] = None) -> int:
    """PatchMatch repair of a depth map (_dense.pyx:566-666): only `active` pixels are re-estimated;
    depth / normal / score are updated in place.  Returns the number of accepted changes.
    fmt "f16" / "bf16" / "f32": the precision variant, with `srcs` in that storage (to_lp).
    stats: if given, receives "evaluated", the propagation cost evaluations actually computed."""
    S = len(srcs)
    if S == 0 or S > 16:
        return 0
    H, W = int(ref.shape[0]), int(ref.shape[1])
    p = np.ascontiguousarray(par.detach().cpu().numpy() if isinstance(par, torch.Tensor) else par, np.float64)
    if p.shape != (S, 18):
        raise ValueError("par must be (n_sources, 18)")
    cr = np.ascontiguousarray(Cr.detach().cpu().numpy() if isinstance(Cr, torch.Tensor) else Cr, np.float64)
    if cr.shape != (3,):
        raise ValueError("Cr must have 3 values")
    if len(src_depths) != S:
        raise ValueError("src_depths must have one entry (or None) per source")
    sdt = torch.float32 if fmt is None or fmt == "f32" else torch.int16
    imgs = (ctypes.c_void_p * S)(*[_need(s, f"srcs[{k}]", sdt) for k, s in enumerate(srcs)])
    deps = (ctypes.c_void_p * S)(*[None if d is None else _need(d, f"src_depths[{k}]")
                                  for k, d in enumerate(src_depths)])
    for k, d in enumerate(src_depths):
        if d is not None and tuple(d.shape) != tuple(srcs[k].shape):
            raise ValueError("a source depth map must have the shape of its image")
    ih = (ctypes.c_int32 * S)(*[int(s.shape[0]) for s in srcs])
    iw = (ctypes.c_int32 * S)(*[int(s.shape[1]) for s in srcs])
    pd = (ctypes.c_double * (18 * S))(*p.ravel().tolist())
    cd = (ctypes.c_double * 3)(*cr.tolist())
    L = _lib.lib()
    nb = L.om_pm_workspace_bytes(H, W)
    ws = _workspace(nb, ref.device)
    changed = ctypes.c_int64(0)
    args = (_need(ref, "ref", shape=(H, W)), H, W, _need(ray, "ray", shape=(3, H, W)), cd, S, imgs, deps, ih, iw, pd,
            _need(depth, "depth", shape=(H, W)), _need(normal, "normal", shape=(3, H, W)),
            _need(score, "score", shape=(H, W)), _need(active, "active", torch.uint8, shape=(H, W)), int(r),
            int(step), int(top_k), float(geo_w), float(geo_tol), int(iters), float(dmin), float(dmax),
            int(seed) & 0xFFFFFFFF, ws.data_ptr(), nb, ctypes.byref(changed))
    if fmt is None:
        rc = L.om_pm_refine(*args, _stream())
    else:
        rc = L.om_pm_refine_lp(*args, FORMATS[fmt], _stream())
    check(rc, "pm_refine")
    if stats is not None:
        ev = ctypes.c_int64(0)
        check(L.om_pm_workspace_counts(ws.data_ptr(), H, W, ctypes.byref(ev)), "pm counts")
        stats["evaluated"] = int(ev.value)
    return int(changed.value)


# ---------------------------------------------------------------- reduced precision (FP16 / BF16)
FORMATS = {"f16": _lib.OM_FMT_F16, "bf16": _lib.OM_FMT_BF16, "f32": _lib.OM_FMT_F32}


def _lp_dtype(fmt: str):
    return torch.float32 if fmt == "f32" else torch.int16


def to_lp(a: torch.Tensor, fmt: str) -> torch.Tensor:
    """float32 -> FP16 / BF16 storage (round to nearest even), returned as an int16 tensor of
    the same shape holding the 16-bit patterns; "f32": a float32 copy (the variants' FP32
    arithmetic without the 16-bit rounding)."""
    f = FORMATS[fmt]
    src = a.contiguous()
    _need(src, "image")
    out = torch.empty(src.shape, dtype=_lp_dtype(fmt), device=src.device)
    check(_lib.lib().om_to_lp(src.data_ptr(), src.numel(), f, out.data_ptr(), _stream()), "to_lp")
    return out


def sweep_lp(ref16, mu, sd, srcs16, ray, geo, hyp: dict, r: int, top_k: int, fmt: str,
             state: Optional[SweepState] = None, i0: int = 0, nh: Optional[int] = None) -> SweepState:
    """The precision variant of `sweep`: images in FP16/BF16 (or FP32) storage (to_lp), FP32 warp, sums and
    NCC, FP64 plane depth and variance guard.  Same winner state; finish with sweep_finalize."""
    H, W = int(ref16.shape[0]), int(ref16.shape[1])
    mode, a, b, off, n_all = _hyp_inputs(hyp, H, W)
    if state is None:
        state = SweepState(H, W, mu.device)
    if nh is None:
        nh = n_all - i0
    for t in [ref16] + list(srcs16):
        _need(t, "stored image", _lp_dtype(fmt))
    d, keep = _sweep_desc(mu, mu, sd, list(srcs16), ray, geo, r, top_k, state, mode, a, b, off, i0, nh,
                          src_dtype=_lp_dtype(fmt))
    check(_lib.lib().om_sweep_lp(ctypes.byref(d), FORMATS[fmt], ref16.data_ptr(), _stream()), "sweep_lp")
    del keep
    return state

