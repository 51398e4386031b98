"""Geometric consistency and depth-map fusion on the device, with the reference's interfaces.

- consistency(k, D, nbrs_k, dm, cams, px_tol, rel_tol) is densify._consistency
  (densify.py:441-478): per pixel of view k's depth map the number of neighbour depth maps that
  agree with it (om_consistency).
- fuse(order, dm, used, cams, nbrs, have, rel_tols, opt, px_tol, min_views, rgb, pts, cols, nv,
  owner) is densify._fuse / _fuse_loop (densify.py:676-762): sequential fusion over the views in
  `order`, appending the fused points (device tensors) to the four lists (om_fuse).

Depth maps, colours and the used maps are device tensors; cameras are any objects with the fields
of densify._Cam (R, C, f, k1, k2, k3, cx, cy, w, h).  The used maps live on the stride grid
(ceil(h / stride), ceil(w / stride)): the reference's full-resolution used map is only ever set and
read at grid pixels (densify.py:698-701, 749-750); used_full() expands one for comparison.
Results equal the reference's bit for bit (FP64 in its order, 3x3 products as NumPy's OpenBLAS
forms them).  There is no CPU fallback: without libom355.so these functions raise.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Mapping, Sequence

import numpy as np
import torch

from . import _lib


def om_cam(cam) -> _lib.OmCam:
    """OmCam of a densify._Cam-like camera (float64 parameters)."""
    c = _lib.OmCam()
    c.R[:] = [float(v) for v in np.asarray(cam.R, np.float64).ravel()]
    c.C[:] = [float(v) for v in np.asarray(cam.C, np.float64).ravel()]
    for k in ("f", "k1", "k2", "k3", "cx", "cy"):
        setattr(c, k, float(getattr(cam, k)))
    c.w, c.h = int(cam.w), int(cam.h)
    return c


def _dev_f32(t, h, w, what):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError(f"{what}: contiguous float32 device tensor expected")
    if tuple(t.shape) != (h, w):
        raise ValueError(f"{what}: shape {tuple(t.shape)}, camera says {(h, w)}")
    return t


def consistency(k, D, nbrs_k, dm, cams, px_tol, rel_tol, stream=None) -> torch.Tensor:
    """densify._consistency (densify.py:441-478) -> (H, W) int16 device tensor.

    D: view k's depth map; dm: mapping j -> depth map (neighbours missing from dm are skipped, as
    `j in dm` does); cams[j]: cameras; rel_tol: mapping j -> relative depth tolerance."""
    cr = cams[k]
    D = _dev_f32(D, int(cr.h), int(cr.w), "consistency: D")
    nb = [j for j in nbrs_k if j in dm]
    if len(nb) > 16:
        raise ValueError(f"consistency: {len(nb)} neighbours (at most 16)")
    deps = [_dev_f32(dm[j], int(cams[j].h), int(cams[j].w), f"consistency: dm[{j}]") for j in nb]
    n = max(1, len(nb))
    ptrs = (ctypes.c_void_p * n)(*[d.data_ptr() for d in deps])
    ncams = (_lib.OmCam * n)(*[om_cam(cams[j]) for j in nb])
    tol = (ctypes.c_double * n)(*[float(rel_tol[j]) for j in nb])
    cnt = torch.empty(D.shape, dtype=torch.int16, device=D.device)
    st = stream if stream is not None else torch.cuda.current_stream(D.device).cuda_stream
    rc = _lib.lib().om_consistency(D.data_ptr(), ctypes.byref(om_cam(cr)), len(nb), ptrs, ncams, tol, float(px_tol),
                                   cnt.data_ptr(), st)
    _lib.check(rc, "om_consistency")
    return cnt


def grid_shape(h: int, w: int, stride: int):
    st = max(1, int(stride))
    return (-(-int(h) // st), -(-int(w) // st))


def new_used(dm: Mapping, cams, stride: int, device="cuda") -> dict:
    """Zeroed used maps on the stride grid for every view of dm (densify.py:657)."""
    return {k: torch.zeros(grid_shape(cams[k].h, cams[k].w, stride), dtype=torch.uint8, device=device) for k in dm}


def used_full(used_grid, h: int, w: int, stride: int) -> np.ndarray:
    """The reference's full-resolution used map of a grid used map."""
    st = max(1, int(stride))
    out = np.zeros((h, w), bool)
    g = used_grid.cpu().numpy() if isinstance(used_grid, torch.Tensor) else np.asarray(used_grid)
    out[::st, ::st] = g.astype(bool)
    return out


class Fuser:
    """One prepared om_fuse call over fixed views: the view table, neighbour table, used maps,
    output buffers and workspace are built once; run() enqueues the whole sequential fusion.

    Arguments as fuse(); used: k -> grid used map, or None for fresh zeroed maps (run() with
    reset_used=True zeroes them again before fusing, as a new densify run starts from zero)."""

    def __init__(self, order, dm: Mapping, cams, nbrs, have, rel_tols: Callable, px_tol, min_views, stride=1,
                 rgb=None, used=None, cap: int | None = None):
        stride = max(1, int(stride or 1))
        order = [int(k) for k in order]
        if not order:
            raise ValueError("Fuser: empty view order")
        views = sorted(set(order) | {j for k in order for j in nbrs[k] if j in have})
        dev = dm[order[0]].device
        self._used_buf = None
        if used is None:                 # all grid maps in one allocation: one fill clears them
            shapes = {k: grid_shape(cams[k].h, cams[k].w, stride) for k in views}
            self._used_buf = torch.zeros(sum(a * b for a, b in shapes.values()), dtype=torch.uint8, device=dev)
            used, o = {}, 0
            for k in views:
                a, b = shapes[k]
                used[k] = self._used_buf[o:o + a * b].view(a, b)
                o += a * b
        n_views = max(views) + 1
        V = (_lib.OmFuseView * n_views)()
        self._keep = []
        for k in views:
            cam = cams[k]
            d = _dev_f32(dm[k], int(cam.h), int(cam.w), f"fuse: dm[{k}]")
            u = used[k]
            if tuple(u.shape) != grid_shape(cam.h, cam.w, stride) or u.dtype != torch.uint8 or not u.is_contiguous():
                raise ValueError(f"fuse: used[{k}] must be a contiguous uint8 grid map of shape "
                                 f"{grid_shape(cam.h, cam.w, stride)}")
            V[k].depth, V[k].used, V[k].cam = d.data_ptr(), u.data_ptr(), om_cam(cam)
            self._keep.append(d)
            if rgb is not None and k in order:
                c = rgb[k]
                if not (isinstance(c, torch.Tensor) and c.is_cuda and c.dtype == torch.uint8 and c.is_contiguous()
                        and tuple(c.shape) == (int(cam.h), int(cam.w), 3)):
                    raise ValueError(f"fuse: rgb[{k}] must be a contiguous (h, w, 3) uint8 device tensor")
                V[k].rgb = c.data_ptr()
                self._keep.append(c)
        max_nb = max(1, max(len(nbrs[k]) for k in order))
        nb = np.full((n_views, max_nb), -1, np.int32)
        tol = np.zeros((n_views, max_nb), np.float64)
        for k in order:
            t = rel_tols(k)
            for s, j in enumerate(nbrs[k]):
                if j in have:
                    nb[k, s] = j
                    tol[k, s] = float(t[j])
        self.candidates = sum(grid_shape(cams[k].h, cams[k].w, stride)[0] * grid_shape(cams[k].h, cams[k].w, stride)[1]
                              for k in order)
        self.pixels = sum(int(cams[k].h) * int(cams[k].w) for k in order)
        if cap is None:
            cap = self.candidates
        L = _lib.lib()
        self.ws_bytes = L.om_fuse_workspace_bytes(V, n_views, stride)
        self.ws = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=dev)
        self.P = torch.empty((cap, 3), dtype=torch.float64, device=dev)
        self.Cl = torch.empty((cap, 3), dtype=torch.uint8, device=dev) if rgb is not None else None
        self.N = torch.empty(cap, dtype=torch.uint8, device=dev)
        self.O = torch.empty(cap, dtype=torch.int32, device=dev)
        self.tot = torch.zeros(2, dtype=torch.int64, device=dev)
        self.used, self.views, self.cap, self.device = used, views, int(cap), dev
        self._V, self._nb, self._tol = V, nb, tol
        self._order = np.ascontiguousarray(order, np.int32)
        self._args = (float(px_tol), int(min_views), stride, max_nb, n_views)

    def run(self, stream=None, reset_used: bool = False) -> None:
        """Enqueue the fusion on `stream` (torch's current stream by default)."""
        px_tol, min_views, stride, max_nb, n_views = self._args
        if reset_used:
            if self._used_buf is not None:
                self._used_buf.zero_()
            else:
                for k in self.views:
                    self.used[k].zero_()
        st = stream if stream is not None else torch.cuda.current_stream(self.device).cuda_stream
        rc = _lib.lib().om_fuse(len(self._order), self._order.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n_views,
                                self._V, self._nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), max_nb,
                                self._tol.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), px_tol, min_views, stride,
                                self.P.data_ptr(), self.Cl.data_ptr() if self.Cl is not None else None,
                                self.N.data_ptr(), self.O.data_ptr(), self.cap, self.tot.data_ptr(),
                                self.ws.data_ptr(), self.ws_bytes, st)
        _lib.check(rc, "om_fuse")

    def evaluated(self) -> int:
        """Grid pixels the last run evaluated (finite and not used when their view came up)."""
        torch.cuda.synchronize(self.device)
        return int(self.tot[1].item())

    def result(self):
        """(pts, cols or None, nv, owner) device tensors of the last run (synchronises)."""
        torch.cuda.synchronize(self.device)
        n = int(self.tot[0].item())
        if n > self.cap:
            raise RuntimeError(f"fuse: {n} points exceed the output capacity {self.cap}")
        return self.P[:n], (self.Cl[:n] if self.Cl is not None else None), self.N[:n], self.O[:n]


def fuse(order: Sequence[int], dm: Mapping, used, cams, nbrs, have, rel_tols: Callable, opt, px_tol, min_views,
         rgb, pts: list, cols: list, nv: list, owner: list, cap: int | None = None, stream=None):
    """densify._fuse_loop (densify.py:688-762) on the device.

    dm: k -> (h, w) float32 depth map; used: k -> grid used map (new_used) or None to start from
    zero; cams[k]: cameras; nbrs[k]: neighbour lists; have: the views with depth maps; rel_tols(k)
    -> {j: tolerance}; opt.point_stride; rgb: k -> (h, w, 3) uint8 device colours, or None for no
    colours.  Appends one tensor each to pts (n, 3) float64, cols (n, 3) uint8, nv (n,) uint8 and
    owner (n,) int32 (the reference appends one per view; the concatenation is the same) and
    returns the used maps."""
    if not list(order):
        return used
    f = Fuser(order, dm, cams, nbrs, have, rel_tols, px_tol, min_views, getattr(opt, "point_stride", 1), rgb, used,
              cap)
    f.run(stream)
    P, C, N, O = f.result()
    pts.append(P)
    if C is not None:
        cols.append(C)
    nv.append(N)
    owner.append(O)
    return f.used
