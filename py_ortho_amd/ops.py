"""Backend array operations of the ground-grid height sweep on the device (libom355 kernels).

Reference: CPUBackend.sample_view / box / topk_mean / argmin0 / argmax0 (orthomosaic/backend.py:67-99),
used by mvs.dense_reconstruct's score_at (mvs.py:254-304), and _mvs.sgm (_mvs.pyx:98-250), its
cost aggregation on the coarse level (mvs.py:325).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .dense import _need, _stream, _workspace


def sample_view(img: torch.Tensor, cam, X0, Y0, gsd, Z: torch.Tensor):
    """img (h, w, ch) float32, cam = (R, C, f, k1, k2, k3, cx, cy), Z (H, W) or (..., H, W) heights
    -> (samples (..., H, W, ch) float32, valid (..., H, W) uint8)."""
    R, C, f, k1, k2, k3, cx, cy = cam
    if img.dim() == 2:
        img = img[:, :, None]
    img = img.contiguous()
    Z = Z.contiguous()
    h, w, nch = (int(v) for v in img.shape)
    H, W = int(Z.shape[-2]), int(Z.shape[-1])
    planes = int(np.prod(Z.shape[:-2])) if Z.dim() > 2 else 1
    out = torch.zeros(tuple(Z.shape) + (nch,), dtype=torch.float32, device=Z.device)
    valid = torch.zeros(tuple(Z.shape), dtype=torch.uint8, device=Z.device)
    Rd = (ctypes.c_double * 9)(*np.asarray(R, np.float64).ravel().tolist())
    Cd = (ctypes.c_double * 3)(*np.asarray(C, np.float64).ravel().tolist())
    check(_lib.lib().om_sample_view(_need(img, "img"), h, w, nch, Rd, Cd, float(f), float(k1), float(k2), float(k3),
                                    float(cx), float(cy), float(X0), float(Y0), float(gsd), _need(Z, "Z"), planes, H,
                                    W, out.data_ptr(), valid.data_ptr(), _stream()), "sample_view")
    return out, valid


def box(a: torch.Tensor, r: int) -> torch.Tensor:
    """Zero-padded (2r+1)^2 window mean over the last two axes from float64 summed-area tables."""
    if r < 0:
        raise ValueError("radius must be nonnegative")
    a = a.contiguous()
    shape = tuple(a.shape)
    H, W = shape[-2], shape[-1]
    B = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    out = torch.empty(shape, dtype=torch.float32, device=a.device)
    L = _lib.lib()
    nb = L.om_box_sat_workspace_bytes(B, H, W, int(r))
    ws = _workspace(nb, a.device)
    check(L.om_box_sat_mean(_need(a, "a"), B, H, W, int(r), out.data_ptr(), ws.data_ptr(), nb, _stream()), "box")
    return out


def topk_mean(st: torch.Tensor, k: int) -> torch.Tensor:
    """Mean of the k largest values along axis 0."""
    st = st.contiguous()
    n = int(st.shape[0])
    M = int(st.numel() // max(n, 1))
    out = torch.empty(tuple(st.shape[1:]), dtype=torch.float32, device=st.device)
    check(_lib.lib().om_topk_mean(_need(st, "st"), n, M, int(k), out.data_ptr(), _stream()), "topk_mean")
    return out


def argminmax0(a: torch.Tensor, want_max: bool):
    """(index of the first min / max along axis 0 (NumPy's NaN rule), the value there)."""
    a = a.contiguous()
    n = int(a.shape[0])
    M = int(a.numel() // max(n, 1))
    idx = torch.empty(tuple(a.shape[1:]), dtype=torch.int64, device=a.device)
    val = torch.empty(tuple(a.shape[1:]), dtype=torch.float32, device=a.device)
    check(_lib.lib().om_argminmax0(_need(a, "a"), n, M, int(bool(want_max)), idx.data_ptr(), val.data_ptr(),
                                   _stream()), "argmin0/argmax0")
    return idx, val


def sgm(cost: torch.Tensor, P1: float, P2: float, workers: int = 1) -> torch.Tensor:
    """_mvs.sgm (_mvs.pyx:98-250): semi-global aggregation of a (D, H, W) float32 cost volume, or of
    a batch (B, D, H, W) in one call, along the 8 path directions; bit-identical to the reference.
    `workers` is accepted for the reference's signature (the device runs every path at once)."""
    if cost.dim() not in (3, 4):
        raise ValueError(f"sgm: cost must be (D, H, W) or (B, D, H, W), got {tuple(cost.shape)}")
    c = cost.contiguous()
    if c.dtype != torch.float32:
        raise ValueError("sgm: float32 cost expected")
    B = int(c.shape[0]) if c.dim() == 4 else 1
    D, H, W = (int(v) for v in c.shape[-3:])
    agg = torch.empty_like(c)
    if c.numel() == 0:
        return agg.zero_()
    L = _lib.lib()
    nb = L.om_sgm_workspace_bytes(B, D, H, W)
    ws = _workspace(nb, c.device)
    check(L.om_sgm(_need(c, "cost"), B, D, H, W, float(P1), float(P2), agg.data_ptr(), ws.data_ptr(), nb, _stream()),
          "sgm")
    return agg
