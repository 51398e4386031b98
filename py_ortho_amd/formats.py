"""The reference's depth-map file set and dense cloud, written and read from device results.

densify keeps three files per view after pass 1 (densify.py:590-592): p{k:05d}.npy (depth,
float32), s{k:05d}.npy (score, float16) and t{k:05d}.npy (low texture, bool); pass 2 loads them
(:607, :613-614), repairs, filters D[(S < min_ncc) | low] = NaN with S the re-widened float16
score (:641) and keeps {k:05d}.npy (:639), which fusion reads (:656).  The fused result is a
DenseCloud (:84-90).  This module produces the same files from device tensors -- the float16
conversion and the filter run on the device, bit-identical to NumPy's -- and reads them back onto
the device, so callers that keep depth maps (DepthOptions.keep_depthmaps) see the same bytes.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check
from .dense import _stream


@dataclass
class DenseCloud:
    """densify.DenseCloud (densify.py:84-90)."""
    xyz: np.ndarray        # (n, 3) float64, local frame
    rgb: np.ndarray        # (n, 3) uint8
    views: np.ndarray      # (n,) uint8: images that agree on the point
    cam: np.ndarray        # (n,) int32: reference camera (index into rec.used)
    spacing: float         # typical ground spacing of the points (m)


def dense_cloud(pts: Sequence[torch.Tensor], cols: Sequence[torch.Tensor], nv: Sequence[torch.Tensor],
                owner: Sequence[torch.Tensor], spacing: float) -> DenseCloud:
    """The DenseCloud of fusion's outputs (fusion.fuse's lists; densify.py:664-666)."""
    cat = [torch.cat(list(x)).cpu().numpy() if len(x) else None for x in (pts, cols, nv, owner)]
    if cat[0] is None:
        raise RuntimeError("Densification produced no consistent points: too little overlap or texture "
                           "(try min_views=2 or a larger max_image_dim)")
    xyz, rgb, views, cam = cat
    if rgb is None:
        rgb = np.zeros((len(xyz), 3), np.uint8)
    return DenseCloud(np.ascontiguousarray(xyz, np.float64), np.ascontiguousarray(rgb, np.uint8),
                      np.ascontiguousarray(views, np.uint8), np.ascontiguousarray(cam, np.int32), float(spacing))


def to_f16(x: torch.Tensor) -> torch.Tensor:
    """x.astype(np.float16) on the device, bit for bit (NumPy's rounding and NaN rule)."""
    x = x.contiguous()
    if x.dtype != torch.float32 or not x.is_cuda:
        raise ValueError("to_f16: float32 device tensor expected")
    out = torch.empty(x.shape, dtype=torch.float16, device=x.device)
    check(_lib.lib().om_f32_to_f16(x.data_ptr(), x.numel(), out.data_ptr(), _stream()), "to_f16")
    return out


def to_f32(h: torch.Tensor) -> torch.Tensor:
    """h.astype(np.float32) on the device (exact)."""
    h = h.contiguous()
    if h.dtype != torch.float16 or not h.is_cuda:
        raise ValueError("to_f32: float16 device tensor expected")
    out = torch.empty(h.shape, dtype=torch.float32, device=h.device)
    check(_lib.lib().om_f16_to_f32(h.data_ptr(), h.numel(), out.data_ptr(), _stream()), "to_f32")
    return out


def filter_depth(depth: torch.Tensor, score16: torch.Tensor, low, min_ncc: float) -> torch.Tensor:
    """depth[(score16.astype(float32) < min_ncc) | low] = NaN in place (densify.py:641)."""
    if depth.dtype != torch.float32 or not depth.is_contiguous() or score16.dtype != torch.float16:
        raise ValueError("filter_depth: contiguous float32 depth and float16 score expected")
    if score16.shape != depth.shape or (low is not None and low.shape != depth.shape):
        raise ValueError("filter_depth: shapes differ")
    s = score16.contiguous()
    lo = None if low is None else low.to(torch.uint8).contiguous()
    check(_lib.lib().om_filter_depth(depth.data_ptr(), s.data_ptr(), lo.data_ptr() if lo is not None else None,
                                     depth.numel(), float(min_ncc), _stream()), "filter_depth")
    return depth


def save_pass1(wd: str, k: int, depth: torch.Tensor, score: torch.Tensor, low: torch.Tensor) -> None:
    """The three files of densify's job(k) (densify.py:590-592)."""
    np.save(os.path.join(wd, f"p{k:05d}.npy"), depth.cpu().numpy())
    np.save(os.path.join(wd, f"s{k:05d}.npy"), to_f16(score).cpu().numpy())
    np.save(os.path.join(wd, f"t{k:05d}.npy"), low.to(torch.bool).cpu().numpy())


def load_pass1(wd: str, k: int, device="cuda"):
    """(depth float32, score float16, low bool) of view k on the device (densify.py:607, :613-614)."""
    d = torch.from_numpy(np.load(os.path.join(wd, f"p{k:05d}.npy"))).to(device)
    s = torch.from_numpy(np.load(os.path.join(wd, f"s{k:05d}.npy"))).to(device)
    t = torch.from_numpy(np.load(os.path.join(wd, f"t{k:05d}.npy"))).to(device)
    return d, s, t


def save_final(wd: str, k: int, depth: torch.Tensor) -> None:
    """The filtered depth map fusion reads (densify.py:639)."""
    np.save(os.path.join(wd, f"{k:05d}.npy"), depth.cpu().numpy())


def load_final(wd: str, k: int, device="cuda") -> torch.Tensor:
    return torch.from_numpy(np.load(os.path.join(wd, f"{k:05d}.npy"))).to(device)
