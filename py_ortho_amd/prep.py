"""View preparation on the device: the matching image of a decoded view and its pyramid.

- depth_view(rgb, gain, bias, mask) is _fast.depth_view (_fast.pyx:577-616), which densify's
  load_view runs on every decoded image (densify.py:528-535): grey (h, w) float32 with NaN where
  the mask is False, and the uint8 colour image (om_depth_view).
- prepare_view(rgb, gain, bias, mask, levels) uploads a decoded uint8 image (3 bytes a pixel
  over PCIe instead of a 4-byte grey plus the colours), runs depth_view on the device and builds
  the coarse-to-fine pyramid of densify.depth_map (_half, densify.py:127-130, :384-386).

Decode itself stays on the CPU (Pillow, imageio.load_rgb).  Results are bit-identical to the
reference; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._lib import check
from .dense import _stream, half


def _f3(v) -> Optional[ctypes.Array]:
    if v is None:
        return None
    a = np.asarray(v, np.float32).reshape(-1)
    if a.size != 3:
        raise ValueError("gain / bias must hold 3 values (one per channel)")
    return (ctypes.c_float * 3)(*[float(x) for x in a])


def depth_view(rgb: torch.Tensor, gain=None, bias=None, mask: Optional[torch.Tensor] = None):
    """_fast.depth_view on the device: rgb (h, w, 3) uint8 -> (grey (h, w) float32, rgb8 (h, w, 3) uint8)."""
    if not (isinstance(rgb, torch.Tensor) and rgb.is_cuda and rgb.dtype == torch.uint8):
        raise ValueError("depth_view: rgb must be a uint8 device tensor")
    if rgb.dim() != 3 or rgb.shape[2] != 3:
        raise ValueError("depth_view expects (h, w, 3) RGB")
    rgb = rgb.contiguous()
    h, w = int(rgb.shape[0]), int(rgb.shape[1])
    m = None
    if mask is not None:
        if tuple(mask.shape) != (h, w):
            raise ValueError("mask shape differs from the image")
        m = mask.to(device=rgb.device, dtype=torch.uint8).contiguous()
    grey = torch.empty((h, w), dtype=torch.float32, device=rgb.device)
    rgb8 = torch.empty((h, w, 3), dtype=torch.uint8, device=rgb.device)
    check(_lib.lib().om_depth_view(rgb.data_ptr(), h, w, _f3(gain), _f3(bias), m.data_ptr() if m is not None else None,
                                   grey.data_ptr(), rgb8.data_ptr(), _stream()), "depth_view")
    return grey, rgb8


def prepare_view(rgb, gain=None, bias=None, mask=None, levels: int = 3, device="cuda"):
    """A decoded view made ready for matching: (grey, rgb8, pyramid) on the device, pyramid =
    [grey, half(grey), half(half(grey)), ...] (`levels` entries).  rgb / mask may be host arrays
    (uploaded here) or device tensors."""
    if isinstance(rgb, np.ndarray):
        rgb = torch.from_numpy(np.ascontiguousarray(rgb, np.uint8)).to(device, non_blocking=False)
    if isinstance(mask, np.ndarray):
        mask = torch.from_numpy(np.ascontiguousarray(mask, np.bool_).view(np.uint8)).to(device)
    grey, rgb8 = depth_view(rgb, gain, bias, mask)
    pyr = [grey]
    for _ in range(1, max(1, int(levels))):
        pyr.append(half(pyr[-1]))
    return grey, rgb8, pyr
