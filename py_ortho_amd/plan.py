"""What a depth map searches, planned from the block's sparse reconstruction.

The reference plans each depth map with four small steps over the tie points (densify.py:183-284):
the source views of every image (`_neighbours`), an image's depth range (`_depth_range`), a
per-block inverse-depth band at the coarse level (`_depth_band`) and the coarse hypothesis count
(`_num_depths`).  Here they are organised around one index of the reconstruction, `TiePoints`,
built once per block:

- the images that share tie points come from one sparse incidence product (images x points), and
  every image's candidates are ranked from one score matrix;
- every image's own observations are one slice of a camera-sorted index;
- the band runs on the device (`om_depth_band`): the host projects the image's own tie points into
  the coarse camera (NumPy float64, the reference's arithmetic), the device bins them into blocks
  with 64-bit atomic min / max, dilates over the 3x3 block neighbourhood and writes the per-pixel
  inverse-depth maps and the linear generator's step directly;
- the hypothesis count stays a host scalar: it sizes the launch, so computing it on the device
  would only add a device-to-host wait.

`rec` is any object with the fields of sfm.Reconstruction these use (sfm.py:35-61): `used`, `R`
(N, 3, 3) world->camera, `C` (N, 3), `X` (P, 3) tie points, `obs_cam` / `obs_pt` (one entry per
observation), `intr` (per camera group: f, cx, cy, width, height) and `cam_group`.
"""
from __future__ import annotations

import collections
import ctypes
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch
from scipy import sparse

from . import _lib
from ._lib import check

__all__ = ["TiePoints", "tie_points", "neighbours", "depth_range", "depth_band", "num_depths"]

_MIN_SHARED = 15          # tie points two images must share to be paired (densify.py:204)
_ANGLE_PREF = 8.0         # degrees of triangulation angle below which a pair is down-weighted
_ANGLE_STEEP = 45.0       # degrees above which a pair is down-weighted to 0.2
_BAND_MIN_OBS = 50        # own observations below which the band is the image-wide range


class TiePoints:
    """Per-block index of the tie points: camera-sorted observations, the shared-point counts
    between images, every image's median own-point depth, and the lowest ground."""

    def __init__(self, rec):
        self.rec = rec
        self.n = len(rec.used)
        cam = np.asarray(rec.obs_cam)
        pt = np.asarray(rec.obs_pt)
        by_cam = np.argsort(cam, kind="stable")          # observation order kept within an image
        self._own_pt = pt[by_cam]
        self._start = np.searchsorted(cam[by_cam], np.arange(self.n + 1))
        self._shared: Optional[np.ndarray] = None
        self._zmed: Optional[np.ndarray] = None
        self._zlow: Optional[float] = None

    def own(self, k: int) -> np.ndarray:
        """Tie-point indices image k observes, in observation order."""
        return self._own_pt[self._start[k]:self._start[k + 1]]

    def shared(self) -> np.ndarray:
        """(N, N) int: tie points two different images both observe."""
        if self._shared is None:
            cam = np.asarray(self.rec.obs_cam)
            inc = sparse.csr_matrix((np.ones(len(cam), np.int32), (cam, np.asarray(self.rec.obs_pt))),
                                    shape=(self.n, len(self.rec.X)))
            inc.data[:] = 1                              # an image observing a point twice counts once
            s = (inc @ inc.T).toarray().astype(np.int32)
            np.fill_diagonal(s, 0)
            self._shared = s
        return self._shared

    def median_depth(self) -> np.ndarray:
        """Per image the median camera-axis depth of its own tie points (1.0 without any)."""
        if self._zmed is None:
            z = np.ones(self.n)
            for k in range(self.n):
                idx = self.own(k)
                if len(idx):
                    cam_frame = self.rec.R[k] @ (self.rec.X[idx] - self.rec.C[k]).T
                    z[k] = np.median(cam_frame[2])
            self._zmed = z
        return self._zmed

    def lowest_ground(self) -> float:
        """Height of the lowest ground of the block (0.5th percentile of the tie-point heights)."""
        if self._zlow is None:
            self._zlow = float(np.percentile(self.rec.X[:, 2], 0.5))
        return self._zlow


_recent: "collections.OrderedDict" = collections.OrderedDict()   # id(rec) -> TiePoints, a few recent blocks
_recent_lock = threading.Lock()


def tie_points(rec) -> TiePoints:
    """The TiePoints index of `rec`, built once per reconstruction object (the last few are kept).
    The index assumes the reconstruction is not modified while it is in use."""
    with _recent_lock:
        tp = _recent.get(id(rec))
        if tp is not None and tp.rec is rec:
            _recent.move_to_end(id(rec))
            return tp
        tp = TiePoints(rec)
        _recent[id(rec)] = tp
        while len(_recent) > 4:
            _recent.popitem(last=False)
        return tp


def neighbours(rec, n_nb: int) -> List[List[int]]:
    """Source images of every image: up to n_nb images sharing at least 15 tie points, ranked by
    shared points weighted by the triangulation angle (quadratic below 8 degrees, 0.2 above 45),
    the angle being the baseline over the image's median tie-point depth (densify.py:183-205)."""
    tp = tie_points(rec)
    shared = tp.shared()
    C = np.asarray(rec.C, np.float64)
    base = np.linalg.norm(C[None, :, :] - C[:, None, :], axis=2)          # (N, N) baselines
    ang = np.degrees(base / np.maximum(tp.median_depth(), 1e-3)[:, None])
    weight = np.clip(ang / _ANGLE_PREF, 0, 1) ** 2 * np.where(ang > _ANGLE_STEEP, 0.2, 1.0)
    score = shared * weight
    out = []
    for k in range(tp.n):
        order = np.argsort(-score[k])
        keep = (score[k] > 0) & (shared[k] >= _MIN_SHARED)
        out.append([int(j) for j in order[keep[order]][:n_nb]])
    return out


def depth_range(rec, k: int) -> Optional[Tuple[float, float]]:
    """(near, far) depth search range of image k: the 1st-99th percentile depth of every tie point
    inside its view (not only its own), the far end reaching the block's lowest ground along the
    view axis, both widened by 20 % of the span (+1 m, +2 m); None with fewer than 10 points
    (densify.py:208-230)."""
    tp = tie_points(rec)
    it = rec.intr[rec.cam_group[k]]
    cam = (rec.X - rec.C[k]) @ rec.R[k].T
    z = cam[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        u = it.f * cam[:, 0] / z + it.cx
        v = it.f * cam[:, 1] / z + it.cy
    d = z[(z > 0) & (u >= 0) & (v >= 0) & (u < it.width) & (v < it.height)]
    if len(d) < 10:
        return None
    near, far = np.percentile(d, [1, 99])
    cos_nadir = max(-float(rec.R[k][2, 2]), 0.2)
    far = max(far, (float(rec.C[k][2]) - tp.lowest_ground()) / cos_nadir)
    span = far - near
    return max(0.3 * near, near - 0.2 * span - 1.0), far + 0.2 * span + 2.0


def _own_projections(rec, k: int, cam) -> Tuple[np.ndarray, int]:
    """Image k's own tie points in front of `cam`, projected: (n, 3) float64 (u, v, z), and the
    number of its observations."""
    idx = tie_points(rec).own(k)
    if len(idx) < _BAND_MIN_OBS:
        return np.zeros((0, 3)), len(idx)
    xc = (rec.X[idx] - cam.C) @ cam.R.T
    front = xc[:, 2] > 0
    u, v = cam.project_cam(xc[front, 0], xc[front, 1], xc[front, 2])
    return np.ascontiguousarray(np.stack([u, v, xc[front, 2]], 1), np.float64), len(idx)


def depth_band(rec, k: int, cam, shape, rng, block: int = 32, margin: float = 1.0, nd: int = 0,
               device="cuda"):
    """Per-pixel inverse-depth search band of image k at `shape` (densify._depth_band,
    densify.py:233-274), computed on the device from k's own tie points: per block of `block`
    pixels the depth span of its tie points (blocks with >= 5), dilated over the 3x3 block
    neighbourhood, widened by `margin` m and 5 % of the span, clipped to `rng`; elsewhere `rng`.
    -> (inv_lo, inv_hi) float32 (H, W) device tensors; with nd > 0 also the linear generator's
    step (inv_hi - inv_lo) / max(nd - 1, 1) (densify.py:411) as a third tensor."""
    H, W = int(shape[0]), int(shape[1])
    uvz, n_obs = _own_projections(rec, k, cam)
    inv_lo = torch.empty((H, W), dtype=torch.float32, device=device)
    inv_hi = torch.empty((H, W), dtype=torch.float32, device=device)
    step = torch.empty((H, W), dtype=torch.float32, device=device) if nd > 0 else None
    L = _lib.lib()
    nb = int(L.om_depth_band_workspace_bytes(H, W, int(block)))
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=device)
    pts = torch.from_numpy(uvz).to(device) if len(uvz) else None
    stream = torch.cuda.current_stream(device).cuda_stream
    check(L.om_depth_band(None if pts is None else pts.data_ptr(), len(uvz), n_obs, H, W, int(block),
                          float(rng[0]), float(rng[1]), float(margin), max(int(nd), 1), inv_lo.data_ptr(),
                          inv_hi.data_ptr(), None if step is None else step.data_ptr(), ws.data_ptr(), nb,
                          ctypes.c_void_p(stream)), "depth_band")
    del pts, ws                       # the caching allocator keeps them until the kernels ran
    return (inv_lo, inv_hi, step) if nd > 0 else (inv_lo, inv_hi)


def num_depths(rec, k: int, nbrs_k: Sequence[int], rng, f_coarse: float, base: int) -> int:
    """Coarse hypotheses: about one per pixel of coarse-level disparity over the range for the
    median baseline to the sources, between base and 3 base (densify.py:277-284)."""
    if not nbrs_k:
        return base
    C = np.asarray(rec.C, np.float64)
    baseline = float(np.median([np.linalg.norm(C[k] - C[j]) for j in nbrs_k]))
    disparity = f_coarse * baseline * (1.0 / rng[0] - 1.0 / rng[1])
    return int(np.clip(np.ceil(disparity), base, 3 * base))
