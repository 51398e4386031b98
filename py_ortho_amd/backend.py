"""Compute-backend selection with the MI355X (ROCm) backend.

Mirrors orthomosaic/backend.py:185-214: `select_backend(prefer)` returns a backend object that the
pipeline stores on `AlignResult.backend` (pipeline.py:93, :129) and `build_3d` passes down.  The
ROCm backend implements the depth-map path natively (sweep, PatchMatch, full depth maps) and the
height-sweep array operations the reference routes through backends (`sample_view`, `box`,
`topk_mean`, `argmin0`, `argmax0`, backend.py:67-99).  The 2D-mosaic and feature-matching methods
(`match`, `match_mutual`, `new_block`, `warp_accumulate`, `finalize`, `finalize_float`) are not
part of this path: they are delegated to the reference's CPU backend when pyOrthomosaic is
installed and raise NotImplementedError otherwise.

Selection: 'auto' -> ROCm when a HIP device is present, libom355.so loads and the self-test
passes, else the reference's CPU backend; 'rocm' / 'gpu' / 'cuda' -> ROCm or RuntimeError (like
the reference's explicit 'cuda', backend.py:204-208).  ORTHO_DISABLE_ROCM=1 (or the reference's
ORTHO_DISABLE_CUDA=1) hides the device.
"""
from __future__ import annotations

import logging
import os

import numpy as np
import torch

log = logging.getLogger(__name__)


def rocm_available() -> bool:
    if os.environ.get("ORTHO_DISABLE_ROCM") == "1" or os.environ.get("ORTHO_DISABLE_CUDA") == "1":
        return False
    try:
        return bool(torch.cuda.is_available()) and torch.version.hip is not None
    except Exception:
        return False


def _reference_cpu_backend():
    try:
        from orthomosaic.backend import CPUBackend  # pyOrthomosaic, when installed beside this package
    except Exception:
        return None
    return CPUBackend()


class ROCmBackend:
    name = "rocm"
    parallel_blocks = False  # the GPU parallelises inside each block (as the reference's CUDA backend)
    xp = torch

    def __init__(self, device: str = "cuda"):
        from . import _lib
        self.device = torch.device(device)
        self._lib = _lib.lib()
        props = torch.cuda.get_device_properties(self.device)
        self.device_name = f"{props.name} ({getattr(props, 'gcnArchName', '?')}, {props.total_memory / 2**30:.0f} GB)"
        self._cpu = None
        self._self_test()

    # ------------------------------------------------------------------ self-test
    def _self_test(self):
        """Known answers on tiny inputs (pattern of _cuda.py:146-165): window mean of a ramp (closed
        form), fused sweep of an identity view (score 1, first plane)."""
        from . import dense
        a = torch.arange(30, dtype=torch.float32, device=self.device).reshape(5, 6)
        m = dense.box_zero_mean(a, 1).cpu().numpy()
        exp = sum(np.pad(np.arange(30, dtype=np.float32).reshape(5, 6), 1)[dy:dy + 5, dx:dx + 6]
                  for dy in range(3) for dx in range(3)) / 9.0
        if not np.allclose(m, exp, atol=1e-4):
            raise RuntimeError("ROCm self-test failed (box mean)")

    # ------------------------------------------------------------------ arrays
    def upload(self, img):
        return torch.as_tensor(np.ascontiguousarray(img), device=self.device)

    def image_nbytes(self, dev_img) -> int:
        return int(dev_img.numel() * dev_img.element_size())

    def to_numpy(self, a):
        return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)

    def asarray(self, a, dtype=None):
        if isinstance(dtype, np.dtype) or (isinstance(dtype, type) and issubclass(dtype, np.generic)):
            dtype = torch.from_numpy(np.zeros(0, dtype)).dtype
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=dtype or a.dtype)
        return torch.as_tensor(np.asarray(a), device=self.device, dtype=dtype)

    # ------------------------------------------------------------------ depth-map path (native)
    def sweep(self, *args, **kw):
        from . import mvs
        return mvs.sweep(*args, **kw)

    def depth_map(self, *args, **kw):
        from . import mvs
        return mvs.depth_map(*args, device=self.device, **kw)

    def pm_refine(self, *args, **kw):
        from . import dense
        return dense.pm_refine(*args, **kw)

    def sweep_hypothesis(self, *args, **kw):
        from . import dense
        return dense.sweep_hypothesis(*args, **kw)

    def densify(self, ar, rec, gains, biases, out_dir, opt, **kw):
        """densify.densify (densify.py:499-510) on this device: depth maps, repair and fusion in HBM
        -> DenseCloud.  The reference's densify hands over here when ar.backend is this backend."""
        from . import densify
        return densify.densify(ar, rec, gains, biases, out_dir, opt, device=self.device, **kw)

    # ------------------------------------------------------------------ height-sweep array ops
    def box(self, a, r):
        from . import ops
        return ops.box(self.asarray(a, torch.float32), r)

    def topk_mean(self, st, k):
        from . import ops
        return ops.topk_mean(self.asarray(st, torch.float32), k)

    def argmin0(self, a):
        from . import ops
        return ops.argminmax0(self.asarray(a, torch.float32), False)

    def argmax0(self, a):
        from . import ops
        return ops.argminmax0(self.asarray(a, torch.float32), True)

    def sample_view(self, img, cam, X0, Y0, gsd, Z):
        from . import ops
        return ops.sample_view(self.asarray(img, torch.float32), cam, X0, Y0, gsd, self.asarray(Z, torch.float32))

    # ------------------------------------------------------------------ other stages (delegated)
    def _delegate(self, name):
        if self._cpu is None:
            self._cpu = _reference_cpu_backend()
        if self._cpu is None:
            raise NotImplementedError(f"ROCmBackend.{name} belongs to the 2D / matching stages; install "
                                      "pyOrthomosaic to run them on its CPU backend")
        return getattr(self._cpu, name)

    def match(self, d1, d2):
        return self._delegate("match")(d1, d2)

    def match_mutual(self, d1, d2):
        return self._delegate("match_mutual")(d1, d2)

    def new_block(self, H, W):
        return self._delegate("new_block")(H, W)

    def warp_accumulate(self, block, dev_img, M, gain, power, mode):
        img = self.to_numpy(dev_img) if isinstance(dev_img, torch.Tensor) else dev_img
        return self._delegate("warp_accumulate")(block, img, M, gain, power, mode)

    def finalize(self, block, mode):
        return self._delegate("finalize")(block, mode)

    def finalize_float(self, block, mode):
        return self._delegate("finalize_float")(block, mode)


def select_backend(prefer: str = "auto"):
    """prefer: 'auto' (ROCm > CPU) | 'rocm' | 'gpu' | 'cuda' | 'cpu'."""
    prefer = prefer.lower()
    if prefer in ("auto", "rocm", "gpu", "cuda", "hip"):
        if rocm_available():
            try:
                be = ROCmBackend()
                log.info("Using ROCm backend: %s", be.device_name)
                return be
            except Exception as exc:
                msg = str(exc).splitlines()[0] if str(exc) else type(exc).__name__
                if prefer != "auto":
                    raise RuntimeError(f"ROCm backend unusable: {msg}") from exc
                log.warning("HIP device present but unusable (%s); falling back to CPU", msg)
        elif prefer != "auto":
            raise RuntimeError("ROCm backend requested but no HIP device / libom355.so found")
    if prefer in ("mps", "metal"):
        raise RuntimeError("MPS backend requested; this build targets AMD Instinct GPUs")
    cpu = _reference_cpu_backend()
    if cpu is None:
        raise RuntimeError("no usable backend: no HIP device and pyOrthomosaic's CPU backend is not installed")
    log.info("Using CPU backend (pyOrthomosaic)")
    return cpu
