"""py_ortho_amd: MI355X-native (gfx950) per-view MVS depth maps for pyOrthomosaic.

Plane sweep with winner-take-all and red-black slanted-plane PatchMatch as hand-written HIP
kernels (libom355.so) behind the reference's own entry points:

- :mod:`py_ortho_amd.dense`   -- the `_dense` / `_fast` kernel API on device tensors
- :mod:`py_ortho_amd.mvs`     -- depth-map entry points (`DepthOptions`, `sweep`, `depth_map`, ...)
- :mod:`py_ortho_amd.backend` -- `select_backend` with the ROCm backend
"""
__version__ = "0.1.0"
