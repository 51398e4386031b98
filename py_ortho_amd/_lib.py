"""ctypes binding of libom355.so (include/om355.h).

The shared library is the only compute path: if it is missing or cannot be loaded, every entry
point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libom355.so")

OM_OK, OM_EINVAL, OM_ESHAPE, OM_ENOMEM, OM_EHIP, OM_ENSRC, OM_EUNSUP = 0, -1, -2, -3, -4, -5, -6
OM_HYP_MAPS, OM_HYP_LINEAR, OM_HYP_REFINE, OM_HYP_CONST = 0, 1, 2, 3
OM_FMT_F16, OM_FMT_BF16, OM_FMT_F32 = 1, 2, 3

c_void_p = ctypes.c_void_p
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_size_t = ctypes.c_size_t
c_double = ctypes.c_double
c_float_p = ctypes.POINTER(ctypes.c_float)


class OmSrc(ctypes.Structure):
    _fields_ = [("img", c_void_p), ("h", c_int32), ("w", c_int32), ("geo", c_double * 18)]


class OmSweepDesc(ctypes.Structure):
    _fields_ = [
        ("ref", c_void_p), ("mu", c_void_p), ("sd", c_void_p), ("ray", c_void_p),
        ("H", c_int32), ("W", c_int32),
        ("n_src", c_int32), ("src", ctypes.POINTER(OmSrc)),
        ("r", c_int32), ("top_k", c_int32),
        ("hyp_mode", c_int32), ("hyp_a", c_void_p), ("hyp_b", c_void_p), ("hyp_offset", c_int32),
        ("i0", c_int32), ("nh", c_int32),
        ("best", c_void_p), ("prev", c_void_p), ("s_prev_best", c_void_p), ("s_next_best", c_void_p),
        ("idx", c_void_p),
        ("workspace", c_void_p), ("workspace_bytes", c_size_t),
        ("chunks", c_int32), ("tile_h", c_int32),
    ]


class OmPmView(ctypes.Structure):
    _fields_ = [
        ("ref", c_void_p), ("H", c_int32), ("W", c_int32), ("ray", c_void_p), ("Cr", c_double * 3),
        ("S", c_int32), ("imgs", c_void_p * 16), ("deps", c_void_p * 16), ("ih", c_int32 * 16), ("iw", c_int32 * 16),
        ("par", (c_double * 18) * 16), ("depth", c_void_p), ("normal", c_void_p), ("score", c_void_p),
        ("active", c_void_p), ("geo_tol", c_double), ("dmin", c_double), ("dmax", c_double),
        ("seed", ctypes.c_uint32), ("depth_from", c_void_p), ("score_from", c_void_p),
    ]


class OmCam(ctypes.Structure):
    _fields_ = [("R", c_double * 9), ("C", c_double * 3), ("f", c_double), ("k1", c_double), ("k2", c_double),
                ("k3", c_double), ("cx", c_double), ("cy", c_double), ("w", c_int32), ("h", c_int32)]


class OmFuseView(ctypes.Structure):
    _fields_ = [("depth", c_void_p), ("rgb", c_void_p), ("used", c_void_p), ("cam", OmCam)]


# name -> (restype, argtypes); every symbol declared in include/om355.h
SIGNATURES = {
    "om_version": (c_int32, []),
    "om_last_error": (ctypes.c_char_p, []),
    "om_profile_enable": (c_int32, [c_int32]),
    "om_profile_read": (c_int32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_int64)]),
    "om_device_info": (c_int32, [ctypes.c_char_p, c_int32, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32),
                                 ctypes.POINTER(c_int64)]),
    "om_sweep_workspace_bytes": (c_size_t, [ctypes.POINTER(OmSweepDesc)]),
    "om_sweep": (c_int32, [ctypes.POINTER(OmSweepDesc), c_void_p]),
    "om_depth_band_workspace_bytes": (c_size_t, [c_int32, c_int32, c_int32]),
    "om_depth_band": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32, c_int32, c_double, c_double, c_double,
                                c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "om_sweep_rows_workspace_bytes": (c_size_t, [ctypes.POINTER(OmSweepDesc)]),
    "om_sweep_rows": (c_int32, [ctypes.POINTER(OmSweepDesc), c_void_p]),
    "om_sweep_noalloc_max_bytes": (c_size_t, []),
    "om_sweep_noalloc": (c_int32, [ctypes.POINTER(OmSweepDesc), c_void_p]),
    "om_sweep_hypothesis": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                      ctypes.POINTER(c_void_p), ctypes.POINTER(c_int32), c_void_p,
                                      ctypes.POINTER(c_double), c_void_p, c_int32, c_int32, c_int32,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "om_sweep_finalize": (c_int32, [ctypes.POINTER(OmSweepDesc), c_void_p, c_void_p]),
    "om_pm_workspace_bytes": (c_size_t, [c_int32, c_int32]),
    "om_pm_refine": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, ctypes.POINTER(c_double), c_int32,
                               ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), ctypes.POINTER(c_int32),
                               ctypes.POINTER(c_int32), ctypes.POINTER(c_double), c_void_p, c_void_p, c_void_p,
                               c_void_p, c_int32, c_int32, c_int32, c_double, c_double, c_int32, c_double,
                               c_double, ctypes.c_uint32, c_void_p, c_size_t, ctypes.POINTER(c_int64), c_void_p]),
    "om_hyp_linear": (c_int32, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "om_hyp_refine": (c_int32, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "om_box_workspace_bytes": (c_size_t, [c_int32, c_int32, c_int32]),
    "om_box_zero_mean": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "om_ref_stats": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_size_t,
                               c_void_p]),
    "om_half": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p]),
    "om_rays_world": (c_int32, [c_int32, c_int32, ctypes.POINTER(c_double), c_double, c_double, c_double, c_double,
                                c_double, c_double, c_void_p, c_void_p]),
    "om_sample_view": (c_int32, [c_void_p, c_int32, c_int32, c_int32, ctypes.POINTER(c_double),
                                 ctypes.POINTER(c_double), c_double, c_double, c_double, c_double, c_double, c_double,
                                 c_double, c_double, c_double, c_void_p, c_int64, c_int32, c_int32, c_void_p, c_void_p,
                                 c_void_p]),
    "om_box_sat_workspace_bytes": (c_size_t, [c_int64, c_int32, c_int32, c_int32]),
    "om_box_sat_mean": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_size_t,
                                  c_void_p]),
    "om_topk_mean": (c_int32, [c_void_p, c_int32, c_int64, c_int32, c_void_p, c_void_p]),
    "om_argminmax0": (c_int32, [c_void_p, c_int32, c_int64, c_int32, c_void_p, c_void_p, c_void_p]),
    "om_fp64_steps": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p]),
    "om_batch_create": (c_void_p, []),
    "om_batch_destroy": (c_int32, [c_void_p]),
    "om_batch_set_sweep": (c_int32, [c_void_p, ctypes.POINTER(OmSweepDesc), c_int32, ctypes.POINTER(c_void_p),
                                     c_int32, c_int32, c_int32]),
    "om_batch_set_pm": (c_int32, [c_void_p, ctypes.POINTER(OmPmView), c_int32, c_int32, c_int32, c_int32, c_double,
                                  c_int32, c_int32]),
    "om_batch_run": (c_int32, [c_void_p, c_void_p]),
    "om_batch_capture": (c_int32, [c_void_p, c_void_p]),
    "om_batch_replay": (c_int32, [c_void_p, c_void_p]),
    "om_batch_pm_changed": (c_int32, [c_void_p, ctypes.POINTER(c_int64)]),
    "om_batch_pm_evaluated": (c_int32, [c_void_p, ctypes.POINTER(c_int64)]),
    "om_init_normals": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p]),
    "om_to_lp": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "om_pm_workspace_counts": (c_int32, [c_void_p, c_int32, c_int32, ctypes.POINTER(c_int64)]),
    "om_pm_refine_lp": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, ctypes.POINTER(c_double), c_int32,
                                  ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), ctypes.POINTER(c_int32),
                                  ctypes.POINTER(c_int32), ctypes.POINTER(c_double), c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_int32, c_int32, c_int32, c_double, c_double, c_int32, c_double,
                                  c_double, ctypes.c_uint32, c_void_p, c_size_t, ctypes.POINTER(c_int64), c_int32,
                                  c_void_p]),
    "om_sweep_lp": (c_int32, [ctypes.POINTER(OmSweepDesc), c_int32, c_void_p, c_void_p]),
    "om_sweep_multi_bytes": (c_size_t, [c_int32]),
    "om_sweep_multi": (c_int32, [ctypes.POINTER(OmSweepDesc), c_int32, ctypes.POINTER(c_void_p), c_int32, c_void_p,
                                 c_size_t, c_void_p, c_void_p]),
    "om_consistency": (c_int32, [c_void_p, ctypes.POINTER(OmCam), c_int32, ctypes.POINTER(c_void_p),
                                 ctypes.POINTER(OmCam), ctypes.POINTER(c_double), c_double, c_void_p, c_void_p]),
    "om_f32_to_f16": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p]),
    "om_f16_to_f32": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p]),
    "om_filter_depth": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_double, c_void_p]),
    "om_depth_view": (c_int32, [c_void_p, c_int32, c_int32, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), c_void_p, c_void_p, c_void_p, c_void_p]),
    "om_sgm_workspace_bytes": (c_size_t, [c_int64, c_int32, c_int32, c_int32]),
    "om_sgm": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32, ctypes.c_float, ctypes.c_float, c_void_p,
                         c_void_p, c_size_t, c_void_p]),
    "om_fuse_workspace_bytes": (c_size_t, [ctypes.POINTER(OmFuseView), c_int32, c_int32]),
    "om_fuse": (c_int32, [c_int32, ctypes.POINTER(c_int32), c_int32, ctypes.POINTER(OmFuseView),
                          ctypes.POINTER(c_int32), c_int32, ctypes.POINTER(c_double), c_double, c_int32, c_int32,
                          c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_size_t, c_void_p]),
}

_lib = None
_lock = threading.Lock()


class NativeLibraryMissing(ImportError):
    pass


def lib():
    """The loaded libom355.so (raises NativeLibraryMissing when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NativeLibraryMissing(
                    f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                    "(hipcc --offload-arch=gfx950)")
            h = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
            _lib = h
    return _lib


def last_error() -> str:
    msg = lib().om_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(rc: int, what: str) -> None:
    """Map an OM_E* code to the exception type the reference raises for the same condition."""
    if rc == OM_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc in (OM_EINVAL, OM_ESHAPE, OM_ENSRC):
        raise ValueError(msg)
    if rc == OM_ENOMEM:
        raise MemoryError(msg)
    if rc == OM_EUNSUP:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)
