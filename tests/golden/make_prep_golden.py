"""Writes tests/golden/prep_views.npz: vectors of the reference's matching-image preparation
_fast.depth_view (_fast.pyx:577-616), called by densify's load_view (densify.py:528-535).

The reference's own test (test_native_parity.py:380-398) compares depth_view with densify's NumPy
path on seeded images; this script records the compiled function's outputs on images of those
shapes and on an odd-sized one, with and without gains, biases and masks.  Only arrays are
committed.  Needs the reference tree and Cython; run from the repository root:

    python tests/golden/make_prep_golden.py [/path/to/reference]
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_depthmap_golden import build_reference_package  # noqa: E402

CASES = (  # h, w, gain, bias, masked
    (29, 41, None, None, False),
    (29, 41, (1.07, 0.93, 1.21), (-3.5, 2.25, 0.0), True),
    (2, 2, None, None, False),
    (2, 2, (1.07, 0.93, 1.21), (-3.5, 2.25, 0.0), True),
    (121, 97, (0.81, 1.33, 1.02), None, True),
    (121, 97, None, (12.0, -40.5, 7.25), False),
)


def main(ref_root: str) -> None:
    with tempfile.TemporaryDirectory() as td:
        build_reference_package(ref_root, td)
        sys.path.insert(0, td)
        from orthomosaic import _fast
        rng = np.random.default_rng(13)
        out = {}
        for i, (h, w, g, b, masked) in enumerate(CASES):
            rgb = rng.integers(0, 256, (h, w, 3)).astype(np.uint8)
            m = rng.random((h, w)) < 0.7 if masked else None
            grey, rgb8 = _fast.depth_view(rgb, g, b, m)
            out[f"rgb_{i}"] = rgb
            if m is not None:
                out[f"mask_{i}"] = m
            if g is not None:
                out[f"gain_{i}"] = np.asarray(g, np.float32)
            if b is not None:
                out[f"bias_{i}"] = np.asarray(b, np.float32)
            out[f"grey_{i}"] = grey
            out[f"rgb8_{i}"] = rgb8
        path = os.path.join(HERE, "prep_views.npz")
        np.savez_compressed(path, **out)
        print(path, len(CASES), "cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
