"""View preparation on the CPU: the oracle's depth_view (densify's NumPy path) against the
compiled reference's _fast.depth_view outputs (tests/golden/prep_views.npz), bit for bit."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "prep_views.npz")


def cases():
    z = np.load(GOLDEN)
    out = []
    for i in range(len([k for k in z.files if k.startswith("rgb_")])):
        def get(k):
            return z[f"{k}_{i}"] if f"{k}_{i}" in z.files else None
        out.append((i, z[f"rgb_{i}"], get("gain"), get("bias"), get("mask"), z[f"grey_{i}"], z[f"rgb8_{i}"]))
    return out


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("case", cases(), ids=lambda c: str(c[0]))
def test_oracle_depth_view_matches_reference(case):
    i, rgb, g, b, m, grey, rgb8 = case
    a, c = oracle.depth_view(rgb, g, b, m)
    assert _same(a, grey) and _same(c, rgb8)
    if m is not None:
        assert np.isnan(a[~m]).all() and np.isfinite(a[m]).all()
