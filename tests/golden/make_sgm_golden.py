"""Writes tests/golden/sgm_volumes.npz: vectors of the reference's semi-global matching
(_mvs.sgm, _mvs.pyx:98-250), the cost aggregation of the height sweep (mvs.py:325).

The reference's own test (test_native_parity.py:425-447) compares sgm with a Python loop on
small volumes; this script records the compiled reference's outputs on volumes of those shapes
(seeded costs in [1, 1.5) with 5 % zeros and a NaN at the first cell) and on two larger ones,
with 1 and 8 workers.  Only arrays are committed.  Needs the reference tree and Cython; run from
the repository root:

    python tests/golden/make_sgm_golden.py [/path/to/reference]
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_depthmap_golden import build_reference_package  # noqa: E402

SHAPES = ((7, 9, 11), (1, 4, 3), (12, 1, 6), (5, 6, 1), (24, 32, 40), (96, 20, 24))
P1, P2 = 0.05, 0.6


def volume(rng, D, H, W):
    cost = (1.0 + rng.uniform(0, 0.5, (D, H, W))).astype(np.float32)
    cost[rng.random(cost.shape) < 0.05] = 0.0
    cost[0, 0, 0] = np.nan
    return cost


def main(ref_root: str) -> None:
    with tempfile.TemporaryDirectory() as td:
        build_reference_package(ref_root, td)
        sys.path.insert(0, td)
        from orthomosaic import _mvs
        rng = np.random.default_rng(15)
        out = {"p": np.array([P1, P2], np.float32)}
        for i, (D, H, W) in enumerate(SHAPES):
            cost = volume(rng, D, H, W)
            a1 = _mvs.sgm(cost, P1, P2, 1)
            a8 = _mvs.sgm(cost, P1, P2, 8)
            assert np.array_equal(a1.view(np.uint32), a8.view(np.uint32))
            out[f"cost_{i}"] = cost
            out[f"agg_{i}"] = a1
        # a smooth cost with a cliff: aggregation should pick a surface, not noise
        D, H, W = 40, 24, 32
        z = (np.arange(W)[None, :] * 0.3 + np.arange(H)[:, None] * 0.2).astype(np.float32)
        z[:, W // 2:] += 10.0
        d = np.arange(D, dtype=np.float32)[:, None, None]
        cost = (np.minimum(np.abs(d - z[None]) * 0.2, 1.0) + rng.uniform(0, 0.3, (D, H, W))).astype(np.float32)
        out["cost_surface"] = cost
        out["agg_surface"] = _mvs.sgm(cost, P1, P2, 1)
        path = os.path.join(HERE, "sgm_volumes.npz")
        np.savez_compressed(path, **out)
        print(path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
