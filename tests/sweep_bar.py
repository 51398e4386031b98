"""The plane sweep's parity bar against the oracle.

The device forms each window's sums over short runs where the reference keeps running sums along
whole rows and columns (_sweep.h:96-113, :139-146).  Both equal the exact sum whenever no FP64
partial sum rounds, so the results are normally bit-identical; where one does round, a score can
move by an FP64 rounding and a near-tie between two planes can resolve the other way.  The bar:

- every hypothesis's combined NCC score within TOL = 1e-4 of the oracle's (checked on the best
  score, on the last hypothesis's score `prev`, and on the parabola neighbours `s_prev_best` /
  `s_next_best` wherever the winner agrees);
- the winner-take-all index identical, except at pixels whose two best oracle scores differ by
  less than TOL; there the device's plane must score within TOL of the oracle's best;
- the validity mask (best score >= min_ncc) identical, except where the oracle's best lies within
  TOL of min_ncc;
- NaN scores at the same pixels.
Bit-identity is reported as a separate statistic, not required.
"""
from __future__ import annotations

import numpy as np

import oracle

TOL = 1e-4
MIN_NCC = 0.3  # DepthOptions.min_ncc (densify.py:57-81)


def oracle_scores(ref, mu, sd, srcs, ray, geo, hyps, r, top_k):
    """The oracle's sweep over `hyps` (a list of (H, W) inverse-depth maps) -> ((best, prev, spb,
    snb), idx, S) with S[i] the combined score of hypothesis i (its `prev` right after it)."""
    H, W = ref.shape
    st, idx = oracle.new_state(H, W)
    S = np.empty((len(hyps), H, W), np.float32)
    for i, inv in enumerate(hyps):
        oracle.sweep_hypothesis(ref, mu, sd, srcs, ray, geo, inv, r, top_k, i, st[0], st[1], st[2], st[3], idx)
        S[i] = st[1]
    return st, idx, S


def _bits_equal(a, b):
    return a.view(np.uint32) == b.view(np.uint32)


def check(S, st0, idx0, st, idx, label, tol=TOL, min_ncc=MIN_NCC) -> dict:
    """Asserts the bar (module docstring) for device state (st = [best, prev, spb, snb], idx) against
    the oracle's (st0, idx0) and its score volume S; returns statistics of the comparison."""
    best0, prev0, spb0, snb0 = st0
    best, prev, spb, snb = st
    stats = {"label": str(label), "pixels": int(best0.size),
             "bit_identical_best": float(np.mean(_bits_equal(best, best0))),
             "bit_identical_state": bool(all(_bits_equal(a, b).all() for a, b in zip(st, st0))
                                         and np.array_equal(idx, idx0))}
    if stats["bit_identical_state"]:
        stats.update(idx_mismatch=0, near_ties=0, max_abs_dbest=0.0)
        return stats
    nan0, nan1 = np.isnan(best0), np.isnan(best)
    assert np.array_equal(nan0, nan1), (label, "NaN scores at different pixels")
    fin = ~nan0
    d = np.abs(best[fin].astype(np.float64) - best0[fin])
    stats["max_abs_dbest"] = float(d.max()) if d.size else 0.0
    assert stats["max_abs_dbest"] <= tol, (label, "best score", stats)
    dp = np.abs(prev.astype(np.float64) - prev0)
    assert np.nanmax(dp, initial=0.0) <= tol and np.array_equal(np.isnan(prev), np.isnan(prev0)), (label, "prev")
    # winner index, up to near-ties of the oracle's two best scores
    D = S.shape[0]
    if D >= 2:
        Sf = np.where(np.isnan(S), -np.inf, S)
        top2 = -np.partition(-Sf, 1, axis=0)[:2]
        gap = top2[0].astype(np.float64) - top2[1]
    else:
        gap = np.full(best0.shape, np.inf)
    tie = gap < tol
    diff = idx != idx0
    stats["near_ties"] = int(tie.sum())
    stats["idx_mismatch"] = int(diff.sum())
    assert not (diff & ~tie).any(), (label, "winner differs away from a near-tie", stats)
    if diff.any():
        yy, xx = np.nonzero(diff)
        got = S[np.clip(idx[yy, xx], 0, D - 1), yy, xx].astype(np.float64)
        assert (idx[yy, xx] >= 0).all() and (got >= best0[yy, xx] - tol).all(), (label, "device winner too poor")
    same = ~diff
    for a, b, name in ((spb, spb0, "s_prev_best"), (snb, snb0, "s_next_best")):
        da = np.abs(a[same].astype(np.float64) - b[same])
        assert np.nanmax(da, initial=0.0) <= tol, (label, name)
    # validity mask
    v0, v1 = best0 >= min_ncc, best >= min_ncc
    edge = np.abs(best0.astype(np.float64) - min_ncc) < tol
    assert not ((v0 != v1) & ~edge).any(), (label, "validity mask")
    stats["validity_edge"] = int(edge.sum())
    return stats
