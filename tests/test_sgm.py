"""Semi-global matching on the CPU: the oracle (oracle/om_sgm.c) against the compiled reference's
outputs (tests/golden/sgm_volumes.npz) bit for bit, and against a direct float32 evaluation of
the path recurrence on tiny volumes."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sgm_volumes.npz")
DIRS = ((0, 1), (0, -1), (1, 0), (-1, 0), (1, 1), (1, -1), (-1, 1), (-1, -1))


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def golden_cases():
    z = np.load(GOLDEN)
    P1, P2 = (float(v) for v in z["p"])
    return [(k[5:], z[k], z["agg_" + k[5:]], P1, P2) for k in z.files if k.startswith("cost_")]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_oracle_matches_reference(case):
    name, cost, agg, P1, P2 = case
    assert _same(oracle.sgm(cost, P1, P2), agg)


def _paths(cost, P1, P2):
    """Each direction's L by walking every path cell by cell from its entry (float32 scalars)."""
    D, H, W = cost.shape
    f = np.float32
    P1, P2 = f(P1), f(P2)
    out = []
    for dy, dx in DIRS:
        L = np.empty_like(cost)
        starts = [(y, x) for y in range(H) for x in range(W)
                  if not (0 <= y - dy < H and 0 <= x - dx < W)]
        for y, x in starts:
            prev = None
            while 0 <= y < H and 0 <= x < W:
                if prev is None:
                    L[:, y, x] = cost[:, y, x]
                else:
                    best = prev[0]
                    for d in range(1, D):
                        if prev[d] < best:
                            best = prev[d]
                    for d in range(D):
                        cand = [prev[d]]
                        if d > 0:
                            cand.append(f(prev[d - 1] + P1))
                        if d < D - 1:
                            cand.append(f(prev[d + 1] + P1))
                        cand.append(f(best + P2))
                        v = cand[0]
                        for c in cand[1:]:
                            if c < v:
                                v = c
                        L[d, y, x] = f(f(cost[d, y, x] + v) - best)
                prev = L[:, y, x].copy()
                y, x = y + dy, x + dx
        out.append(L)
    return out


@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 3, 3), (6, 1, 4), (4, 5, 1), (5, 3, 4)])
def test_oracle_matches_path_walk(shape):
    rng = np.random.default_rng(sum(shape))
    cost = (1.0 + rng.uniform(0, 0.5, shape)).astype(np.float32)
    cost[rng.random(shape) < 0.2] = 0.0
    cost.flat[int(rng.integers(cost.size))] = np.nan
    if cost.size > 3:
        cost.flat[1] = -0.0
    agg = np.zeros(shape, np.float32)
    for L in _paths(cost, 0.05, 0.6):
        agg = (agg + L).astype(np.float32)
    assert _same(oracle.sgm(cost, 0.05, 0.6), agg)


def test_oracle_aggregation_recovers_a_surface():
    """The smooth-with-cliff golden volume: the argmin of the aggregated cost follows the surface
    the noisy raw cost hides (a property of the method, not of the implementation)."""
    z = np.load(GOLDEN)
    cost = z["cost_surface"]
    D, H, W = cost.shape
    truth = (np.arange(W)[None, :] * 0.3 + np.arange(H)[:, None] * 0.2)
    truth[:, W // 2:] += 10.0
    raw = np.abs(np.argmin(cost, 0) - truth)
    agg = np.abs(np.argmin(oracle.sgm(cost, 0.05, 0.6), 0) - truth)
    assert (agg <= 1.0).mean() > 0.95 and (agg <= 1.0).mean() > (raw <= 1.0).mean()
