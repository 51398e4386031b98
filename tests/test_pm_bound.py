"""The margin of the random refinement's geometric bound (pm_eps_view in pm.hip), checked on the
CPU: a photometric score computed with the reference's operations (_dense.pyx:520-541: window
sums in sample order, moments over the count, cv / sqrt(vi vj) when both variances reach 1e-8)
never exceeds 1 by more than the margin's bound, on windows built to round as badly as they can --
perfectly correlated samples with a large mean and a variance just above the 1e-8 guard.  A
planted bound ten thousand times smaller fails on the same windows."""
import math

import numpy as np

U = 2.0 ** -53


def margin_d(MI, MJ, r, step, kk):
    """The bound by which a photometric score can exceed 1 (before the final doubling of pm.hip)."""
    nw = 2 * r // step + 1
    m = nw * nw + 4.0
    gam = m * U / (1.0 - m * U)
    s = MI + MJ
    d = 2.0 * gam * 1.0000001e8 * s * s
    return d + (kk + 16) * U * (1.0 + d)


def ncc_reference(I, J):
    """One source's NCC in the reference's operation order (Python floats are IEEE doubles)."""
    sI = sJ = sII = sJJ = sIJ = 0.0
    for a, b in zip(I, J):
        sI += a
        sJ += b
        sII += a * a
        sJJ += b * b
        sIJ += a * b
    m = float(len(I))
    mi, mj = sI / m, sJ / m
    vi = sII / m - mi * mi
    vj = sJJ / m - mj * mj
    cv = sIJ / m - mi * mj
    if vi < 1e-8 or vj < 1e-8:
        return 0.0
    return cv / math.sqrt(vi * vj)


def _worst_excess(M, n_trials, rng):
    worst = 0.0
    for _ in range(n_trials):
        n = int(rng.integers(4, 37))
        sd = 1e-4 * rng.uniform(1.0, 1.6)                      # variance just above 1e-8
        I = (M + sd * rng.standard_normal(n)).astype(np.float32).astype(np.float64)
        alpha = rng.uniform(0.5, 2.0)
        J = alpha * (I - M) + M * rng.uniform(0.5, 1.0)       # exactly correlated, mean near M
        if np.var(I) < 1e-8 or np.var(J) < 1e-8:
            continue
        top = sorted((ncc_reference(I, J), ncc_reference(I, J * 1.0000001)), reverse=True)[:2]
        photo = (top[0] + top[1]) / 2
        worst = max(worst, photo - 1.0)
    return worst


def test_photometric_score_stays_within_the_margin():
    rng = np.random.default_rng(7)
    for M in (0.5, 1.0, 2.0, 4.0):
        d = margin_d(M, M, 5, 2, 2)
        worst = _worst_excess(M, 3000, rng)
        assert worst <= d, (M, worst, d)
        assert worst > 0.0 or M < 1.0  # the windows do round above 1


def test_a_planted_smaller_margin_fails():
    rng = np.random.default_rng(7)
    M = 4.0
    worst = _worst_excess(M, 3000, rng)
    assert worst > margin_d(M, M, 5, 2, 2) * 1e-4
