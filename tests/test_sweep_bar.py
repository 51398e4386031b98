"""The sweep parity bar itself (tests/sweep_bar.py), on the CPU: it accepts the oracle's own
result and results that differ only within its tolerance, and rejects planted mistakes."""
import numpy as np
import pytest

import cases
import sweep_bar


@pytest.fixture(scope="module")
def swept():
    rng = np.random.default_rng(31)
    ref, mu, sd, srcs, ray, geo, hyps = cases.sweep_case(rng, 40, 56, 4, 2, 9)
    return sweep_bar.oracle_scores(ref, mu, sd, srcs, ray, geo, hyps, 2, 2)


def _copy(orc):
    st0, idx0, S = orc
    return [a.copy() for a in st0], idx0.copy()


def test_bar_accepts_oracle_and_small_moves(swept):
    st0, idx0, S = swept
    st, idx = _copy(swept)
    assert sweep_bar.check(S, st0, idx0, st, idx, "self")["bit_identical_state"]
    fin = np.isfinite(st[0]) & (np.abs(st[0] - sweep_bar.MIN_NCC) > 1e-3)
    st[0][fin] += np.float32(3e-5)
    stats = sweep_bar.check(S, st0, idx0, st, idx, "moved")
    assert not stats["bit_identical_state"] and stats["max_abs_dbest"] < 1e-4


def test_bar_rejects_score_beyond_tolerance(swept):
    st0, idx0, S = swept
    st, idx = _copy(swept)
    y, x = np.argwhere(np.isfinite(st[0]))[0]
    st[0][y, x] += np.float32(2e-4)
    with pytest.raises(AssertionError):
        sweep_bar.check(S, st0, idx0, st, idx, "planted")


def test_bar_rejects_winner_change_away_from_ties(swept):
    st0, idx0, S = swept
    st, idx = _copy(swept)
    Sf = np.where(np.isnan(S), -np.inf, S)
    top2 = -np.partition(-Sf, 1, axis=0)[:2]
    clear = (top2[0] - top2[1] > 1e-2) & (idx0 >= 1)
    y, x = np.argwhere(clear)[0]
    idx[y, x] = idx0[y, x] - 1
    with pytest.raises(AssertionError):
        sweep_bar.check(S, st0, idx0, st, idx, "planted winner")


def test_bar_allows_winner_change_at_near_tie():
    S = np.array([[[0.5]], [[0.50004]], [[0.1]]], np.float32)
    st0 = [np.array([[0.50004]], np.float32), np.array([[0.1]], np.float32), np.array([[0.5]], np.float32),
           np.array([[0.1]], np.float32)]
    idx0 = np.array([[1]], np.int32)
    st = [np.array([[0.5]], np.float32), np.array([[0.1]], np.float32), np.array([[-2.0]], np.float32),
          np.array([[0.50004]], np.float32)]
    idx = np.array([[0]], np.int32)
    stats = sweep_bar.check(S, st0, idx0, st, idx, "tie")
    assert stats["idx_mismatch"] == 1 and stats["near_ties"] == 1
