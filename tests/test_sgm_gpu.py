"""Semi-global matching on the device (om_sgm through ops.sgm): bit for bit against the compiled
reference's outputs and against the oracle for every lane layout (D up to 512), NaN / zero / -0 /
inf costs, thin volumes and batches."""
import numpy as np
import pytest
import torch

import oracle
from test_sgm import golden_cases

pytestmark = pytest.mark.gpu


def _same(a, b):
    """Bit-identical, except that any NaN equals any NaN: the reference's NaNs carry x86's payloads
    (an operand's, or 0xffc00000 from an invalid operation), the device's its own quiet NaN, and
    neither is part of the result's meaning."""
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    diff = int((a[~na & ~nb].view(np.uint32) != b[~na & ~nb].view(np.uint32)).sum())
    if diff or not np.array_equal(na, nb):
        print(f"  differing values {diff}, NaN positions differ {int((na != nb).sum())}", flush=True)
        return False
    return True


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_sgm_matches_reference(case):
    from py_ortho_amd import ops
    name, cost, agg, P1, P2 = case
    print(f"sgm golden {name} {cost.shape}", flush=True)
    out = ops.sgm(torch.from_numpy(cost).cuda(), P1, P2).cpu().numpy()
    assert _same(out, agg)


def _volume(rng, D, H, W):
    cost = (1.0 + rng.uniform(0, 0.5, (D, H, W))).astype(np.float32)
    cost[rng.random(cost.shape) < 0.05] = 0.0
    cost[rng.random(cost.shape) < 0.01] = -0.0
    cost[rng.random(cost.shape) < 0.005] = np.nan
    cost[rng.random(cost.shape) < 0.002] = np.inf
    return cost


@pytest.mark.parametrize("D,H,W", [(1, 7, 9), (3, 1, 17), (64, 13, 11), (65, 9, 20), (128, 16, 5), (129, 6, 6),
                                   (200, 11, 13), (256, 5, 9), (300, 4, 7), (512, 3, 5)])
def test_sgm_matches_oracle(D, H, W):
    from py_ortho_amd import ops
    print(f"sgm oracle {D}x{H}x{W}", flush=True)
    rng = np.random.default_rng(D * 1000 + H * 10 + W)
    cost = _volume(rng, D, H, W)
    cost[0, 0, 0] = np.nan
    out = ops.sgm(torch.from_numpy(cost).cuda(), 0.05, 0.6).cpu().numpy()
    assert _same(out, oracle.sgm(cost, 0.05, 0.6))


def test_sgm_batch_equals_single_and_zero_planes():
    from py_ortho_amd import ops
    print("sgm batch", flush=True)
    rng = np.random.default_rng(7)
    vols = np.stack([_volume(rng, 48, 20, 24) for _ in range(5)])
    vols[2, :, 3, :] = 0.0                           # a row of zero costs: minima are zeros
    vols[3, :, :, 4] = -0.0
    out = ops.sgm(torch.from_numpy(vols).cuda(), 0.1, 1.2).cpu().numpy()
    for b in range(5):
        assert _same(out[b], oracle.sgm(vols[b], 0.1, 1.2)), b


def test_sgm_limits():
    from py_ortho_amd import ops
    with pytest.raises(NotImplementedError):
        ops.sgm(torch.zeros((513, 2, 2), device="cuda"), 0.05, 0.6)
    assert ops.sgm(torch.zeros((0, 4, 4), device="cuda"), 0.05, 0.6).shape == (0, 4, 4)
