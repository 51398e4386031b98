"""View preparation on the device (om_depth_view through prep.depth_view / prepare_view): bit for
bit against the compiled reference's outputs and the oracle, for the vector and the unaligned
paths and odd pixel counts."""
import numpy as np
import pytest
import torch

import oracle
from test_prep import _same, cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", cases(), ids=lambda c: str(c[0]))
def test_depth_view_matches_reference(case):
    from py_ortho_amd import prep
    i, rgb, g, b, m, grey, rgb8 = case
    print(f"depth_view golden {i} {rgb.shape}", flush=True)
    mt = None if m is None else torch.from_numpy(m).cuda()
    a, c = prep.depth_view(torch.from_numpy(rgb).cuda(), g, b, mt)
    assert _same(a.cpu().numpy(), grey) and _same(c.cpu().numpy(), rgb8)


@pytest.mark.parametrize("h,w,offset", [(1200, 1600, 0), (301, 333, 0), (77, 5, 1), (1, 3, 3)])
def test_depth_view_matches_oracle(h, w, offset):
    from py_ortho_amd import prep
    print(f"depth_view oracle {h}x{w} offset {offset}", flush=True)
    rng = np.random.default_rng(h * w + offset)
    rgb = rng.integers(0, 256, (h, w, 3)).astype(np.uint8)
    m = rng.random((h, w)) < 0.8
    g, b = (0.9, 1.17, 1.004), (-20.0, 3.5, 0.125)
    buf = torch.empty(h * w * 3 + offset, dtype=torch.uint8, device="cuda")
    dev = buf[offset:].view(h, w, 3)                       # offset 1 or 3: the per-pixel path
    dev.copy_(torch.from_numpy(rgb))
    a, c = prep.depth_view(dev, g, b, torch.from_numpy(m).cuda())
    ra, rc = oracle.depth_view(rgb, g, b, m)
    assert _same(a.cpu().numpy(), ra) and _same(c.cpu().numpy(), rc)
    a2, c2 = prep.depth_view(dev, None, None, None)
    ra2, rc2 = oracle.depth_view(rgb)
    assert _same(a2.cpu().numpy(), ra2) and _same(c2.cpu().numpy(), rc2)


def test_prepare_view_pyramid():
    from py_ortho_amd import prep
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, (121, 203, 3)).astype(np.uint8)
    m = rng.random((121, 203)) < 0.95
    grey, rgb8, pyr = prep.prepare_view(rgb, (1.1, 0.9, 1.0), None, m, levels=3)
    g0, _ = oracle.depth_view(rgb, (1.1, 0.9, 1.0), None, m)
    assert _same(grey.cpu().numpy(), g0)
    assert [tuple(p.shape) for p in pyr] == [(121, 203), (60, 101), (30, 50)]
    assert _same(pyr[1].cpu().numpy(), oracle.half(g0)) and _same(pyr[2].cpu().numpy(), oracle.half(oracle.half(g0)))
