"""The depth-map file set on the device (formats.py, om_f32_to_f16 / om_f16_to_f32 /
om_filter_depth): the conversions equal NumPy's astype bit for bit over every float16 and millions
of float32 bit patterns (NumPy's conversion is the one the reference stores with, densify.py:591),
the filter equals densify.py:641 in NumPy, and the files written are byte-identical to the
reference's np.save of the same results."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _specials():
    bits = [0x00000000, 0x80000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000, 0x7F800001, 0xFF800001,
            0x7FA00000, 0x7F801FFF, 0x7F802000, 0x477FE000, 0x477FEFFF, 0x477FF000, 0x477FF001, 0x47800000,
            0x33800000, 0x33000000, 0x33000001, 0x387FC000, 0x38800000, 0x387FFFFF, 0x00000001, 0x3EAAAAAB]
    return np.array(bits, np.uint32)


def test_f32_to_f16_matches_numpy():
    from py_ortho_amd import formats
    print("f32 -> f16 bit patterns", flush=True)
    rng = np.random.default_rng(1)
    u = np.concatenate([_specials(), rng.integers(0, 2 ** 32, 4_000_000, dtype=np.uint32),
                        # dense around the float16 range, including its subnormals
                        (rng.integers(0x32000000, 0x47900000, 2_000_000, dtype=np.uint32) |
                         (rng.integers(0, 2, 2_000_000, dtype=np.uint32) << 31))])
    x = u.view(np.float32)
    with np.errstate(all="ignore"):
        ref = x.astype(np.float16)
    out = formats.to_f16(torch.from_numpy(x).cuda()).cpu().numpy()
    bad = np.flatnonzero(out.view(np.uint16) != ref.view(np.uint16))
    assert bad.size == 0, [(hex(u[i]), hex(out.view(np.uint16)[i]), hex(ref.view(np.uint16)[i])) for i in bad[:8]]


def test_f16_to_f32_matches_numpy():
    from py_ortho_amd import formats
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16)
    ref = h.astype(np.float32)
    out = formats.to_f32(torch.from_numpy(h).cuda()).cpu().numpy()
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_filter_matches_densify():
    from py_ortho_amd import formats
    rng = np.random.default_rng(2)
    D = rng.uniform(50, 90, (301, 407)).astype(np.float32)
    D[rng.random(D.shape) < 0.05] = np.nan
    S = rng.uniform(-1, 1, D.shape).astype(np.float32)
    S[rng.random(D.shape) < 0.01] = 0.3                     # at the threshold
    S16 = S.astype(np.float16)
    low = rng.random(D.shape) < 0.1
    ref = D.copy()
    Sf = S16.astype(np.float32)
    ref[(Sf < 0.3) | low] = np.nan
    d = torch.from_numpy(D).cuda()
    formats.filter_depth(d, torch.from_numpy(S16).cuda(), torch.from_numpy(low).cuda(), 0.3)
    assert np.array_equal(d.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_files_equal_the_reference_bytes(tmp_path):
    from py_ortho_amd import formats
    rng = np.random.default_rng(3)
    depth = rng.uniform(50, 90, (120, 160)).astype(np.float32)
    depth[5:9, 7:30] = np.nan
    score = rng.uniform(-1, 1, depth.shape).astype(np.float32)
    low = rng.random(depth.shape) < 0.2
    a, b = tmp_path / "ours", tmp_path / "ref"
    a.mkdir()
    b.mkdir()
    formats.save_pass1(str(a), 7, torch.from_numpy(depth).cuda(), torch.from_numpy(score).cuda(),
                       torch.from_numpy(low).cuda())
    np.save(b / "p00007.npy", depth)                        # densify.py:590-592
    np.save(b / "s00007.npy", score.astype(np.float16))
    np.save(b / "t00007.npy", low)
    for f in ("p00007.npy", "s00007.npy", "t00007.npy"):
        assert (a / f).read_bytes() == (b / f).read_bytes(), f
    d, s, t = formats.load_pass1(str(a), 7)
    assert np.array_equal(d.cpu().numpy().view(np.uint32), depth.view(np.uint32))
    assert s.dtype == torch.float16 and t.dtype == torch.bool
    formats.filter_depth(d, s, t, 0.3)
    formats.save_final(str(a), 7, d)
    D = np.load(b / "p00007.npy").astype(np.float32)
    D[(np.load(b / "s00007.npy").astype(np.float32) < 0.3) | np.load(b / "t00007.npy")] = np.nan   # densify.py:641
    np.save(b / "00007.npy", D)
    assert (a / "00007.npy").read_bytes() == (b / "00007.npy").read_bytes()
    assert np.array_equal(formats.load_final(str(a), 7).cpu().numpy().view(np.uint32), D.view(np.uint32))


def test_dense_cloud_of_fusion():
    from py_ortho_amd import formats
    pts = [torch.randn(10, 3, dtype=torch.float64, device="cuda"), torch.randn(5, 3, dtype=torch.float64,
                                                                                    device="cuda")]
    cols = [torch.zeros(10, 3, dtype=torch.uint8, device="cuda"), torch.ones(5, 3, dtype=torch.uint8, device="cuda")]
    nv = [torch.full((10,), 3, dtype=torch.uint8, device="cuda"), torch.full((5,), 2, dtype=torch.uint8,
                                                                               device="cuda")]
    ow = [torch.zeros(10, dtype=torch.int32, device="cuda"), torch.ones(5, dtype=torch.int32, device="cuda")]
    c = formats.dense_cloud(pts, cols, nv, ow, 0.04)
    assert c.xyz.shape == (15, 3) and c.xyz.dtype == np.float64 and c.rgb.dtype == np.uint8
    assert c.views.dtype == np.uint8 and c.cam.dtype == np.int32 and c.spacing == 0.04
    assert np.array_equal(c.xyz, torch.cat(pts).cpu().numpy())
    with pytest.raises(RuntimeError):
        formats.dense_cloud([], [], [], [], 0.04)
