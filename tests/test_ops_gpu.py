"""GPU height-sweep backend operations and ray maps against the reference's NumPy expressions and
the oracle (bit-exact where the reference's arithmetic is sequential, tolerance where NumPy's own
summation order is unspecified)."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_rays_world_matches_reference_expression():
    from py_ortho_amd import dense
    rng = np.random.default_rng(1)
    for H, W, k in ((30, 40, (0.0, 0.0, 0.0)), (37, 53, (-0.05, 0.01, 0.002))):
        w = rng.normal(0, 0.1, 3) + np.array([np.pi, 0, 0])
        th = np.linalg.norm(w)
        kk = w / th
        K = np.array([[0, -kk[2], kk[1]], [kk[2], 0, -kk[0]], [-kk[1], kk[0], 0]])
        R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
        f, cx, cy = 45.0, (W - 1) / 2 + 0.3, (H - 1) / 2 - 0.2
        exp = oracle.rays_world(H, W, R, cx, cy, f, *k)
        got = dense.rays_world(H, W, R, cx, cy, f, *k).cpu().numpy()
        assert np.max(np.abs(got - exp)) <= 1e-7 * np.max(np.abs(exp)), float(np.max(np.abs(got - exp)))
        frac = np.mean(got.view(np.uint32) == exp.view(np.uint32))
        assert frac > 0.999, frac


def test_box_sat_mean_matches_float64_cumsum():
    """Same closed form as the reference's test_native_parity.py:520-539."""
    from py_ortho_amd import ops
    rng = np.random.default_rng(73)
    for shape in [(7, 11), (2, 3, 19, 13), (1, 1)]:
        a = rng.normal(0, 100, shape).astype(np.float32)
        if a.size > 10:
            a.flat[0] = -0.0
            a.flat[-2] = np.nan
        for r in (0, 1, 5, 23):
            k = 2 * r + 1
            padded = np.pad(a.astype(np.float64), [(0, 0)] * (a.ndim - 2) + [(r + 1, r)] * 2)
            c = padded.cumsum(axis=-2).cumsum(axis=-1)
            ref = ((c[..., k:, k:] - c[..., :-k, k:] - c[..., k:, :-k] + c[..., :-k, :-k]) / (k * k)).astype(np.float32)
            got = ops.box(_dev(a), r).cpu().numpy()
            assert _same(ref, got), (shape, r)


def test_topk_and_argminmax_match_numpy():
    from py_ortho_amd import ops
    rng = np.random.default_rng(0)
    st = rng.uniform(-1, 1, (6, 4, 30, 40)).astype(np.float32)
    for k in (1, 3, 6):
        exp = np.partition(st, 6 - k, axis=0)[6 - k:].mean(axis=0)
        got = ops.topk_mean(_dev(st), k).cpu().numpy()
        assert np.abs(exp - got).max() < 1e-5
    st[2, 0, 0, 0] = st[4, 0, 0, 0] = 5.0      # ties: first index
    st[1, 1, 2, 3] = np.nan                    # NaN wins (NumPy)
    for want_max, fn in ((False, np.argmin), (True, np.argmax)):
        i0 = fn(st, axis=0)
        v0 = np.take_along_axis(st, i0[None], axis=0)[0]
        i1, v1 = ops.argminmax0(_dev(st), want_max)
        assert np.array_equal(i0, i1.cpu().numpy()) and _same(v0, v1.cpu().numpy())


def test_sample_view_matches_oracle_and_known_answer():
    from py_ortho_amd import ops
    rng = np.random.default_rng(0)
    im = rng.uniform(0, 255, (200, 240, 3)).astype(np.float32)
    th = 0.05
    R = np.array([[1, 0, 0], [0, np.cos(np.pi + th), -np.sin(np.pi + th)], [0, np.sin(np.pi + th), np.cos(np.pi + th)]])
    cam = (R, np.array([3.0, -2.0, 40.0]), 300.0, -0.05, 0.01, 0.02, 119.5, 99.5)
    Z = rng.uniform(0, 8, (4, 60, 70)).astype(np.float32)
    oc, vc = oracle.sample_view(im, cam, -10.0, 8.0, 0.25, Z)
    og, vg = ops.sample_view(_dev(im), cam, -10.0, 8.0, 0.25, _dev(Z))
    assert _same(vc, vg.cpu().numpy()) and _same(oc, og.cpu().numpy())
    # known answer (test_3d.py:50-61): a bright pixel seen straight down
    img = np.zeros((100, 100, 1), np.float32)
    img[50, 60, 0] = 1.0
    Rn = np.diag([1.0, -1.0, -1.0])
    f = 100.0
    X, Y = (60 - 49.5) / f * 10, -(50 - 49.5) / f * 10
    out, valid = ops.sample_view(_dev(img), (Rn, np.array([0.0, 0.0, 10.0]), f, 0.0, 0.0, 0.0, 49.5, 49.5),
                                 X - 0.005, Y + 0.005, 0.01, _dev(np.zeros((1, 1), np.float32)))
    assert int(valid[0, 0]) == 1 and abs(float(out[0, 0, 0]) - 1.0) < 1e-4


def _sample_view_golden():
    import test_oracle
    return list(test_oracle.sample_view_cases())


@pytest.mark.parametrize("case", _sample_view_golden(), ids=lambda c: c[0])
def test_sample_view_matches_reference_golden(case):
    """om_sample_view == the compiled reference's CPUBackend.sample_view (_mvs.pyx:13-54,
    backend.py:67-77), bit for bit: distorted, tilted and oblique cameras, (D, H, W) stacks, cells
    behind the camera and on the last pixel row and column."""
    from py_ortho_amd import ops
    name, img, cam, X0, Y0, gsd, Z, want, wvalid = case
    out, valid = ops.sample_view(_dev(img), cam, X0, Y0, gsd, _dev(Z))
    assert np.array_equal(valid.cpu().numpy(), wvalid)
    assert _same(out.cpu().numpy(), want)
