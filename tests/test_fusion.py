"""Consistency and fusion on the CPU: the oracle (oracle/om_fuse.c) against the vectors the compiled
reference wrote (tests/golden/fusion_grid3x3.npz), bit for bit, and the host helpers of
py_ortho_amd.fusion."""
import numpy as np
import pytest

import fusion_cases
import oracle


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool((a.view(f"u{a.itemsize}") == b.view(f"u{b.itemsize}")).all())
    return bool((a == b).all())


@pytest.mark.parametrize("name", ["a", "b"])
def test_oracle_consistency_matches_reference(name):
    """densify._consistency counts of every view, as the compiled reference computed them."""
    g = fusion_cases.golden()
    s = g.settings[name]
    for k in g.have:
        nb = [j for j in g.nbrs[k] if j in g.depths]
        c = oracle.consistency(g.depths[k], g.cams[k], [g.depths[j] for j in nb], [g.cams[j] for j in nb],
                               [s.rel_tol[k][j] for j in nb], s.px_tol)
        assert _same(c, s.cnt[k]), k
        assert c.max() > 0


@pytest.mark.parametrize("name", ["a", "b"])
def test_oracle_fusion_matches_reference(name):
    """densify._fuse_loop points, colours, view counts, owners and used maps, bit for bit."""
    g = fusion_cases.golden()
    s = g.settings[name]
    n = len(g.cams)
    nb, tol = fusion_cases.nbr_table(g.nbrs, set(g.have), s.rel_tol, n)
    depths = [g.depths.get(k) for k in range(n)]
    pts, cols, nv, owner, used = oracle.fuse(g.have, depths, g.cams, list(g.rgb), nb, tol, s.px_tol, s.min_views,
                                             s.stride)
    assert len(pts) == len(s.pts) > 1000
    assert _same(pts, s.pts) and _same(cols, s.cols) and _same(nv, s.nv) and _same(owner, s.owner)
    for k in g.have:
        assert _same(used[k].astype(bool), s.used[k]), k
    assert (nv >= s.min_views).all()


def test_golden_pins_the_product_order():
    """The golden points need the 3x3 products formed as OpenBLAS does (a0*b0, then two fused
    multiply-adds): with a plain three-term sum planted in the oracle they differ."""
    g = fusion_cases.golden()
    s = g.settings["a"]
    n = len(g.cams)
    nb, tol = fusion_cases.nbr_table(g.nbrs, set(g.have), s.rel_tol, n)
    depths = [g.depths.get(k) for k in range(n)]
    L = oracle.orc()
    L.orc_fuse_plain_products(1)
    try:
        pts = oracle.fuse(g.have, depths, g.cams, list(g.rgb), nb, tol, s.px_tol, s.min_views, s.stride)[0]
    finally:
        L.orc_fuse_plain_products(0)
    assert not (len(pts) == len(s.pts) and _same(pts, s.pts))
    # NumPy's own matmul here agrees with the fused order, a plain sum does not
    rng = np.random.default_rng(3)
    M = rng.standard_normal((4096, 3)) * 50
    R = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    plain = (M[:, :1] * R[0] + M[:, 1:2] * R[1]) + M[:, 2:3] * R[2]
    assert not np.array_equal(M @ R, plain)


def test_used_map_helpers():
    from py_ortho_amd import fusion
    assert fusion.grid_shape(120, 160, 2) == (60, 80)
    assert fusion.grid_shape(121, 161, 3) == (41, 54)
    assert fusion.grid_shape(5, 7, 0) == (5, 7)
    g = np.zeros((41, 54), np.uint8)
    g[3, 5] = 1
    full = fusion.used_full(g, 121, 161, 3)
    assert full.shape == (121, 161) and full.sum() == 1 and full[9, 15]


def test_dense_cloud_fields():
    """formats.DenseCloud has densify.DenseCloud's fields, in order (densify.py:84-90)."""
    import dataclasses

    from py_ortho_amd import formats
    assert [f.name for f in dataclasses.fields(formats.DenseCloud)] == ["xyz", "rgb", "views", "cam", "spacing"]
