"""The FP64 kernels' arithmetic shortcuts against IEEE double arithmetic, bit for bit.

- Quotients over one divisor share its reciprocal when the operands lie in the range where the
  compiler's own division would neither scale its operands nor fix up its result, and divide
  plainly otherwise: the projection's pair xc / zc, yc / zc (om_fp64_steps: qa, qb) and the sweep
  NCC's three means (qs, the middle of three).  NumPy's float64 division is IEEE correctly
  rounded, so it is the exact expectation.
- The bilinear tap's fraction u - (int)u, computed as v_fract_f64 (fr), for u >= 0 (the tap only
  runs inside [0, w - 1.001]).
- The sweep NCC's square root of the source variance (sq): for 1e-6 < v < inf the compiler's
  Newton sequence without its operand scaling, sqrt otherwise.  NumPy's sqrt is IEEE correctly
  rounded.

Operands are drawn inside the shared-reciprocal range, at its edges and far beyond it, with zeros,
signed zeros, infinities and NaNs mixed in; the divisor always satisfies the projection's own
precondition zc > 1e-6 (or is NaN).
"""
import numpy as np
import pytest
import torch

from py_ortho_amd import _lib
from py_ortho_amd._lib import check
from py_ortho_amd.dense import _stream

pytestmark = pytest.mark.gpu

N = 1 << 20


def _run(a, b, c):
    A, B, C = (torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).cuda() for v in (a, b, c))
    qa, qb, qs, fr, sq = (torch.empty_like(A) for _ in range(5))
    check(_lib.lib().om_fp64_steps(A.data_ptr(), B.data_ptr(), C.data_ptr(), A.numel(), qa.data_ptr(),
                                   qb.data_ptr(), qs.data_ptr(), fr.data_ptr(), sq.data_ptr(), _stream()),
          "om_fp64_steps")
    torch.cuda.synchronize()
    return qa.cpu().numpy(), qb.cpu().numpy(), qs.cpu().numpy(), fr.cpu().numpy(), sq.cpu().numpy()


def _same(x, y):
    """bit-identical, or NaN in both"""
    return (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))


def _operands(kind, rng):
    s = lambda: rng.choice([-1.0, 1.0], N) * rng.uniform(0.5, 1.0, N)  # noqa: E731
    if kind == "typical":      # world coordinates in a camera frame: |xc|, |yc| ~ 1e-9 .. 1e9, zc ~ 1e-6 .. 1e6
        ea, eb, ec = rng.uniform(-30, 30, N), rng.uniform(-30, 30, N), rng.uniform(-19, 20, N)
    elif kind == "lower_edge":  # numerators around 2^-900, the lower end of the shared-reciprocal range
        ea, eb, ec = -900 + rng.uniform(-6, 6, N), rng.uniform(-40, 40, N), rng.uniform(-19, 4, N)
    elif kind == "upper_edge":  # numerators and divisor around 2^100, the upper end
        ea, eb, ec = 100 + rng.uniform(-4, 4, N), rng.uniform(-40, 100, N), 100 + rng.uniform(-3, 3, N)
    else:                       # anything a double holds
        ea, eb, ec = rng.uniform(-1070, 1023, N), rng.uniform(-1070, 1023, N), rng.uniform(-19, 1023, N)
    a = s() * np.exp2(ea)
    b = s() * np.exp2(eb)
    c = np.abs(s()) * np.exp2(ec)
    c = np.where(c > 1e-6, c, 1e-6 * (1 + rng.uniform(0, 1, N)))
    # zeros, signed zeros and non-finite operands in a few slots
    a[::997] = 0.0
    b[1::991] = -0.0
    a[2::4093] = np.inf
    b[3::4091] = -np.inf
    a[4::8191] = np.nan
    c[5::8179] = np.nan
    return a, b, c


@pytest.mark.parametrize("kind", ["typical", "lower_edge", "upper_edge", "any"])
def test_projection_quotients_are_ieee_division(kind):
    rng = np.random.default_rng(["typical", "lower_edge", "upper_edge", "any"].index(kind) + 11)
    a, b, c = _operands(kind, rng)
    qa, qb, qs, _, _ = _run(a, b, c)
    with np.errstate(all="ignore"):
        ea, eb, es = a / c, b / c, (a + b) / c
    bad_a, bad_b, bad_s = ~_same(qa, ea), ~_same(qb, eb), ~_same(qs, es)
    assert not bad_a.any(), (kind, a[bad_a][:3], c[bad_a][:3], qa[bad_a][:3], ea[bad_a][:3])
    assert not bad_b.any(), (kind, b[bad_b][:3], c[bad_b][:3], qb[bad_b][:3], eb[bad_b][:3])
    assert not bad_s.any(), (kind, a[bad_s][:3], b[bad_s][:3], c[bad_s][:3], qs[bad_s][:3], es[bad_s][:3])


def test_tap_fraction_is_u_minus_its_integer_part():
    rng = np.random.default_rng(5)
    u = np.concatenate([
        rng.uniform(0, 1, N // 4),                         # inside the first pixel
        rng.uniform(0, 8192, N // 4),                      # image coordinates
        np.exp2(rng.uniform(-60, 30, N // 4)),             # tiny to 2^30
        np.arange(N // 4, dtype=np.float64) / 4.0,         # exact quarters, whole numbers among them
    ])
    u[::1009] = np.nextafter(np.floor(u[::1009]) + 1.0, 0.0)  # just below the next integer
    _, _, _, fr, _ = _run(u, u, np.ones_like(u))
    exp = u - np.trunc(u)
    bad = ~_same(fr, exp)
    assert not bad.any(), (u[bad][:3], fr[bad][:3], exp[bad][:3])


def test_variance_square_root_is_ieee_sqrt():
    rng = np.random.default_rng(17)
    q = N // 8
    v = np.concatenate([
        rng.uniform(0, 1, q),                              # variances of [0, 1] images
        np.exp2(rng.uniform(-20, 20, q)),                  # 1e-6 .. 1e6
        1e-6 * (1 + np.exp2(rng.uniform(-52, 0, q))),      # just above the guard
        np.exp2(rng.uniform(-1074, 1024, q)),              # anything positive a double holds
        np.exp2(rng.uniform(1000, 1024, q)),               # up to the largest doubles
        rng.integers(1, 1 << 40, q).astype(np.float64) ** 2,  # exact squares
        np.nextafter(rng.integers(1, 1 << 20, q).astype(np.float64) ** 2, np.inf),  # an ulp off a square
        rng.uniform(-1, 1e-6, q),                          # at or below the guard (plain sqrt)
    ])
    v[::4099] = np.inf
    v[1::4093] = 0.0
    v[2::4091] = np.nan
    v[3::4089] = np.finfo(np.float64).max
    _, _, _, _, sq = _run(v, v, v)
    with np.errstate(all="ignore"):
        exp = np.sqrt(v)
    bad = ~_same(sq, exp)
    assert not bad.any(), (v[bad][:3], sq[bad][:3], exp[bad][:3])
