"""CPU proofs behind the reference kernels' short inverse distance
(utc_amd/csrc/ref_inv.h), exhaustive over every float significand.

1. Tuckerman's test: from any s0 within 1 ulp of sqrt(x), the choice
   "s0 - ulp if x <= s0 (s0 - ulp); s0 + ulp if x > s0 (s0 + ulp); else s0"
   is the correctly rounded sqrtf(x).  Checked for every x in [1, 4) (every
   significand at both exponent parities) and every such s0.  The signs of
   x - s0 (s0 -/+ ulp) are exact in float64 (the 48-bit products are exact
   and x lies within a factor two of them), as the device's FMA computes them.
2. 1/s for float s stays at least 2^-48 (relative) away from every float
   midpoint, so any approximation of 1/s within 2^-49 rounds to the float
   that the reference's (float)(1.0 / (double)s) gives.  Checked for every s
   in [1, 2): |1 - s m| is exact in float64 (a 24-by-25-bit product).
3. The FP32 reciprocal of the device path: y1 = fma(y0, e, y0) with
   e = fma(-s, y0, 1), from y0 = v_rcp_f32(s).  Evaluated exactly (the
   products and the residual in float64, the one or two ambiguous roundings
   per start in rational arithmetic): from y0 = RN(1/s) the step returns
   RN(1/s) for every s in [1, 2); from a y0 one ulp off it fails for 33
   significands (32 above, 1 below), so the device path is exact only
   because v_rcp_f32 does not land on those; the device probe checks that
   over every input.
The device-side check of the whole input range is
tests/test_gpu_forward.py::test_reference_inverse_distance_exhaustive.
"""
from fractions import Fraction

import numpy as np


def floats_between(lo, hi):
    a = np.float32(lo).view(np.uint32)
    b = np.float32(hi).view(np.uint32)
    return np.arange(a, b, dtype=np.uint32)


def test_tuckerman_choice_is_correct_rounding_of_sqrt():
    bits = floats_between(1.0, 4.0)
    x = bits.view(np.float32).astype(np.float64)
    want = np.sqrt(x).astype(np.float32)  # sqrt rounded to double then float: correctly rounded
    wbits = want.view(np.uint32)
    for d in (-1, 0, 1):
        s0b = (wbits.astype(np.int64) + d).astype(np.uint32)
        s0 = s0b.view(np.float32).astype(np.float64)
        lo = (s0b - np.uint32(1)).view(np.float32).astype(np.float64)
        hi = (s0b + np.uint32(1)).view(np.float32).astype(np.float64)
        s = s0b.copy()
        s = np.where(x - lo * s0 <= 0.0, s0b - np.uint32(1), s)
        s = np.where(x - hi * s0 > 0.0, s0b + np.uint32(1), s)
        assert np.array_equal(s, wbits), f"start offset {d}: {np.count_nonzero(s != wbits)} wrong"


def test_reciprocal_keeps_2_to_minus_48_from_float_midpoints():
    bits = floats_between(1.0, 2.0)
    s = bits.view(np.float32).astype(np.float64)
    q = (1.0 / s).astype(np.float32)
    qb = q.view(np.uint32)
    up = (qb + np.uint32(1)).view(np.float32).astype(np.float64)
    dn = (qb - np.uint32(1)).view(np.float32).astype(np.float64)
    qd = q.astype(np.float64)
    m1, m2 = (qd + up) / 2, (qd + dn) / 2  # the two midpoints around q (exact in float64)
    dist = np.minimum(np.abs(1.0 - s * m1), np.abs(1.0 - s * m2))
    assert dist.min() == 2.0 ** -48
    # so an approximation within 2^-49 of 1/s rounds to q: spot-check both sides
    for f in (1 + 2.0 ** -49, 1 - 2.0 ** -49):
        approx = (1.0 / s) * f
        assert np.array_equal(approx.astype(np.float32), q)


def _rn32(x):
    """Nearest float32 to the rational x (ties to even), x positive."""
    c = np.float32(float(x))
    cb = int(np.array([c]).view(np.uint32)[0])
    best = None
    for k in (cb - 1, cb, cb + 1):
        v = np.array([k], np.uint32).view(np.float32)[0]
        key = (abs(Fraction(float(v)) - x), k & 1)
        if best is None or key < best[0]:
            best = (key, v)
    return best[1]


def test_fp32_reciprocal_correction_from_within_one_ulp():
    bits = floats_between(1.0, 2.0)
    s = bits.view(np.float32).astype(np.float64)
    q = (1.0 / s).astype(np.float32)
    qb = q.view(np.uint32)
    bad = {}
    for d in (-1, 0, 1):
        y0 = (qb.astype(np.int64) + d).astype(np.uint32).view(np.float32).astype(np.float64)
        e = (1.0 - s * y0).astype(np.float32).astype(np.float64)  # the 48-bit product is exact: one rounding
        t = y0 + y0 * e  # float64; rounding it to float32 is exact unless it sits on a float32 midpoint
        y1 = t.astype(np.float32)
        t32 = y1.astype(np.float64)
        up = (y1.view(np.uint32) + np.uint32(1)).view(np.float32).astype(np.float64)
        dn = (y1.view(np.uint32) - np.uint32(1)).view(np.float32).astype(np.float64)
        for i in np.nonzero((t == (t32 + up) / 2) | (t == (t32 + dn) / 2))[0]:
            y1[i] = _rn32(Fraction(float(y0[i])) * (1 + Fraction(float(e[i]))))
        bad[d] = int(np.count_nonzero(y1.view(np.uint32) != qb))
    assert bad == {-1: 1, 0: 0, 1: 32}, bad
