"""The half-angle trig of the draw (sincospi_half in device_common.hpp).

The device evaluates sin(pi x), cos(pi x) for |x| < 1/2 with a range swap at
|x| = 1/4 and two short polynomials.  This test reads the coefficients from
the header (so a change there is what gets tested), evaluates the same float
steps on the CPU (each FMA rounded once, from the exact FP64 product-sum of
float operands) for every x the draw can produce, and bounds the error
against FP64 sin/cos: at most 2 ulp.  The GPU parity tests cover the device
execution end to end.
"""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "twisty_amd", "csrc",
                   "device_common.hpp")


def coefficients():
    src = open(HDR).read()
    body = src[src.index("void sincospi_half"):]
    body = body[:body.index("\n}\n")]
    lit = [float.fromhex(m.rstrip("f")) for m in re.findall(r"-?0x1\.[0-9a-f]+p[+-]\d+f", body)]
    assert len(lit) == 8, lit
    return lit[:4], lit[4:]  # P from highest degree down, then Q from highest degree down (constant 1 implicit)


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c).astype(np.float32)


def sincospi_half(x):
    """The device steps of sincospi_half on float32 x, |x| <= 1/2."""
    P, Q = coefficients()
    f32 = np.float32
    a = np.abs(x)
    swap = a > f32(0.25)
    y = np.where(swap, (f32(0.5) - a).astype(np.float32), a)
    t = (y.astype(np.float64) * y.astype(np.float64)).astype(np.float32)
    p = np.full_like(t, f32(P[0]))
    for c in P[1:]:
        p = fma32(p, t, float(f32(c)))
    sy = (p.astype(np.float64) * y.astype(np.float64)).astype(np.float32)
    q = np.full_like(t, f32(Q[0]))
    for c in Q[1:] + [1.0]:
        q = fma32(q, t, float(f32(c)))
    s = np.where(swap, q, sy)
    s = np.where(x < 0, -s, s)
    c = np.where(swap, sy, q)
    return s, c


def ulp_errors(s, c, ref_s, ref_c):
    ulp_s = np.spacing(np.abs(ref_s).astype(np.float32)).astype(np.float64)
    ulp_c = np.spacing(np.abs(ref_c).astype(np.float32)).astype(np.float64)
    return np.abs(s.astype(np.float64) - ref_s) / ulp_s, np.abs(c.astype(np.float64) - ref_c) / ulp_c


def test_sincospi_half_error_bound():
    odd = np.arange(-(1 << 24) + 1, 1 << 24, 2, dtype=np.int64)
    x = (odd.astype(np.float64) * 2.0 ** -25).astype(np.float32)
    s, c = sincospi_half(x)
    xs = np.pi * x.astype(np.float64)
    ref_s, ref_c = np.sin(xs), np.cos(xs)
    es, ec = ulp_errors(s, c, ref_s, ref_c)
    print(f"[sincospi_half] sin max {es.max():.2f} ulp mean {es.mean():.3f}; cos max {ec.max():.2f} ulp "
          f"mean {ec.mean():.3f}")
    assert es.max() <= 2.0 and ec.max() <= 2.0
    assert es.mean() < 0.5 and ec.mean() < 0.5
    # the swap point and both ends: exact symmetries of the reduction
    assert np.all(np.sign(s) == np.sign(x)) and np.all(c > 0)


def test_sincospi_turn_error_bound():
    """sincospi_turn (the generator's azimuth): sin, cos of 2 pi u for every
    u = k 2^-24 in [0, 1), through the exact reductions x = 2u - 1 and
    z = 1 - |x| into sincospi_half; at most 2 ulp, exact at the quarter turns
    and signs matching the quadrant."""
    f32 = np.float32
    u = (np.arange(1 << 24, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    x = (f32(2.0) * u - f32(1.0)).astype(np.float32)
    a = np.abs(x)
    far = a > f32(0.5)
    z = np.where(far, (f32(1.0) - a).astype(np.float32), a)
    sz, cz = sincospi_half(z)
    s = np.where(x < 0, sz, -sz)
    c = np.where(far, cz, -cz)
    th = 2.0 * np.pi * u.astype(np.float64)
    ref_s, ref_c = np.sin(th), np.cos(th)
    # near the zeros of sin and cos compare absolutely: an ulp of the FP64 value
    # is not meaningful where the float argument itself carries the rounding
    es, ec = ulp_errors(s, c, ref_s, ref_c)
    es = np.where(np.abs(ref_s) < 1e-6, 0.0, es)
    ec = np.where(np.abs(ref_c) < 1e-6, 0.0, ec)
    print(f"[sincospi_turn] sin max {es.max():.2f} ulp; cos max {ec.max():.2f} ulp")
    assert es.max() <= 2.0 and ec.max() <= 2.0
    assert np.abs(s - ref_s).max() < 2e-7 and np.abs(c - ref_c).max() < 2e-7
    q = np.arange(4) << 22  # u = 0, 1/4, 1/2, 3/4
    assert list(s[q]) == [0.0, 1.0, 0.0, -1.0] and list(c[q]) == [1.0, 0.0, -1.0, 0.0]
    assert np.all(np.sign(s[1:1 << 23]) == 1) and np.all(np.sign(s[(1 << 23) + 1:]) == -1)
