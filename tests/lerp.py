"""Vectorised per-joint segment weights with the reference's float rounding.

Restates PathWeighters.cpp:41-88 for arrays of curvatures (float32 arithmetic
step by step, as the reference does in float): returns each joint's log10
weight and its table bin, used by the parity tests to locate joints whose
curvature crosses a bin edge.
"""
import math

import numpy as np


def joint_log10w(curv, table, absorption):
    k = np.asarray(curv, np.float32)
    T = np.asarray(table.table, np.float32)
    minK, maxK, step = np.float32(table.minK), np.float32(table.maxK), np.float32(table.step)
    k = np.clip(k, minK, maxK)
    dist = (k - minK).astype(np.float32)
    real_idx = (dist / step).astype(np.float32)
    left = np.floor(real_idx).astype(np.int64)
    left = np.clip(left, 0, len(T) - 1)
    right = np.where(left == len(T) - 1, left, left + 1)
    interp = (dist - (left.astype(np.float32) * step).astype(np.float32)).astype(np.float32)
    a = (T[left] * (np.float32(1.0) - interp)).astype(np.float32)
    b = (T[right] * interp).astype(np.float32)
    w = (a + b).astype(np.float32).astype(np.float64)
    # PathWeighters.cpp:37 takes a float exp; glibc's expf is correctly rounded,
    # so round the double exp once (NumPy's float32 exp may differ by an ulp).
    absorb = float(np.float32(math.exp(float(np.float32(-np.float32(absorption) * np.float32(table.ds))))))
    with np.errstate(divide="ignore", invalid="ignore"):
        lw = np.log10(w * absorb)
    return lw, left


def joint_log10w_spread(curv, table, absorption, delta):
    """For curvatures known only to within +-delta (per joint, absolute):
    each joint's log10 weight at curv and the spread (max - min) of the log10
    weight over [curv - delta, curv + delta].  The weight is the lerp of a
    table, linear inside a bin and jumping at each bin edge (the reference's
    fraction is in curvature units), so its extremes on the interval lie at
    the two ends or on either side of an edge inside it: those are the points
    sampled (each edge with its float neighbours two ulps either way, since
    the bin index is formed in float arithmetic)."""
    k = np.asarray(curv, np.float64)
    d = np.broadcast_to(np.asarray(delta, np.float64), k.shape)
    minK, step = float(np.float32(table.minK)), float(np.float32(table.step))
    lw, _ = joint_log10w(k, table, absorption)
    pts = [k - d, k + d]
    lo_bin = np.floor((k - d - minK) / step)
    hi_bin = np.floor((k + d - minK) / step)
    # an interval of 2 delta spans at most a few edges; sample each of them
    for m in range(int(np.max(hi_bin - lo_bin, initial=0)) + 1):
        e = (minK + (lo_bin + 1 + m) * step).astype(np.float32)
        inside = (e >= k - d) & (e <= k + d)
        lo, hi = e, e
        pts.append(np.where(inside, e.astype(np.float64), k))
        for _ in range(2):
            lo = np.nextafter(lo, np.float32(-np.inf))
            hi = np.nextafter(hi, np.float32(np.inf))
            pts.append(np.where(inside, lo.astype(np.float64), k))
            pts.append(np.where(inside, hi.astype(np.float64), k))
    vals = np.stack([joint_log10w(p, table, absorption)[0] for p in pts] + [lw])
    with np.errstate(invalid="ignore"):
        spread = np.nanmax(vals, axis=0) - np.nanmin(vals, axis=0)
    return lw, spread
