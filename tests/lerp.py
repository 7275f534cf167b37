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
