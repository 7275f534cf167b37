"""Per-path parity rule of the GPU tests (FP32 kernel vs FP64 oracle on the
same Philox draws, same float initial curve and table).

Every valid path must satisfy

    |g - r| <= 1e-5 |r| + sum of the lerp jumps at its straddling joints,

where a joint straddles when its FP32 (kernel) and FP64 (oracle) curvatures
fall in different table bins, and its jump is the difference of the two lerped
log10 segment weights.  The reference's lerp fraction is in curvature units
(PathWeighters.cpp:79-82), so the segment weight is discontinuous at every bin
edge and no finite precision reproduces FP64 across one; a path with no
straddling joint is held to 1e-5 relative.  The sum of weights over the same
paths is then bounded by the worst per-path linear error:
|S_g / S_r - 1| <= max_p |10^(g_p - r_p) - 1| (S is a weighted mean of the
per-path ratios), plus 1e-6 for the two sides' summation rounding.
"""
import math

import numpy as np

from tests.lerp import joint_log10w

REL_PATH = 1e-5


def path_rule(g, r, valid, curv_g, curv_r, table, absorption):
    """Evaluates the rule on arrays of per-path log10 weights g (kernel) and r
    (oracle) and their per-joint curvatures ([paths, N-1]); valid selects the
    paths with a finite oracle weight.  Returns a dict of the figures the
    assertions use (see assert_path_rule)."""
    g = np.asarray(g, np.float64)
    r = np.asarray(r, np.float64)
    lg, bg = joint_log10w(curv_g, table, absorption)
    lr, br = joint_log10w(curv_r, table, absorption)
    straddle = bg != br
    jump = np.where(straddle, np.abs(lg - lr), 0.0).sum(axis=1)
    clean = valid & ~straddle.any(axis=1)
    err = np.abs(g - r)
    scale = np.maximum(np.abs(r), 1e-30)
    rel_clean = float((err[clean] / scale[clean]).max()) if clean.any() else 0.0
    excess = float((err[valid] - REL_PATH * scale[valid] - jump[valid]).max()) if valid.any() else -1.0
    ratio = float(np.max(np.abs(10.0 ** (g[valid] - r[valid]) - 1.0))) if valid.any() else 0.0
    return {"paths": int(len(g)), "valid": int(valid.sum()), "straddle_paths": int((valid & ~clean).sum()),
            "max_rel_clean": rel_clean, "max_excess": excess, "max_ratio_err": ratio,
            "max_rel_all": float((err[valid] / scale[valid]).max()) if valid.any() else 0.0}


def assert_path_rule(stats):
    assert stats["max_rel_clean"] <= REL_PATH, stats
    assert stats["max_excess"] <= 0.0, stats


def lse10(v):
    """log10 of sum 10^v over the finite entries of v (-inf when none)."""
    v = np.asarray(v, np.float64)
    v = v[np.isfinite(v)]
    if v.size == 0:
        return -math.inf
    m = v.max()
    return m + math.log10(np.sum(10.0 ** (v - m)))


def merge10(a, b):
    """log10(10^a + 10^b)."""
    if a == -math.inf:
        return b
    if b == -math.inf:
        return a
    m = max(a, b)
    return m + math.log10(10.0 ** (a - m) + 10.0 ** (b - m))
