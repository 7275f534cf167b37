"""The division-free bin index of the table lookup (bin_floor in device_common.hpp).

PathWeighters.cpp:67-68 takes the bin as floor(RN(d / step)) with
d = RN(curvature - minK) >= 0.  The device computes q = RN(d * RN(1 / step))
and uses floor(q) unless q lies within q * 2^-22 of an integer (or d is NaN),
where it falls back to the IEEE division.  The claim behind it: |q - RN(d /
step)| < q * 2^-22 for every d with q in the normal range (below it both are
far below 1/2), so outside that band no integer separates q from RN(d / step)
and both floors agree.  This test evaluates the same float
steps in C for EVERY float d in [0, maxK - minK] of the workloads' table grids
and checks both the premise and the conclusion, and that the fallback is rare.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests import cases

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>
/* every float d in [0, maxd]: counts (premise violations, floor mismatches on
   the fast path, fallbacks) */
void check(float step, float maxd, long long *out)
{
    const float inv = 1.0f / step;
    uint32_t lim;
    memcpy(&lim, &maxd, 4);
    long long bad_premise = 0, bad_floor = 0, slow = 0;
    for (uint32_t u = 0; u <= lim; ++u) {
        float d;
        memcpy(&d, &u, 4);
        const float q = d * inv;
        const float r = d / step;
        /* the relative bound holds for normal q; below, q and r are < 1/2 */
        if (q >= 0x1p-100f && !(fabsf(q - r) < q * 0x1p-22f)) ++bad_premise;
        const float k = rintf(q);
        const int fast = (d == 0.0f) | (fabsf(q - k) > q * 0x1p-22f);
        if (!fast) {
            ++slow;
            continue;
        }
        if (floorf(q) != floorf(r)) ++bad_floor;
    }
    out[0] = bad_premise;
    out[1] = bad_floor;
    out[2] = slow;
}
"""


@pytest.fixture(scope="module")
def checker():
    d = tempfile.mkdtemp(prefix="binfloor_")
    c, so = os.path.join(d, "bf.c"), os.path.join(d, "bf.so")
    open(c, "w").write(SRC)
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", "-o", so, c,
                           "-lm"])
    lib = C.CDLL(so)
    lib.check.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_longlong)]
    return lib


@pytest.mark.parametrize("case", [cases.C1, cases.C2, cases.C4, cases.C5], ids=lambda c: c.name)
def test_bin_floor_exhaustive(checker, oracle, case):
    t = oracle.build_table(cases.ds_of(case), eps=float(np.float32(case.sigma) ** 2), numStepsInt=200,
                           numCurvatureSteps=case.bins)
    step = np.float32(t.step)
    maxd = np.float32(np.float32(t.maxK) - np.float32(t.minK))
    out = (C.c_longlong * 3)()
    checker.check(step, maxd, out)
    n = int(maxd.view(np.uint32)) + 1
    print(f"[bin_floor {case.name}] step {float(step):.6g} range {float(maxd):.6g}: {n} floats, "
          f"premise violations {out[0]}, floor mismatches {out[1]}, fallbacks {out[2]} ({out[2] / n:.2e})")
    assert out[0] == 0 and out[1] == 0
    assert out[2] < 1e-4 * n
