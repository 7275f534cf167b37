"""GPU exit-direction histogram against the oracle.

The histogram is build-defined (the reference bins only weight magnitudes):
bin (it, ip) of the last free tangent t_{N-2}, cos theta = t.e over nTheta bins
of [-1, 1], phi = atan2(t.f2, t.f1) over nPhi bins of [-pi, pi), each bin the
share of the weight sum of its paths.

Checks:
  * per-path bins equal the oracle's bins of the FP64 exit tangent, except for
    paths whose FP64 coordinate lies within 1e-3 of a bin edge (FP32 tangents
    differ from FP64 by ~1e-6);
  * bin arithmetic: the GPU histogram equals an FP64 histogram of the GPU's
    own per-path weights, |dh| <= 3e-5 h + 1e-12 (the float log10 output
    carries ~1e-5 relative weight error at log10 w ~ -40);
  * against the oracle's weights in the GPU's bins: L1 distance no larger than
    the per-path weight differences allow, sum_p |w_gpu - w_ref| / W (the
    per-path log10 weights themselves are checked to 1e-5 in the forward
    tests, with the bin-edge allowance of the reference's lerp);
  * invariants: sum of bins = 1, bins >= 0, a constant table gives the path
    fraction of every bin.
"""
import math

import numpy as np
import pytest

from tests import cases
from tests.parity import assert_path_rule, path_rule
from twisty_amd import capi

pytestmark = pytest.mark.gpu


def frame(e):
    e = np.asarray(e, np.float64)
    e = e / np.linalg.norm(e)
    k = 0
    if abs(e[1]) < abs(e[0]):
        k = 1
    if abs(e[2]) < min(abs(e[0]), abs(e[1])):
        k = 2
    a = np.zeros(3)
    a[k] = 1.0
    f1 = a - (a @ e) * e
    f1 /= np.linalg.norm(f1)
    return e, f1, np.cross(e, f1)


def edge_distance(t, e, nt, nph):
    """Distance (in bins) of each FP64 tangent's coordinates to the nearest bin edge."""
    e, f1, f2 = frame(e)
    t = t / np.linalg.norm(t, axis=1, keepdims=True)
    u = (t @ e + 1.0) * 0.5 * nt
    v = (np.arctan2(t @ f2, t @ f1) + math.pi) / (2 * math.pi) * nph
    return np.minimum(np.abs(u - np.round(u)), np.abs(v - np.round(v)))


def reference_histogram(log10w, bins, nb):
    ok = np.isfinite(log10w)
    w = np.zeros(len(log10w))
    m = log10w[ok].max()
    w[ok] = 10.0 ** (log10w[ok] - m)
    h = np.bincount(bins[ok], weights=w[ok], minlength=nb)
    return h / h.sum()


@pytest.mark.parametrize("case,nt,nph,paths", [(cases.C2, 64, 64, 1 << 15), (cases.MID, 32, 16, 1 << 14),
                                               (cases.C1, 256, 256, 1 << 15)])
def test_exit_histogram_matches_oracle(hip_device, oracle, case, nt, nph, paths):
    print(f"[hist {case.name} {nt}x{nph} paths={paths}]", flush=True)
    ctx = cases.hip_context(case, oracle)
    ctx.set_exit_histogram(nt, nph)
    res = ctx.run(paths, case.K, seed=case.seed, per_path=True)
    h, gbins = ctx.exit_histogram(paths)
    table, pos32 = cases.setup(case, oracle)
    ref = oracle.simulate(pos32.astype(np.float64), case.N, case.L, case.K, paths, table, case.absorption,
                          seed=case.seed, want_exit=True)
    e = (0.0, 0.0, 1.0)
    obins = oracle.exit_bins(ref["exit"].astype(np.float64), e, nt, nph)
    near = edge_distance(ref["exit"].astype(np.float64), e, nt, nph) < 1e-3
    mism = (gbins != obins) & ref["valid"]
    print(f"  bin mismatches {mism.sum()} (near-edge paths {near.sum()})", flush=True)
    assert not np.any(mism & ~near)
    assert h.shape == (nt, nph) and np.all(h >= 0)
    assert abs(h.sum() - 1.0) < 1e-5
    g = res["log10w"].astype(np.float64)
    hown = reference_histogram(g, gbins.astype(np.int64), nt * nph).reshape(nt, nph)
    err = np.abs(h - hown)
    print(f"  own weights: max |dh| {err.max():.3e} (max h {h.max():.3e})", flush=True)
    assert np.all(err <= 3e-5 * hown + 1e-12)
    href = reference_histogram(ref["log10w"], gbins.astype(np.int64), nt * nph).reshape(nt, nph)
    ok = ref["valid"] & np.isfinite(ref["log10w"])
    m = ref["log10w"][ok].max()
    wr = 10.0 ** (ref["log10w"][ok] - m)
    wg = 10.0 ** (g[ok] - m)
    allow = np.abs(wg - wr).sum() / wr.sum()
    l1 = np.abs(h - href).sum()
    print(f"  oracle weights: L1 {l1:.3e} allowance {allow:.3e}", flush=True)
    assert l1 <= 2 * allow + 1e-6
    ctx.close()


def test_exit_histogram_constant_table_counts(hip_device, oracle):
    """Equal path weights: every bin is its path fraction."""
    case = cases.MID
    table, pos32 = cases.setup(case, oracle)
    ctx = capi.Context(0)
    ctx.set_tables(np.full(len(table.table), 0.5, np.float32), table.minK, table.maxK, table.step, table.ds,
                   case.absorption)
    ctx.set_curve(pos32, cases.bc13(case))
    ctx.set_exit_histogram(16, 8)
    paths = 1 << 14
    ctx.run(paths, case.K, seed=3)
    h, b = ctx.exit_histogram(paths)
    counts = np.bincount(b, minlength=128).reshape(16, 8) / paths
    assert np.abs(h - counts).max() < 1e-6
    ctx.close()


def test_exit_histogram_c5_three_tables(hip_device, oracle):
    """N=1024, K=200, the sum of 3 tables (sigma 0.02/0.05/0.1), 256x256 bins.
    Every path meets the per-path rule against the FP64 oracle on the summed
    table (tests/parity.py); the histogram's L1 distance to the oracle-weighted
    one is bounded by the worst per-path weight ratio error eps: each bin is a
    weighted mean of ratios, so sum_b |h_b - h'_b| <= 2 eps / (1 - eps)."""
    case = cases.C5
    ds = cases.ds_of(case)
    tabs = [oracle.build_table(ds, eps=float(np.float32(s) ** 2), numCurvatureSteps=4096) for s in (0.02, 0.05, 0.1)]
    summed = np.sum([t.table.astype(np.float64) for t in tabs], axis=0).astype(np.float32)
    _, pos32 = cases.setup(case, oracle)
    ctx = capi.Context(0)
    ctx.set_tables(np.stack([t.table for t in tabs]), tabs[0].minK, tabs[0].maxK, tabs[0].step, ds, case.absorption)
    ctx.set_curve(pos32, cases.bc13(case))
    ctx.set_exit_histogram(256, 256)
    paths = 1 << 13
    print(f"[hist C5 3 tables paths={paths}]", flush=True)
    res = ctx.run(paths, case.K, seed=case.seed, per_path=True, per_joint=True)
    h, gbins = ctx.exit_histogram(paths)

    class T:
        pass

    t = T()
    t.table, t.ds, t.minK, t.maxK, t.step = summed, ds, tabs[0].minK, tabs[0].maxK, tabs[0].step
    ref = oracle.simulate(pos32.astype(np.float64), case.N, case.L, case.K, paths, t, case.absorption,
                          seed=case.seed, want_exit=True, want_curv=True)
    assert np.array_equal(np.isnan(res["log10w"]), ~ref["valid"])
    valid = ref["valid"] & np.isfinite(ref["log10w"])
    st = path_rule(res["log10w"], ref["log10w"], valid, res["curv"], ref["curv"], t, case.absorption)
    print(f"  per-path rule: {st}", flush=True)
    assert_path_rule(st)
    href = reference_histogram(ref["log10w"], gbins.astype(np.int64), 256 * 256).reshape(256, 256)
    eps = st["max_ratio_err"]
    err = np.abs(h - href).sum()
    print(f"  L1 histogram difference {err:.3e} (bound {2 * eps / (1 - eps) + 1e-5:.3e})", flush=True)
    assert err <= 2 * eps / (1 - eps) + 1e-5
    assert abs(h.sum() - 1.0) < 1e-5
    ctx.close()


def _weight_bins(log10w, lo, hi, nb):
    """The reference's rule (FullExperimentRunnerOptimalPerturb.cpp:562-575) in FP64."""
    w = 10.0 ** log10w
    b = np.floor((w - lo) / (hi - lo) * nb)
    b = np.where(w < lo, 0, np.where(w > hi, nb - 1, b))
    return np.minimum(b, nb - 1).astype(np.int64)


def test_weight_histogram(hip_device, oracle):
    """Path counts per linear weight bin: against the same rule applied to the
    kernel's own per-path weights (paths within 3e-5 of a bin edge in relative
    weight may move one bin: the float log10 output carries ~1e-5), and the
    reference's default range, which every path of this workload falls below."""
    case = cases.MID
    paths = 1 << 14
    ctx = cases.hip_context(case, oracle)
    ref = ctx.run(paths, case.K, seed=case.seed, per_path=True)
    g = ref["log10w"].astype(np.float64)
    ok = np.isfinite(g)
    w = 10.0 ** g[ok]
    lo, hi, nb = np.quantile(w, 0.05), np.quantile(w, 0.95), 64
    ctx.set_weight_histogram(nb, lo, hi, 0.0)
    r = ctx.run(paths, case.K, seed=case.seed)
    counts = ctx.weight_histogram().astype(np.int64)
    assert counts.sum() == r["num_valid"] == ok.sum()
    mine = np.bincount(_weight_bins(g[ok], lo, hi, nb), minlength=nb)
    u = (w - lo) / (hi - lo) * nb
    near = np.abs(u - np.round(u)) < 3e-5 * np.maximum(u, 1.0)
    print(f"[weight histogram] |diff| {np.abs(counts - mine).sum()} near-edge paths {near.sum()}", flush=True)
    assert np.abs(counts - mine).sum() <= 2 * near.sum()
    ctx.set_weight_histogram(500)  # the reference's range: weights ~1e-40 all land in bin 0
    r = ctx.run(paths, case.K, seed=case.seed)
    counts = ctx.weight_histogram()
    assert counts[0] == r["num_valid"] and counts[1:].sum() == 0
    ctx.close()
