"""Independent-path estimator (twisty_hip_run_generated) against the FP64 oracle.

ExperimentBase::MSegmentPathGenerationMC (Experiments/ExperimentBase.cpp:
1246-1386): both sides build every path from the same Philox words by the
recursion of :617-983 and weigh it with the same table.  Every path is held to
the forward pass's rule (check_generated): 1e-5 relative, or, where a joint's
FP32 and FP64 curvatures fall in different table bins, 1e-5 plus the lerp jumps
there.  One exception is allowed, and only where it can occur: a path whose
recursion made a branch decision (a 0.001 tolerance, a sphere-meeting test, a
frame's axis limit) within MARGIN of its threshold in FP64 may take the other
branch in FP32 (other geometry, or valid against invalid); such paths must
stay under 0.1%.  The weighting stage is checked on the kernel's own
curvatures, and path offsets and the aggregate follow.
"""
import math
import types

import numpy as np
import pytest

from tests.lerp import joint_log10w
from tests.parity import assert_path_rule, path_rule
from twisty_amd import capi, host

pytestmark = pytest.mark.gpu

REL_STAGE = 2e-6     # log10 w against the sum of its joints' lerped log weights
# A branch decision this close to its threshold in FP64 can go the other way in
# FP32: the geometry's points carry errors of at most a few 1e-7 (hardware
# reciprocal and square root, about 1 ulp each, over the recursion's levels).
MARGIN = 1e-6
# Largest FP32-vs-FP64 joint curvature difference on paths whose branches all
# agree (hardware-precision geometry: tangent errors of ~1e-6 over ds ~ 0.02).
KAPPA_TOL = 2e-2


def check_generated(tag, gpu, ref, tables, tab_idx, absorption):
    """The per-path rule (tests/parity.py) on every path whose branch decisions
    are not marginal; tables[tab_idx[p]] weighs path p."""
    g, r = gpu["log10w"].astype(np.float64), ref["log10w"]
    n = len(r)
    gv, rv = ~np.isnan(g), ref["valid"]
    marginal = ref["margin"] < MARGIN
    # valid against invalid, and zero weight against nonzero, only where a decision is marginal
    assert np.all((gv == rv) | marginal), np.flatnonzero((gv != rv) & ~marginal)[:10]
    both = gv & rv
    assert np.all((np.isneginf(g) == np.isneginf(r)) | ~both | marginal)
    fin = both & np.isfinite(r) & np.isfinite(g)
    dk = np.where(fin[:, None], np.abs(gpu["curv"].astype(np.float64) - ref["curv"].astype(np.float64)), 0.0).max(1)
    apart = fin & (dk > KAPPA_TOL)
    # geometry apart only where a decision is marginal
    assert not np.any(apart & ~marginal), np.flatnonzero(apart & ~marginal)[:10]
    held = fin & ~apart
    worst = {"straddle_paths": 0, "max_rel_clean": 0.0, "max_excess": -1.0}
    stage = 0.0
    for t in np.unique(tab_idx):
        sel = tab_idx == t
        st = path_rule(g[sel], r[sel], held[sel], gpu["curv"][sel], ref["curv"][sel], tables[t], absorption)
        assert_path_rule(st)
        worst["straddle_paths"] += st["straddle_paths"]
        worst["max_rel_clean"] = max(worst["max_rel_clean"], st["max_rel_clean"])
        worst["max_excess"] = max(worst["max_excess"], st["max_excess"])
        # weighting stage: the kernel's log10 w against its own curvatures' joint weights
        gf = (gv & np.isfinite(g))[sel]
        if gf.any():
            lg, _ = joint_log10w(gpu["curv"][sel][gf], tables[t], absorption)
            stage = max(stage, float((np.abs(g[sel][gf] - lg.sum(1)) / np.abs(g[sel][gf])).max()))
    # the exceptions taken: validity, zero weight or geometry apart
    taken = (gv != rv) | (both & (np.isneginf(g) != np.isneginf(r))) | apart
    print(f"[{tag}] paths {n}, valid gpu {int(gv.sum())} oracle {int(rv.sum())}, marginal {int(marginal.sum())}, "
          f"exceptions {int(taken.sum())} (margins {np.sort(ref['margin'][taken])[:5]}), {worst}, "
          f"max dkappa held {dk[held].max() if held.any() else 0.0:.2e}, stage {stage:.2e}", flush=True)
    assert stage <= REL_STAGE
    assert taken.sum() <= max(2, n // 1000)


CASES = [(6, 1.4, ((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1))),
         (10, 2.5, ((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1))),
         (14, 1.4, ((0, 0, 0), (1, 0, 0), (0.5, 0.3, 0.2), (0, 1, 0))),
         (26, 1.8, ((0, 0, 0), (0, 0, 1), (1, 0, 0.2), (0.6, 0.8, 0.0))),
         (66, 1.6, ((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1))),
         (130, 2.0, ((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1)))]


def _setup(oracle, N, L, geom):
    s, sd, e, ed = (np.array(v, np.float64) for v in geom)
    sd, ed = sd / np.linalg.norm(sd), ed / np.linalg.norm(ed)
    bc13 = host.boundary13(s, sd, e, ed, L)
    bc12 = bc13[:12].astype(np.float64)
    ds = float(np.float32(L) / np.float32(N))
    t = oracle.build_table(ds, eps=float(np.float32(0.08) ** 2), numStepsInt=2000, numCurvatureSteps=2048)
    ctx = capi.Context(0)
    ctx.set_tables(t.table, t.minK, t.maxK, t.step, t.ds, 0.004)
    ctx.set_curve(host.initial_arc(bc13, N), bc13)
    return ctx, t, bc12


@pytest.mark.parametrize("N,L,geom", CASES, ids=[f"N{c[0]}" for c in CASES])
def test_generated_paths_match_oracle(hip_device, oracle, N, L, geom):
    n = 4096
    ctx, t, bc12 = _setup(oracle, N, L, geom)
    g = ctx.run_generated(n, seed=7, per_path=True, per_joint=True)
    ref = oracle.generate(bc12, L, N, n, t, 0.004, seed=7, want_joints=True)
    gw = g["log10w"].astype(np.float64)
    check_generated(f"generated N={N}", g, ref, [t], np.zeros(n, np.int64), 0.004)
    assert g["num_paths"] == n and g["num_valid"] + g["num_invalid"] == n
    # aggregates over the same valid paths
    lse = lambda v: v.max() + math.log10(np.sum(10.0 ** (v - v.max())))
    assert abs(lse(gw[~np.isnan(gw)]) - g["log10_sum_w"]) < 1e-5 * abs(g["log10_sum_w"])
    assert g["min_log10_w"] == pytest.approx(np.nanmin(gw), rel=1e-6)
    assert g["max_log10_w"] == pytest.approx(np.nanmax(gw), rel=1e-6)
    assert abs(g["log10_sum_w"] - ref["log10_sum_w"]) < 1e-4 * abs(ref["log10_sum_w"])
    ctx.close()


def test_generated_path_offset_and_determinism(hip_device, oracle):
    N, L, geom = CASES[4]
    ctx, t, bc12 = _setup(oracle, N, L, geom)
    full = ctx.run_generated(3000, seed=3, per_path=True)
    part = ctx.run_generated(1000, seed=3, path_offset=1500, per_path=True)
    again = ctx.run_generated(3000, seed=3, per_path=True)
    a, b = full["log10w"][1500:2500], part["log10w"]
    assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    assert np.array_equal(full["log10w"], again["log10w"], equal_nan=True)
    ctx.close()


def test_generated_rejects_unsupported_segment_counts(hip_device, oracle):
    N, L = 64, 1.5
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    ds = float(np.float32(L) / np.float32(N))
    t = oracle.build_table(ds, eps=0.0025, numStepsInt=2000, numCurvatureSteps=1024)
    ctx = capi.Context(0)
    ctx.set_tables(t.table, t.minK, t.maxK, t.step, t.ds, 0.004)
    ctx.set_curve(host.initial_arc(bc13, N), bc13)
    with pytest.raises(RuntimeError, match="numSegments must be 2 \\+ 2\\^k"):
        ctx.run_generated(100)
    ctx.close()


def test_generated_full_size_properties(hip_device, oracle):
    """2^22 paths at N=66: the partials equal the log-sum-exp of the per-path
    weights, and the first 20000 paths (the same draws in the oracle: a path's
    draws depend only on its index) meet the per-path rule against the FP64
    oracle."""
    N, L, geom = CASES[4]
    ctx, t, bc12 = _setup(oracle, N, L, geom)
    n = 1 << 22
    g = ctx.run_generated(n, seed=1, per_path=True)
    ms = ctx.last_timing()[0]
    w = g["log10w"].astype(np.float64)
    ok = ~np.isnan(w)
    m = w[ok].max()
    lse = m + math.log10(np.sum(10.0 ** (w[ok] - m)))
    print(f"[generated 2^22 N=66] {ms:.3f} ms, {n * (N - 1) / (ms / 1e3):.3e} joint-weights/s, valid "
          f"{ok.mean():.4f}", flush=True)
    assert g["num_valid"] == int(ok.sum()) and abs(lse - g["log10_sum_w"]) < 1e-6 * abs(lse)
    n_ref = 20000
    ref = oracle.generate(bc12, L, N, n_ref, t, 0.004, seed=1, want_joints=True)
    head = ctx.run_generated(n_ref, seed=1, per_path=True, per_joint=True)
    assert np.array_equal(head["log10w"], g["log10w"][:n_ref], equal_nan=True)
    check_generated(f"generated 2^22 N={N}, first {n_ref}", head, ref, [t], np.zeros(n_ref, np.int64), 0.004)
    ctx.close()


def test_host_path_generation_mc(hip_device, oracle):
    """The C++ host estimator on a config: Result = sum of valid 10^(w + norm) / num_valid
    (ExperimentBase.cpp:1364-1384), min and max with the normaliser."""
    N, L = 26, 1.8
    cfg = {"experiment": {
        "experimentParams": {
            "pathsToGenerate": 50000, "pathsToSkip": 0, "numSegments": N, "name": "gen", "experimentDir": "/tmp",
            "maxPerturbThreads": 0, "maxWeightThreads": 0, "outputBigFloatWeights": False,
            "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
            "random": {"bootstrapSeed": 1, "perturbSeed": 1},
            "weighting": {"weightFunction": 0, "mu": 1.0, "eps": float(np.float32(0.08) ** 2), "numStepsInt": 2000,
                          "numCurvatureSteps": 2048, "absorption": 0.004, "scatter": 0.1}},
        "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1], "endDir": [0, 0, 1],
                            "arclength": L}}}
    r = host.path_generation_mc(cfg, seed=11, norm_log10=2.5)
    ctx, t, bc12 = _setup(oracle, N, L, CASES[0][2])
    g = ctx.run_generated(50000, seed=11, per_path=True)
    w = g["log10w"][~np.isnan(g["log10w"])].astype(np.float64)
    m = w.max()
    want = m + math.log10(np.sum(10.0 ** (w - m))) + 2.5 - math.log10(len(w))
    print(f"[path_generation_mc] {r}", flush=True)
    assert r["num_total"] == 50000 and r["num_valid"] == len(w)
    assert abs(r["log10_total_weight"] - want) < 1e-6 * abs(want)
    assert r["min_log10_w"] == pytest.approx(w.min() + 2.5, rel=1e-6)
    assert r["max_log10_w"] == pytest.approx(w.max() + 2.5, rel=1e-6)
    ctx.close()


def test_generated_multi_ds_matches_oracle(hip_device, oracle, tmp_path):
    """Multi-ds form against the oracle: per-path ds, nearest cached table
    (CachedMultiArclengthWeightLookupTable from the host), curvature with
    arclength ds N, the table's own absorption."""
    N = 34
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0.4, 0, 1), (0.6, 0, 0.8), 2.0)
    bc13[9:12] /= np.linalg.norm(bc13[9:12])
    tabs, _, _ = host.cached_rt_tables(0.035, 0.08, 12, 1.0, float(np.float32(0.08) ** 2), 0.1, 0.004,
                                       numStepsInt=2000, numCurvatureSteps=1024, cache_root=tmp_path)
    tables = np.stack([t.table for t in tabs])
    meta = np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32)
    ctx = capi.Context(0)
    ctx.set_curve(host.initial_arc(bc13, N), bc13)
    ctx.set_multi_tables(tables, meta, 0.004, N)
    lo, hi = 0.045, 0.075
    n = 4096
    g = ctx.run_generated_multi(n, lo, hi, seed=5, per_path=True, per_joint=True)
    ref = oracle.generate_multi(bc13[:12].astype(np.float64), N, n, tables, meta, 0.004, lo, hi, seed=5)
    views = [types.SimpleNamespace(table=tables[i], ds=float(meta[i, 0]), minK=float(meta[i, 1]),
                                   maxK=float(meta[i, 2]), step=float(meta[i, 3])) for i in range(len(tables))]
    check_generated(f"generated multi-ds N={N}", g, ref, views, ref["table"].astype(np.int64), 0.004)
    assert abs(g["log10_sum_w"] - ref["log10_sum_w"]) < 1e-4 * abs(ref["log10_sum_w"])
    ctx.close()


def test_host_path_generation_mc_multi(hip_device, tmp_path):
    """The C++ multi-arclength overload on a config equals the device call it makes."""
    N, L = 34, 2.0
    cfg = {"experiment": {
        "experimentParams": {
            "pathsToGenerate": 20000, "pathsToSkip": 0, "numSegments": N, "name": "gen", "experimentDir": "/tmp",
            "maxPerturbThreads": 0, "maxWeightThreads": 0, "outputBigFloatWeights": False,
            "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
            "random": {"bootstrapSeed": 1, "perturbSeed": 1},
            "weighting": {"weightFunction": 0, "mu": 1.0, "eps": float(np.float32(0.08) ** 2), "numStepsInt": 2000,
                          "numCurvatureSteps": 1024, "absorption": 0.004, "scatter": 0.1}},
        "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1], "endDir": [0, 0, 1],
                            "arclength": L}}}
    r = host.path_generation_mc_multi(cfg, 0.03, 0.08, 11, 0.07, seed=3, cache_root=tmp_path)
    tabs, loaded, _ = host.cached_rt_tables(0.03, 0.08, 11, 1.0, float(np.float32(0.08) ** 2), 0.1, 0.004,
                                            numStepsInt=2000, numCurvatureSteps=1024, cache_root=tmp_path)
    assert loaded
    min_ds = float(np.float32(np.float32(1.0) / np.float32(N)))  # the quadratic's root is the end distance here
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 0.07 * N)
    ctx = capi.Context(0)
    ctx.set_curve(np.stack([np.zeros(N + 1), np.zeros(N + 1), np.linspace(0, 1, N + 1)], 1).astype(np.float32), bc13)
    ctx.set_multi_tables(np.stack([t.table for t in tabs]), np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs],
                                                                     np.float32), 0.004, N)
    g = ctx.run_generated_multi(20000, min_ds, 0.07, seed=3)
    ctx.close()
    want = g["log10_sum_w"] - math.log10(g["num_valid"])
    print(f"[path_generation_mc_multi] {r}", flush=True)
    assert r["num_total"] == 20000 and r["num_valid"] == g["num_valid"]
    assert abs(r["log10_total_weight"] - want) < 1e-5 * abs(want)


def test_host_path_generation_mc_chains(hip_device, tmp_path):
    """The C++ two-table overload on a config equals the device call it makes:
    ceil(paths / chainLength) chains, minDs the boundary's minimum arclength
    over N, the cached tables with absorption + scatter."""
    N, L, paths, length = 34, 2.0, 20000, 64
    cfg = {"experiment": {
        "experimentParams": {
            "pathsToGenerate": paths, "pathsToSkip": 0, "numSegments": N, "name": "gen", "experimentDir": "/tmp",
            "maxPerturbThreads": 0, "maxWeightThreads": 0, "outputBigFloatWeights": False,
            "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
            "random": {"bootstrapSeed": 1, "perturbSeed": 1},
            "weighting": {"weightFunction": 0, "mu": 1.0, "eps": float(np.float32(0.08) ** 2), "numStepsInt": 2000,
                          "numCurvatureSteps": 1024, "absorption": 0.004, "scatter": 0.1}},
        "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1], "endDir": [0, 0, 1],
                            "arclength": L}}}
    # the reference returns nothing unless maxDs >= minDs (ExperimentBase.cpp:1639-1642), although its
    # draws then span [minDs, minDs + maxDs]
    max_ds = 0.03
    r = host.path_generation_mc_chains(cfg, 0.025, 0.06, 11, max_ds, chain_length=length, seed=3, norm_log10=1.5,
                                       cache_root=tmp_path)
    tabs, loaded, _ = host.cached_rt_tables(0.025, 0.06, 11, 1.0, float(np.float32(0.08) ** 2), 0.1, 0.004,
                                            numStepsInt=2000, numCurvatureSteps=1024, cache_root=tmp_path)
    assert loaded
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    min_ds = float(np.float32(host.minimum_arclength(bc13, N)) / np.float32(N))
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    chains = (paths + length - 1) // length
    ctx = capi.Context(0)
    ctx.set_curve(np.stack([np.zeros(N + 1), np.zeros(N + 1), np.linspace(0, 1, N + 1)], 1).astype(np.float32), bc13)
    ctx.set_multi_tables(np.stack([t.table for t in tabs]), np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs],
                                                                     np.float32), ext, N)
    g = ctx.run_generated_chains(chains, length, min_ds, max_ds, ext, norm_log10=1.5, seed=3)
    ctx.close()
    want = g["log10_sum_w"] + 1.5 - math.log10(g["num_valid"])
    print(f"[path_generation_mc_chains] {r}", flush=True)
    assert r["num_total"] == chains * length and r["num_valid"] == g["num_valid"]
    assert abs(r["log10_total_weight"] - want) < 1e-9 * abs(want)


def test_generated_chains_match_oracle(hip_device, oracle, tmp_path):
    """The two-table Metropolis-Hastings form (ExperimentBase.cpp:1599-1818)
    against the oracle on the same draws: every proposal by the per-path rule
    (truncated-exponential ds, nearest cached table, loss absorption + scatter
    on every joint; the sampling weight, a function of ds alone, taken off
    both sides first); every chain step against a replay of the acceptance rule
    on the kernel's own proposal weights and the draws u; the partials against
    the chain weights they sum; and the chains against the oracle's."""
    N = 34
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0.4, 0, 1), (0.6, 0, 0.8), 2.0)
    bc13[9:12] /= np.linalg.norm(bc13[9:12])
    lo = np.float32(host.minimum_arclength(bc13, N) / N)
    hi = np.float32(0.03)
    tabs, _, _ = host.cached_rt_tables(float(lo) - 0.001, float(lo + hi) + 0.001, 12, 1.0,
                                       float(np.float32(0.08) ** 2), 0.1, 0.004, numStepsInt=2000,
                                       numCurvatureSteps=1024, cache_root=tmp_path)
    tables = np.stack([t.table for t in tabs])
    meta = np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32)
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    C, L, norm = 256, 16, 2.0
    ctx = capi.Context(0)
    ctx.set_curve(host.initial_arc(bc13, N), bc13)
    ctx.set_multi_tables(tables, meta, ext, N)
    g = ctx.run_generated_chains(C, L, float(lo), float(hi), ext, norm_log10=norm, seed=9, chain_offset=5,
                                 per_path=True, per_joint=True)
    ref = oracle.generate_chains(bc13[:12].astype(np.float64), N, C, L, tables, meta, ext, float(lo), float(hi),
                                 norm_log10=norm, seed=9, chain_offset=5)
    n = C * L
    assert g["num_paths"] == n and g["num_valid"] + g["num_invalid"] == n
    # proposals: the per-path rule, the sampling weight taken off both sides
    sw = np.array([oracle.chain_log10_sampling_weight(ext, float(d), float(lo), N) for d in ref["ds"]])
    views = [types.SimpleNamespace(table=tables[i], ds=float(meta[i, 0]), minK=float(meta[i, 1]),
                                   maxK=float(meta[i, 2]), step=float(meta[i, 3])) for i in range(len(tables))]
    gp = {"log10w": (g["log10w"].astype(np.float64) - sw), "curv": g["curv"]}
    rp = dict(ref, log10w=ref["log10w"] - sw)
    check_generated(f"generated chains N={N}", gp, rp, views, ref["table"].astype(np.int64), ext)
    # the chains: the acceptance rule replayed on the kernel's proposals and draws
    w = g["log10w"].astype(np.float64).reshape(C, L) + norm
    u = ref["u"].reshape(C, L)
    ch = g["chain"].astype(np.float64).reshape(C, L) + norm
    close, resync, total, nv = 0, 0, -math.inf, 0
    for c in range(C):
        cur = None
        for s in range(L):
            if not np.isnan(w[c, s]):
                if cur is None:
                    cur = w[c, s]
                else:
                    ratio = 10.0 ** w[c, s] / 10.0 ** cur
                    take = float(np.float32(u[c, s])) <= ratio
                    if abs(u[c, s] - ratio) <= 1e-5 * ratio:  # the kernel tests its FP64 weight, not the float
                        close += 1
                        take = ch[c, s] == w[c, s]
                    if take:
                        cur = w[c, s]
                nv += 1
                m = max(total, cur)
                total = m + math.log10(10.0 ** (total - m) + 10.0 ** (cur - m))
            if cur is None:
                assert np.isnan(ch[c, s]), (c, s)
            else:
                assert ch[c, s] == cur, (c, s, ch[c, s], cur)
    assert nv == g["num_valid"] and close <= 2
    assert abs(g["log10_sum_w"] + norm - total) < 1e-6 * abs(total)
    # against the oracle's chains: the same decision (valid or not, proposal
    # taken or not) at every step until one falls within the two sides' weight
    # difference (the oracle's acceptance margin) or a marginal geometry
    gv, rv = ~np.isnan(w), ref["valid"].reshape(C, L)
    gtake = gv & (ch == w)
    rtake = rv & np.isclose(ref["chain"].reshape(C, L), ref["log10w"].reshape(C, L), rtol=1e-13, atol=0)
    am, gm = ref["accept_margin"].reshape(C, L), ref["margin"].reshape(C, L)
    diverged = 0
    for c in range(C):
        for s in range(L):
            if gv[c, s] != rv[c, s]:
                assert gm[c, s] < MARGIN, (c, s)
            elif gtake[c, s] != rtake[c, s]:
                assert am[c, s] < 2e-2, (c, s, am[c, s])
            else:
                continue
            diverged += 1
            break
    print(f"[generated chains] {n} proposals, valid {g['num_valid']}, close replay decisions {close}, chains "
          f"leaving the oracle's {diverged} of {C}", flush=True)
    assert diverged <= C // 10
    ctx.close()


def test_generated_chains_argument_errors(hip_device, tmp_path):
    """The chains entry point refuses, before any launch: an extinction other
    than the absorption coefficient the cached tables were given, bad ds bounds,
    an empty or oversized chain set, and a run before the tables."""
    N = 34
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 2.0)
    tabs, _, _ = host.cached_rt_tables(0.025, 0.06, 4, 1.0, float(np.float32(0.08) ** 2), 0.1, 0.004,
                                       numStepsInt=500, numCurvatureSteps=256, cache_root=tmp_path)
    tables = np.stack([t.table for t in tabs])
    meta = np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32)
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    ctx = capi.Context(0)
    ctx.set_curve(host.initial_arc(bc13, N), bc13)
    with pytest.raises(capi.TwistyHipError) as e:
        ctx.run_generated_chains(8, 8, 0.03, 0.02, ext)
    assert e.value.code == capi.ERR_STATE  # no cached tables yet
    ctx.set_multi_tables(tables, meta, 0.004, N)
    for args, code in [((8, 8, 0.03, 0.02, ext), capi.ERR_ARG),        # tables given another coefficient
                       ((8, 8, 0.03, 0.02, 0.004), None),             # matching: runs
                       ((8, 8, 0.0, 0.02, 0.004), capi.ERR_ARG),      # minDs
                       ((8, 8, 0.03, -1.0, 0.004), capi.ERR_ARG),     # maxDs
                       ((8, 8, 0.03, 0.02, 0.0), capi.ERR_ARG),       # extinction
                       ((0, 8, 0.03, 0.02, 0.004), capi.ERR_ARG),     # no chains
                       ((8, 0, 0.03, 0.02, 0.004), capi.ERR_ARG),     # empty chains
                       ((1 << 40, 2, 0.03, 0.02, 0.004), capi.ERR_ARG)]:  # more than 2^40 proposals
        if code is None:
            r = ctx.run_generated_chains(*args, seed=2)
            assert r["num_paths"] == 64 and r["num_valid"] + r["num_invalid"] == 64
        else:
            with pytest.raises(capi.TwistyHipError) as e:
                ctx.run_generated_chains(*args)
            assert e.value.code == code, args
    ctx.close()


@pytest.mark.gpu
def test_generated_chains_split_equals_whole(hip_device, tmp_path):
    """A chain set run in two parts by chain_offset is the run of the whole
    set: every proposal weight, chain weight and joint curvature bit for bit
    (each chain's draws are keyed by its global index), the path counts add
    up, and the two partial sums combine (log-sum-exp) to the whole run's."""
    N = 34
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0.4, 0, 1), (0.6, 0, 0.8), 2.0)
    bc13[9:12] /= np.linalg.norm(bc13[9:12])
    lo = np.float32(host.minimum_arclength(bc13, N) / N)
    hi = np.float32(0.03)
    tabs, _, _ = host.cached_rt_tables(float(lo) - 0.001, float(lo + hi) + 0.001, 6, 1.0,
                                       float(np.float32(0.08) ** 2), 0.1, 0.004, numStepsInt=1000,
                                       numCurvatureSteps=512, cache_root=tmp_path)
    tables = np.stack([t.table for t in tabs])
    meta = np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32)
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    C, L, cut = 3000, 24, 1234  # the cut is not a multiple of any block's chain count
    ctx = capi.Context(0)
    ctx.set_curve(host.initial_arc(bc13, N), bc13)
    ctx.set_multi_tables(tables, meta, ext, N)
    run = lambda c, off: ctx.run_generated_chains(c, L, float(lo), float(hi), ext, norm_log10=1.5, seed=4,
                                                  chain_offset=off, per_path=True, per_joint=True)
    whole, a, b = run(C, 0), run(cut, 0), run(C - cut, cut)
    for key in ("log10w", "chain", "curv"):
        got = np.concatenate([a[key], b[key]])
        assert np.array_equal(got, whole[key], equal_nan=True), key
    assert a["num_paths"] + b["num_paths"] == whole["num_paths"] == C * L
    assert a["num_valid"] + b["num_valid"] == whole["num_valid"]
    assert a["num_invalid"] + b["num_invalid"] == whole["num_invalid"]
    combined = np.logaddexp(a["log10_sum_w"] * math.log(10), b["log10_sum_w"] * math.log(10)) / math.log(10)
    print(f"[chains split] whole {whole['log10_sum_w']:.12f} parts {combined:.12f} valid {whole['num_valid']}",
          flush=True)
    assert abs(combined - whole["log10_sum_w"]) <= 1e-9 * abs(whole["log10_sum_w"]) + 1e-9
