"""GPU parity of the plain forward kernel against the FP64 oracle.

Tolerances (FP32 kernel vs FP64 oracle, same Philox draws, same float initial
curve and table):
  * per-path log10 weight: |g - r| <= 1e-5 |r| on every path whose joint
    curvatures all fall in the same table bin as the FP64 curvatures;
  * a path with a joint on the other side of a bin edge than FP64 meets
    |g - r| <= 1e-5 |r| + sum of the jumps of the reference's lerp at those
    joints.  The reference's lerp fraction is in curvature units
    (PathWeighters.cpp:79-82), so the segment weight is discontinuous at every
    bin edge and no finite precision reproduces FP64 across one;
  * the lookup/log stage alone (oracle weight function evaluated on the GPU's
    own curvatures): |g - r| <= 2e-6 |r|, which isolates geometry error from
    weighting error;
  * joint curvatures: max |dk| no larger than the reference's own FP32-position
    arithmetic on the same paths (oracle precision 32), floor 1e-4: FP32 and
    FP64 geometry drift apart as the rotation axes feed back into the state;
  * sum of weights over the same paths: bounded by the worst per-path linear
    error, |10^{dS} - 1| <= max_p |10^{dw_p} - 1| + 1e-6, over subsamples and
    over all 2^22 paths of the C2 workload.
"""
import math

import numpy as np
import pytest

from tests import cases
from tests.lerp import joint_log10w
from tests.parity import assert_path_rule, merge10, path_rule
from twisty_amd import capi

pytestmark = pytest.mark.gpu

REL_PATH = 1e-5
REL_WEIGHT_STAGE = 2e-6
KAPPA_FLOOR = 1e-4


def both_refs(case, oracle, n, **kw):
    """FP64 oracle and the reference's own FP32-position arithmetic, same paths."""
    return (cases.oracle_run(case, oracle, n, want_curv=True, **kw),
            cases.oracle_run(case, oracle, n, want_curv=True, precision=32, **kw))


def _report(name, **kw):
    print(f"[{name}] " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def check_against_oracle(case, oracle, gpu, ref, ref32, table, tag):
    g, r = gpu["log10w"].astype(np.float64), ref["log10w"]
    assert gpu["num_paths"] == len(r)
    assert gpu["num_valid"] == ref["n_valid"] and gpu["num_invalid"] == ref["n_invalid"]
    assert np.array_equal(np.isnan(g), ~ref["valid"])
    # zero-weight paths (a segment weight of exactly 0) are -inf on both sides
    assert np.array_equal(np.isneginf(g), ref["valid"] & np.isneginf(r))
    valid = ref["valid"] & np.isfinite(r)
    # per-joint bins and lerped log weights of both curvature sets
    lg, bg = joint_log10w(gpu["curv"], table, case.absorption)
    lr, br = joint_log10w(ref["curv"], table, case.absorption)
    straddle = bg != br
    with np.errstate(invalid="ignore"):  # invalid (NaN) and zero-weight (-inf) paths are selected out below
        jump = np.where(straddle, np.abs(lg - lr), 0.0).sum(axis=1)
        err = np.abs(g - r)
    clean = valid & ~straddle.any(axis=1)
    scale = np.maximum(np.abs(r), 1e-30)
    rel_clean = (err[clean] / scale[clean]).max() if clean.any() else 0.0
    excess = (err[valid] - REL_PATH * scale[valid] - jump[valid]).max()
    dk = np.abs(gpu["curv"].astype(np.float64) - ref["curv"].astype(np.float64)).max()
    dk_ref32 = np.abs(ref32["curv"].astype(np.float64) - ref["curv"].astype(np.float64)).max()
    # weighting stage on the GPU's own curvatures
    worst_stage = 0.0
    for i in np.flatnonzero(valid)[:512]:
        ok, w = oracle.weight_curvatures(gpu["curv"][i], table, case.absorption)
        assert ok
        worst_stage = max(worst_stage, abs(g[i] - w) / abs(w))
    gv, rv = g[valid], r[valid]
    dS = gpu["log10_sum_w"] - ref["log10_sum_w_exact"]
    bound = np.max(np.abs(10.0 ** (gv - rv) - 1.0)) + 1e-6
    _report(tag, paths=len(r), straddle_paths=int((valid & ~clean).sum()), max_rel_clean=f"{rel_clean:.3e}",
            max_rel_all=f"{(err[valid] / scale[valid]).max():.3e}", max_dkappa=f"{dk:.3e}", ref32_dkappa=f"{dk_ref32:.3e}",
            weight_stage=f"{worst_stage:.3e}", dlog10S=f"{dS:.3e}", agg_bound=f"{bound:.3e}")
    assert rel_clean <= REL_PATH
    assert excess <= 0.0
    assert dk <= max(dk_ref32, KAPPA_FLOOR)
    assert worst_stage <= REL_WEIGHT_STAGE
    assert abs(10.0 ** dS - 1.0) <= bound
    return {"straddle_paths": int((valid & ~clean).sum()), "max_rel_clean": float(rel_clean)}


def test_c2_subsample_per_path(hip_device, oracle):
    case = cases.C2
    table, _ = cases.setup(case, oracle)
    n = 4096
    with cases.hip_context(case, oracle) as ctx:
        gpu = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
        _report("schedule", **ctx.get_schedule())
    ref, ref32 = both_refs(case, oracle, n)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, "C2 2^12")


def test_c1_all_paths(hip_device, oracle):
    case = cases.C1
    table, _ = cases.setup(case, oracle)
    n = 1 << 16
    with cases.hip_context(case, oracle) as ctx:
        gpu = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
    ref, ref32 = both_refs(case, oracle, n)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, "C1 2^16")


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_schedules(hip_device, oracle, lanes):
    case = cases.C2
    table, _ = cases.setup(case, oracle)
    n = 1024
    with cases.hip_context(case, oracle, schedule=(lanes, 0, 0)) as ctx:
        gpu = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
        _report(f"lanes={lanes}", **ctx.get_schedule())
    ref, ref32 = both_refs(case, oracle, n)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, f"C2 lanes={lanes}")


def test_lean_instance_matches_lds_instance(hip_device, oracle):
    """C2 resolves to the lean G=4 instance (768 lanes, table read from global
    memory, draw records exchanged with ds_bpermute).  It computes what the
    512-lane instance with the table and records in LDS computes: per-path
    weights, per-joint curvatures and weight-histogram counts bit-identical,
    with chains and burn-in; aggregates differ only through the grid's
    summation order."""
    case = cases.C2
    n, L = 4096, 16
    out = {}
    for block in (0, 512):
        with cases.hip_context(case, oracle, schedule=(0, block, 0) if block else None) as ctx:
            ctx.set_weight_histogram(500, 1e-90, 1e-80, 0.0)
            ctx.set_burn_in(5)
            r = ctx.run(n, 2, seed=case.seed, chain_length=L, per_path=True, per_joint=True)
            r["hist"] = ctx.weight_histogram()
            r["schedule"] = ctx.get_schedule()
            out[block] = r
        _report(f"block={block}", **out[block]["schedule"])
    assert out[0]["schedule"]["block_size"] == 768 and out[0]["schedule"]["lanes_per_path"] == 4
    a, b = out[0], out[512]
    assert np.array_equal(a["log10w"].view(np.uint32), b["log10w"].view(np.uint32))
    assert np.array_equal(a["curv"].view(np.uint32), b["curv"].view(np.uint32))
    assert np.array_equal(a["hist"], b["hist"]) and a["hist"].sum() > 0
    # FP32 rescaling of the block accumulators: ~1e-7 relative, order dependent
    assert a["num_valid"] == b["num_valid"] and abs(a["log10_sum_w"] - b["log10_sum_w"]) <= 1e-6


@pytest.mark.parametrize("case", [cases.SMALL, cases.MID, cases.C4, cases.BIG], ids=lambda c: c.name)
def test_other_shapes(hip_device, oracle, case):
    table, _ = cases.setup(case, oracle)
    n = 2048
    with cases.hip_context(case, oracle) as ctx:
        gpu = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
    ref, ref32 = both_refs(case, oracle, n)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, case.name)


def test_large_n_1024(hip_device, oracle):
    case = cases.C5
    table, _ = cases.setup(case, oracle)
    n = 4096
    with cases.hip_context(case, oracle) as ctx:
        gpu = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
        _report("C5 schedule", **ctx.get_schedule())
    ref, ref32 = both_refs(case, oracle, n)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, "C5 4096 paths")


def test_large_n_instances_agree(hip_device, oracle):
    """N = 1024 on every G = 32 instance that fits (blocks 256, 320, 352 with
    the table in LDS; 384 with the table in global memory): per-path weights,
    per-joint curvatures and exit bins bit-identical, with chains."""
    case = cases.C5
    n, L = 2048, 4
    out = {}
    for block in (256, 320, 352, 384):
        with cases.hip_context(case, oracle, schedule=(32, block, 0)) as ctx:
            ctx.set_exit_histogram(64, 64)
            r = ctx.run(n, 20, seed=case.seed, chain_length=L, per_path=True, per_joint=True)
            r["hist"], r["bins"] = ctx.exit_histogram(n)
            r["schedule"] = ctx.get_schedule()
            out[block] = r
        _report(f"C5 block={block}", **out[block]["schedule"])
    a = out[256]
    for block in (320, 352, 384):
        b = out[block]
        assert np.array_equal(a["log10w"].view(np.uint32), b["log10w"].view(np.uint32)), block
        assert np.array_equal(a["curv"].view(np.uint32), b["curv"].view(np.uint32)), block
        assert np.array_equal(a["bins"], b["bins"]), block
        assert np.allclose(a["hist"], b["hist"], rtol=1e-6, atol=0), block
        assert a["num_valid"] == b["num_valid"] and abs(a["log10_sum_w"] - b["log10_sum_w"]) <= 1e-6


def test_path_offset(hip_device, oracle):
    case = cases.C2
    table, _ = cases.setup(case, oracle)
    off, n = 1_000_003, 777
    with cases.hip_context(case, oracle) as ctx:
        gpu = ctx.run(n, case.K, seed=case.seed, path_offset=off, per_path=True, per_joint=True)
    ref, ref32 = both_refs(case, oracle, n, path_offset=off)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, "C2 offset")


def test_chain_mode(hip_device, oracle):
    """K=1 with the state carried through a chain: the reference's own loop."""
    case = cases.C2
    table, _ = cases.setup(case, oracle)
    L, chains = 64, 48
    n = L * chains
    with cases.hip_context(case, oracle) as ctx:
        gpu = ctx.run(n, 1, seed=case.seed, chain_length=L, per_path=True, per_joint=True)
    ref, ref32 = both_refs(case, oracle, n, chain_length=L, K=1)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, "C2 chain K=1 L=64")


@pytest.mark.parametrize("K,L,burn", [(1, 32, 100), (3, 1, 7)])
def test_burn_in(hip_device, oracle, K, L, burn):
    """pathsToSkip: burn-in paths open every chain (path offset a multiple of L,
    so the chain index of the draws is global)."""
    case = cases.C2
    table, _ = cases.setup(case, oracle)
    n, off = 2048, 1000 * L
    with cases.hip_context(case, oracle) as ctx:
        ctx.set_burn_in(burn)
        gpu = ctx.run(n, K, seed=case.seed, chain_length=L, path_offset=off, per_path=True, per_joint=True)
        ctx.set_burn_in(0)
        plain = ctx.run(n, K, seed=case.seed, chain_length=L, path_offset=off, per_path=True)
    ref, ref32 = both_refs(case, oracle, n, chain_length=L, K=K, path_offset=off, burn_in=burn)
    check_against_oracle(case, oracle, gpu, ref, ref32, table, f"C2 burn-in {burn} K={K} L={L}")
    assert not np.array_equal(gpu["log10w"], plain["log10w"])


def test_invalid_and_zero_weights(hip_device, oracle):
    """NaN table entries make paths invalid; zero entries give zero weight."""
    case = cases.MID
    table, pos32 = cases.setup(case, oracle)
    bad = table.table.copy()
    bad[40:60] = np.nan
    bad[200:260] = 0.0
    bad_t = oracle.Table(bad, [table.ds, table.minK, table.maxK, table.step, table.minW, table.maxW])
    n = 2048
    with capi.Context(0) as ctx:
        ctx.set_tables(bad, table.minK, table.maxK, table.step, table.ds, case.absorption)
        ctx.set_curve(pos32, cases.bc13(case))
        gpu = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
    ref, ref32 = (oracle.simulate(pos32.astype(np.float64), case.N, case.L, case.K, n, bad_t, case.absorption,
                                  seed=case.seed, want_curv=True, precision=prec) for prec in (64, 32))
    _report("invalid", gpu_invalid=gpu["num_invalid"], ref_invalid=ref["n_invalid"],
            zero_weight=int(np.isneginf(ref["log10w"]).sum()))
    assert gpu["num_invalid"] == ref["n_invalid"] > 0
    check_against_oracle(case, oracle, gpu, ref, ref32, bad_t, "bad table")


def test_all_zero_table(hip_device, oracle):
    case = cases.SMALL
    table, pos32 = cases.setup(case, oracle)
    with capi.Context(0) as ctx:
        ctx.set_tables(np.zeros_like(table.table), table.minK, table.maxK, table.step, table.ds, case.absorption)
        ctx.set_curve(pos32, cases.bc13(case))
        gpu = ctx.run(512, case.K, seed=case.seed, per_path=True)
    assert gpu["num_valid"] == 512 and gpu["num_invalid"] == 0
    assert gpu["log10_sum_w"] == -math.inf
    assert np.all(gpu["log10w"] == -np.inf)


def test_c2_all_paths_same_draws_against_oracle(hip_device, oracle):
    """Every one of the 2^22 C2 paths against the FP64 oracle on the same draws
    (16 chunks of 2^18 path ids through path offsets): each path meets the
    per-path rule, each chunk's per-path weights are bit-identical to the same
    paths of one full 2^22 run, and that run's sum of weights matches the
    oracle's sum over all 2^22 paths within the bound the per-path errors imply
    (tests/parity.py).  The Monte Carlo standard error of the estimate is
    reported beside it: the agreement is far tighter than sigma_MC."""
    case = cases.C2
    table, _ = cases.setup(case, oracle)
    n, chunk = 1 << 22, 1 << 18
    S_ref, worst, n_valid, n_straddle, rel_clean = -math.inf, 0.0, 0, 0, 0.0
    with cases.hip_context(case, oracle) as ctx:
        full = ctx.run(n, case.K, seed=case.seed, per_path=True)
        for c in range(n // chunk):
            off = c * chunk
            part = ctx.run(chunk, case.K, seed=case.seed, path_offset=off, per_path=True, per_joint=True)
            assert np.array_equal(part["log10w"], full["log10w"][off:off + chunk], equal_nan=True)
            ref = cases.oracle_run(case, oracle, chunk, path_offset=off, want_curv=True)
            assert np.array_equal(np.isnan(part["log10w"]), ~ref["valid"])
            valid = ref["valid"] & np.isfinite(ref["log10w"])
            st = path_rule(part["log10w"], ref["log10w"], valid, part["curv"], ref["curv"], table, case.absorption)
            assert_path_rule(st)
            S_ref = merge10(S_ref, ref["log10_sum_w_exact"])
            worst = max(worst, st["max_ratio_err"])
            n_valid += ref["n_valid"]
            n_straddle += st["straddle_paths"]
            rel_clean = max(rel_clean, st["max_rel_clean"])
    dS = full["log10_sum_w"] - S_ref
    mean, se = capi.estimator(full)
    _report("C2 2^22 same draws", straddle_paths=n_straddle, max_rel_clean=f"{rel_clean:.3e}",
            max_path_ratio_err=f"{worst:.3e}", dlog10S=f"{dS:.3e}", rel_sum_err=f"{10.0 ** dS - 1.0:.3e}",
            sigma_MC_rel=f"{se:.3e}", log10_avg_w=f"{mean:.6f}")
    assert full["num_valid"] == n_valid
    assert abs(10.0 ** dS - 1.0) <= worst + 1e-6


def test_full_size_c2_properties(hip_device, oracle):
    """2^22 paths: counts, determinism and split invariance of the aggregate."""
    case = cases.C2
    n = 1 << 22
    with cases.hip_context(case, oracle) as ctx:
        a = ctx.run(n, case.K, seed=case.seed)
        b = ctx.run(n, case.K, seed=case.seed)
        h1 = ctx.run(n // 2, case.K, seed=case.seed)
        h2 = ctx.run(n // 2, case.K, seed=case.seed, path_offset=n // 2)
        fwd_ms, tot_ms = ctx.last_timing()
    assert a["num_paths"] == n and a["num_valid"] == n and a["num_invalid"] == 0
    assert a["log10_sum_w"] == b["log10_sum_w"] and a["log10_sum_w2"] == b["log10_sum_w2"]
    merged = merge10(h1["log10_sum_w"], h2["log10_sum_w"])
    _report("C2 2^22", log10S=a["log10_sum_w"], merged=merged, half_ms=f"{fwd_ms:.3f}")
    assert abs(merged - a["log10_sum_w"]) <= 1e-6


LONGEST = cases.Case("N3072", N=3072, L=3.0, K=30, sigma=0.05, bins=2048, numStepsInt=2000)
TOO_LONG = cases.Case("N3136", N=3136, L=3.0, K=30, sigma=0.05, bins=2048, numStepsInt=2000)


def test_size_limits(hip_device, oracle):
    """The longest curve whose state fits one block's LDS (N = 3072: 64 lanes
    per path, 48 segment rows per lane, block 256) runs and matches the oracle
    path by path; one more row per lane (N = 3136) is reported as unsupported
    before any launch.  No paths, no perturbations and more than 2^40 paths
    are argument errors."""
    table, _ = cases.setup(LONGEST, oracle)
    n = 64
    with cases.hip_context(LONGEST, oracle) as ctx:
        sched = ctx.get_schedule()
        _report("N3072 schedule", **sched)
        assert sched["lanes_per_path"] == 64
        gpu = ctx.run(n, LONGEST.K, seed=LONGEST.seed, per_path=True, per_joint=True)
        for bad in [(0, 1), (16, 0), ((1 << 40) + 1, 1)]:
            with pytest.raises(capi.TwistyHipError) as e:
                ctx.run(*bad)
            assert e.value.code == capi.ERR_ARG
    ref, ref32 = both_refs(LONGEST, oracle, n)
    check_against_oracle(LONGEST, oracle, gpu, ref, ref32, table, "N3072 64 paths")
    with cases.hip_context(TOO_LONG, oracle) as ctx:
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.get_schedule()
        assert e.value.code == capi.ERR_UNSUPPORTED
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.run(16, 1)
        assert e.value.code == capi.ERR_UNSUPPORTED


def test_error_codes(hip_device, oracle):
    case = cases.SMALL
    with capi.Context(0) as ctx:
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.N = case.N
            ctx.run(16, 1)
        assert e.value.code == capi.ERR_STATE
        table, pos32 = cases.setup(case, oracle)
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.set_tables(np.full(8, np.inf, np.float32), 0.0, 1.0, 0.1, 0.1, 0.0)
        assert e.value.code == capi.ERR_ARG
        ctx.set_tables(table.table, table.minK, table.maxK, table.step, table.ds, case.absorption)
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.set_curve(np.zeros((3, 3), np.float32), cases.bc13(case))
        assert e.value.code == capi.ERR_ARG
        ctx.set_curve(pos32, cases.bc13(case))
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.run(10, 1, chain_length=4, path_offset=2)
        assert e.value.code == capi.ERR_ARG
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.set_schedule(3, 0, 0)
        assert e.value.code == capi.ERR_ARG
        r = ctx.run(64, case.K)
        assert r["num_paths"] == 64
