"""Persistent batched kernel (twisty_hip_batch_*) against plain runs and the oracle.

Each experiment of a batch runs the same per-path code as twisty_hip_run with
the same draws (Philox key = the experiment's seed, path ids 0..n-1), so when
both use the same lanes per path the per-path log10 weights are bit-identical
and the FP64 partials agree to rounding of the merge order (1e-9 relative).
The plain run itself is checked against the oracle in test_gpu_forward; one
experiment is also checked against the oracle directly here.
"""
import math

import numpy as np
import pytest

from tests import cases
from tests.parity import assert_path_rule, path_rule
from twisty_amd import capi, host, workloads

pytestmark = pytest.mark.gpu


def plain_lanes(N):
    G = 1
    while G < 64 and (N + G - 1) // G > 16:
        G <<= 1
    return G


def batch_lanes(N, spl):
    G = 2
    while G < 64 and (N + G - 1) // G > spl:
        G <<= 1
    return G


def plain_run(ctx, x, tables, meta, G=0, per_joint=False):
    """A plain run of one experiment: its table's grid, its own ds (absorption);
    G lanes per path (0: the plain resolver's choice)."""
    m = meta[x["table"]]
    ds = float(np.float32(x["arclength"]) / np.float32(x["N"]))
    ctx.set_tables(tables[x["table"]], float(m[0]), float(m[1]), float(m[2]), ds, float(m[4]))
    ctx.set_curve(x["positions"], x["boundary"])
    ctx.set_schedule(G, 0, 0)
    try:
        return ctx.run(x["paths"], x["K"], seed=x["seed"], per_path=True, per_joint=per_joint)
    finally:
        ctx.set_schedule(0, 0, 0)


def hetero_batch():
    return workloads.batch_experiments(n_exp=14, seed=7, K=30, n_range=(5, 300), paths_range=(200, 6000), n_tables=3,
                                       bins=1024, numStepsInt=2000)


class _Table:
    def __init__(self, table, meta_row):
        self.table = np.asarray(table, np.float32)
        self.minK, self.maxK, self.step, self.ds = (float(v) for v in meta_row[:4])


def check_experiment_against_oracle(ctx, oracle, x, tables, meta, batched_log10w, G, n_check, tag):
    """One experiment of a batch against the FP64 oracle with the per-path rule
    (tests/parity.py): its batched per-path weights are bit-identical to a plain
    run at the same lanes per path, whose per-joint curvatures then meet the
    rule against the oracle's on the same draws (first n_check paths)."""
    m = meta[x["table"]]
    ds = float(np.float32(x["arclength"]) / np.float32(x["N"]))
    ctx.set_tables(tables[x["table"]], float(m[0]), float(m[1]), float(m[2]), ds, float(m[4]))
    ctx.set_curve(x["positions"], x["boundary"])
    ctx.set_schedule(G, 0, 0)
    n = min(int(x["paths"]), n_check)
    plain = ctx.run(n, x["K"], seed=x["seed"], per_path=True, per_joint=True)
    ctx.set_schedule(0, 0, 0)
    assert np.array_equal(plain["log10w"], np.asarray(batched_log10w[:n]), equal_nan=True), tag
    t = _Table(tables[x["table"]], m)
    ref = oracle.simulate(np.asarray(x["positions"], np.float32).astype(np.float64), x["N"], x["arclength"], x["K"], n,
                          t, float(m[4]), seed=x["seed"], want_curv=True)
    assert np.array_equal(np.isnan(plain["log10w"]), ~ref["valid"]), tag
    valid = ref["valid"] & np.isfinite(ref["log10w"])
    st = path_rule(plain["log10w"], ref["log10w"], valid, plain["curv"], ref["curv"], t, float(m[4]))
    print(f"[{tag}] N={x['N']} ds={ds:.4g} G={G} paths={n}: {st}", flush=True)
    assert_path_rule(st)
    return st


def test_batched_matches_plain_runs(hip_device):
    """Every experiment of a heterogeneous batch is bit-identical per path to a
    plain run at the batch's lanes per path; where the plain resolver picks
    another lane count, the two plain runs differ only in the order of the
    group sums and meet the per-path rule (tests/parity.py) against each other
    on their own per-joint curvatures."""
    w = hetero_batch()
    ctx = capi.Context(0)
    b = capi.Batch(ctx, w["experiments"], w["tables"], w["meta"])
    sched = b.get_schedule()
    print(f"[batched hetero] schedule {sched}", flush=True)
    res, pp, ms = b.run(per_path=True)
    other_g = 0
    for x, r, g in zip(w["experiments"], res, pp):
        gb = batch_lanes(x["N"], sched["segments_per_lane"])
        same = plain_run(ctx, x, w["tables"], w["meta"], G=gb, per_joint=gb != plain_lanes(x["N"]))
        a = g.astype(np.float64)
        c = same["log10w"].astype(np.float64)
        assert np.array_equal(a, c, equal_nan=True), f"N={x['N']} G={gb}"
        assert r["num_paths"] == x["paths"] and r["num_valid"] == same["num_valid"]
        assert abs(r["log10_sum_w"] - same["log10_sum_w"]) <= 1e-6 * abs(same["log10_sum_w"]) + 1e-9
        if gb != plain_lanes(x["N"]):
            other_g += 1
            dflt = plain_run(ctx, x, w["tables"], w["meta"], per_joint=True)
            valid = np.isfinite(dflt["log10w"])
            assert np.array_equal(valid, np.isfinite(c))
            st = path_rule(c, dflt["log10w"], valid, same["curv"], dflt["curv"], _Table(w["tables"][x["table"]],
                           w["meta"][x["table"]]), float(w["meta"][x["table"]][4]))
            print(f"  N={x['N']} G {gb} vs {plain_lanes(x['N'])}: {st}", flush=True)
            assert_path_rule(st)
    print(f"  {len(res)} experiments bit-identical to plain runs at the batch's lanes per path "
          f"({other_g} also checked against the plain default lanes); replay {ms:.3f} ms", flush=True)
    b.close()
    ctx.close()


def test_batched_chains_and_burn_in_match_plain_runs(hip_device):
    """Chains and burn-in: every experiment bit-identical per path to
    twisty_hip_run with the same chain length, burn-in and lanes per path."""
    w = workloads.batch_experiments(n_exp=6, seed=11, K=3, n_range=(5, 120), paths_range=(500, 3000), n_tables=2,
                                    bins=1024, numStepsInt=2000)
    shapes = [(1, 0), (16, 0), (16, 20), (1, 5), (250, 3), (7, 1)]
    for x, (L, B) in zip(w["experiments"], shapes):
        x["chain"], x["burn_in"] = L, B
    ctx = capi.Context(0)
    b = capi.Batch(ctx, w["experiments"], w["tables"], w["meta"])
    sched = b.get_schedule()
    res, pp, _ = b.run(per_path=True)
    for x, r, g in zip(w["experiments"], res, pp):
        m = w["meta"][x["table"]]
        ds = float(np.float32(x["arclength"]) / np.float32(x["N"]))
        ctx.set_tables(w["tables"][x["table"]], float(m[0]), float(m[1]), float(m[2]), ds, float(m[4]))
        ctx.set_curve(x["positions"], x["boundary"])
        ctx.set_burn_in(x["burn_in"])
        ctx.set_schedule(batch_lanes(x["N"], sched["segments_per_lane"]), 0, 0)
        ref = ctx.run(x["paths"], x["K"], seed=x["seed"], chain_length=x["chain"], per_path=True)
        ctx.set_schedule(0, 0, 0)
        a, c = g.astype(np.float64), ref["log10w"].astype(np.float64)
        assert np.array_equal(a, c, equal_nan=True), f"N={x['N']} chain={x['chain']} burn={x['burn_in']}"
        assert r["num_valid"] == ref["num_valid"]
        assert abs(r["log10_sum_w"] - ref["log10_sum_w"]) <= 1e-6 * abs(ref["log10_sum_w"]) + 1e-9
    ctx.set_burn_in(0)
    b.close()
    ctx.close()


def test_batched_replays_are_deterministic(hip_device):
    w = hetero_batch()
    ctx = capi.Context(0)
    b = capi.Batch(ctx, w["experiments"], w["tables"], w["meta"])
    first, _, _ = b.run()
    for _ in range(3):
        again, _, _ = b.run()
        assert again == first
    b.close()
    ctx.close()


@pytest.mark.parametrize("sched", [(256, 0, 8, 1024), (512, 1, 16, 333), (256, 2, 16, 4096), (256, 0, 32, 100000),
                                   (256, 0, 16, 4096, 1), (256, 3, 16, 333, 1), (512, 0, 8, 1024, 1),
                                   (256, 0, 16, 4096, 2), (256, 0, 8, 1024, 2), (256, 2, 8, 333, 2),
                                   (256, 3, 32, 4096, 2)])
def test_batched_schedules_agree(hip_device, sched):
    w = hetero_batch()
    ctx = capi.Context(0)
    b = capi.Batch(ctx, w["experiments"], w["tables"], w["meta"])
    base, pb, _ = b.run(per_path=True)
    b.set_schedule(*sched)
    other, po, ms = b.run(per_path=True)
    print(f"[batched schedule {sched}] {ms:.3f} ms", flush=True)
    for x, r0, r1, a, c in zip(w["experiments"], base, other, pb, po):
        assert r0["num_valid"] == r1["num_valid"]
        assert abs(r0["log10_sum_w"] - r1["log10_sum_w"]) <= 1e-5 * abs(r0["log10_sum_w"])
        ok = np.isfinite(a)
        if batch_lanes(x["N"], sched[2]) == batch_lanes(x["N"], 16):
            assert np.array_equal(a[ok], c[ok])
    b.close()
    ctx.close()


def test_batched_unsupported_schedule_rejected(hip_device):
    """A block size the placement is not compiled for, or an unknown placement,
    fails before any launch and the batch keeps its previous schedule."""
    w = hetero_batch()
    ctx = capi.Context(0)
    b = capi.Batch(ctx, w["experiments"], w["tables"], w["meta"])
    before = b.get_schedule()
    with pytest.raises(capi.TwistyHipError) as e:
        b.set_schedule(512, 0, 16, 4096, 2)
    assert e.value.code == capi.ERR_UNSUPPORTED
    with pytest.raises(capi.TwistyHipError) as e:
        b.set_schedule(256, 0, 16, 4096, 3)
    assert e.value.code == capi.ERR_ARG
    assert b.get_schedule() == before
    b.close()
    ctx.close()


def test_batched_experiment_vs_oracle(hip_device, oracle):
    """Two experiments of one batch, every path against the FP64 oracle with the
    per-path rule of the forward tests, chained through a plain run at the
    batch's lanes per path (bit-identical per path)."""
    case = cases.MID
    table, pos32 = cases.setup(case, oracle)
    x = {"N": case.N, "paths": 4000, "K": case.K, "seed": case.seed, "arclength": case.L, "table": 1,
         "positions": pos32, "boundary": cases.bc13(case)}
    bc20 = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 1.2)
    other = dict(x, N=20, positions=host.initial_arc(bc20, 20), boundary=bc20, arclength=1.2, table=0, paths=1500,
                 seed=5)
    t20 = oracle.build_table(float(np.float32(1.2) / np.float32(20)), eps=float(np.float32(case.sigma) ** 2),
                             numStepsInt=case.numStepsInt, numCurvatureSteps=case.bins)
    tables = np.stack([t20.table, table.table])
    meta = np.array([(t20.minK, t20.maxK, t20.step, t20.ds, case.absorption),
                     (table.minK, table.maxK, table.step, table.ds, case.absorption)], np.float32)
    ctx = capi.Context(0)
    b = capi.Batch(ctx, [other, x], tables, meta)
    spl = b.get_schedule()["segments_per_lane"]
    res, pp, _ = b.run(per_path=True)
    b.close()
    for e, y in enumerate((other, x)):
        check_experiment_against_oracle(ctx, oracle, y, tables, meta, pp[e], batch_lanes(y["N"], spl), y["paths"],
                                        f"batched vs oracle exp {e}")
        assert res[e]["num_paths"] == y["paths"]
    ctx.close()


def test_batched_autotune(hip_device):
    w = workloads.batch_experiments(n_exp=64, seed=3, K=20, paths_range=(1 << 8, 1 << 12), bins=1024,
                                    numStepsInt=2000)
    ctx = capi.Context(0)
    b = capi.Batch(ctx, w["experiments"], w["tables"], w["meta"])
    base, _, _ = b.run()
    best, tried = b.autotune(reps=1)
    print(f"[batched autotune] best {best}; {len(tried)} candidates, "
          f"{min(t for _, t in tried):.3f}..{max(t for _, t in tried):.3f} ms", flush=True)
    assert len(tried) >= 8 and b.get_schedule()["block_size"] == best["block_size"]
    assert {t["lean"] for t, _ in tried} == {0, 1, 2}  # every placement was timed
    after, _, ms = b.run()
    assert abs(ms - min(t for _, t in tried)) < 0.5 * max(ms, 1e-3) + 1.0
    for r0, r1 in zip(base, after):
        assert r0["num_valid"] == r1["num_valid"]
        assert abs(r0["log10_sum_w"] - r1["log10_sum_w"]) <= 1e-5 * abs(r0["log10_sum_w"])
    b.close()
    ctx.close()
