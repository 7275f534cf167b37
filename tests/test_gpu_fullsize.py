"""GPU checks at the full sizes of the benchmark workloads, through properties
that hold at any size (the per-path oracle comparisons run on subsamples in
the other GPU tests):

* C4 (N=100, arclength 1.3, K=100, sigma 0.08, 2^20 paths), adjoint:
  - table homogeneity: the lerped joint weights are linear in the table, so
    E(lambda T) = E(T) + (N-1) log10(lambda) and sum_b T_b dE/dT_b = (N-1)/ln 10;
  - scale and rotation invariance: every curvature is unchanged when all initial
    segments are scaled or rotated together (the perturbation axes are
    normalised chords), so sum_i d_i . dE/dd_i = 0 and sum_i d_i x dE/dd_i = 0.
  A wrong sign, index or transpose in the reverse sweep breaks them at O(1).
* C5 (N=1024, arclength 2, K=200, 3 summed tables, 2^24 paths, 256x256 exit
  histogram): the histogram sums to 1, two runs give the same estimate bit for
  bit and the same histogram to the rounding of its float atomics, and the
  estimate of two halves (path offsets) merges into the full one.
* C3 (4096 heterogeneous experiments, one graph-replayed launch): replays agree
  bit for bit, every experiment reports its path count, every experiment's
  table is the one built at its own ds, experiments spread over the range of N
  match plain runs of the same draws, and the experiments with the largest and
  smallest ds meet the per-path rule against the FP64 oracle.
"""
import math

import numpy as np
import pytest

from tests.test_gpu_batched import batch_lanes, check_experiment_against_oracle, plain_run
from twisty_amd import capi, host, workloads

pytestmark = pytest.mark.gpu


def context(N, L, sigmas, paths_hint=None):
    ds = float(np.float32(L) / np.float32(N))
    tabs = [host.build_rt_table(ds, 1.0, float(np.float32(s) ** 2), 0.1, 0.004, numStepsInt=20000,
                                numCurvatureSteps=4096) for s in sigmas]
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    ctx = capi.Context(0)
    t0 = tabs[0]
    table = t0.table if len(tabs) == 1 else np.stack([t.table for t in tabs])
    ctx.set_tables(table, t0.minK, t0.maxK, t0.step, t0.ds, 0.004)
    pos = host.initial_arc(bc, N)
    ctx.set_curve(pos, bc)
    return ctx, tabs, pos


def test_c4_full_size_adjoint_identities(hip_device):
    N, L, K, P = 100, 1.3, 100, 1 << 20
    ctx, tabs, pos = context(N, L, [0.08])
    r = ctx.run_adjoint(P, K, seed=1, segments=True, kappa=0.0)
    ctx.close()
    assert r["num_paths"] == P and r["num_valid"] + r["num_invalid"] == P
    T = tabs[0].table.astype(np.float64)
    hom = float(np.dot(T, r["d_table"]))
    want = (N - 1) / math.log(10.0)
    d = np.diff(pos.astype(np.float64), axis=0)
    g = r["d_seg"]
    scale = float(np.sum(np.abs(d) * np.abs(g)))
    dot = float(np.sum(d * g))
    torque = np.cross(d, g).sum(axis=0)
    print(f"[C4 2^20 adjoint] sum T dE/dT = {hom:.9f} (want {want:.9f}); sum d.g = {dot:.3e}, "
          f"|sum d x g| = {np.linalg.norm(torque):.3e}, scale {scale:.3e}", flush=True)
    assert abs(hom / want - 1.0) < 1e-4
    assert abs(dot) < 1e-4 * scale and np.linalg.norm(torque) < 1e-4 * scale


def test_c5_full_size_properties(hip_device):
    N, L, K, P = 1024, 2.0, 200, 1 << 24
    ctx, _, _ = context(N, L, [0.02, 0.05, 0.1])
    ctx.set_exit_histogram(256, 256)
    a = ctx.run(P, K, seed=1)
    ha, _ = ctx.exit_histogram()
    b = ctx.run(P, K, seed=1)
    hb, _ = ctx.exit_histogram()
    h1 = ctx.run(P // 2, K, seed=1)
    h2 = ctx.run(P // 2, K, seed=1, path_offset=P // 2)
    ctx.close()
    m = max(h1["log10_sum_w"], h2["log10_sum_w"])
    merged = m + math.log10(10 ** (h1["log10_sum_w"] - m) + 10 ** (h2["log10_sum_w"] - m))
    mean, se = capi.estimator(a)
    print(f"[C5 2^24] valid {a['num_valid']} invalid {a['num_invalid']}, log10 avg {mean:.6f} +- {se:.2e}, "
          f"histogram sum {ha.sum():.9f}, halves merged - full {merged - a['log10_sum_w']:.2e}", flush=True)
    assert a["num_paths"] == P and a["num_valid"] + a["num_invalid"] == P and a["num_valid"] > 0.99 * P
    assert abs(ha.sum() - 1.0) < 1e-6 and np.all(ha >= 0.0)
    # the estimate is bit-reproducible; the bins are summed with LDS float atomics,
    # so a second run agrees to their rounding
    assert a["log10_sum_w"] == b["log10_sum_w"]
    assert np.abs(ha - hb).max() <= 1e-6 * ha.max()
    assert abs(merged - a["log10_sum_w"]) <= 1e-6 * abs(a["log10_sum_w"])


def test_c3_full_size_properties(hip_device, oracle):
    w = workloads.batch_experiments(n_exp=4096, seed=1, K=100)
    exps = w["experiments"]
    ctx = capi.Context(0)
    b = capi.Batch(ctx, exps, w["tables"], w["meta"])
    spl = b.get_schedule()["segments_per_lane"]
    first, pp, ms = b.run(per_path=True)
    again, _, _ = b.run()
    b.close()
    assert first == again
    assert [r["num_paths"] for r in first] == [x["paths"] for x in exps]
    assert all(r["num_valid"] + r["num_invalid"] == r["num_paths"] for r in first)
    # every experiment is weighted with its own table, built at its own ds
    for x in exps:
        assert w["meta"][x["table"]][3] == np.float32(x["arclength"]) / np.float32(x["N"])
    order = np.argsort([x["N"] for x in exps])
    picks = [int(order[k]) for k in (0, len(order) // 3, 2 * len(order) // 3, len(order) - 1)]
    for e in picks:
        x, r = exps[e], first[e]
        ref = plain_run(ctx, x, w["tables"], w["meta"])
        print(f"[C3 4096] exp {e} N={x['N']} paths={x['paths']}: batched {r['log10_sum_w']:.9f} "
              f"plain {ref['log10_sum_w']:.9f}", flush=True)
        assert r["num_valid"] == ref["num_valid"]
        assert abs(r["log10_sum_w"] - ref["log10_sum_w"]) <= 1e-6 * abs(ref["log10_sum_w"]) + 1e-9
    # the experiments with the largest and the smallest ds, path by path against the oracle
    ds = np.array([np.float32(x["arclength"]) / np.float32(x["N"]) for x in exps])
    for e in (int(np.argmax(ds)), int(np.argmin(ds))):
        x = exps[e]
        check_experiment_against_oracle(ctx, oracle, x, w["tables"], w["meta"], pp[e], batch_lanes(x["N"], spl),
                                        4096, f"C3 4096 exp {e}")
    print(f"[C3 4096] replay {ms:.2f} ms", flush=True)
    ctx.close()
