#!/usr/bin/env python3
"""Benchmark of the optimal-perturb hot path on one MI355X.

One step = one twisty_hip_run over the whole workload: N=64 segments, arclength
1.5 between (0,0,0) and (0,0,1) with both end tangents +z, circular-arc initial
curve, K=100 perturbations per path, 2^22 paths, one 4096-bin radiative-transfer
table (Gaussian phase sigma=0.05 -> eps=sigma^2, mu 1.0, scatter 0.1,
absorption 0.004, numStepsInt 20000), Philox seed 1.  Inputs are built on the
host and uploaded before timing (table construction and H2D are excluded).

Prints one JSON line.  `value` is joint-weights/s = paths * (N-1) / wall
seconds per step; `roofline` prices the forward kernel's HIP-event time against
the FP32 vector peak with the algorithmic flop count per path
F = K (40 + 18 r) + 30 N - 14, r = E[diff] - 1; `cpu_baseline` times the
reference CPU path restated in oracle/ (FP32 positions, FP64 weights, same
draws) on the host cores over a bounded sample.
"""
import argparse
import json
import math
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "joint-weights/s at N=64, 2^22 paths; fwd+adjoint ms; batched vs plain"
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, AMD published
HBM_PEAK_GBS = 8000.0

WORKLOAD = {"N": 64, "L": 1.5, "K": 100, "paths": 1 << 22, "sigma": 0.05, "bins": 4096, "mu": 1.0,
            "scatter": 0.1, "absorption": 0.004, "numStepsInt": 20000, "seed": 1}


def progress(msg):
    """One line per finished leg, on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def mean_rotated_points(N):
    """E[diff] - 1 for diff ~ U{2..min(N-2,25)}."""
    md = min(N - 2, 25)
    return (2 + md) / 2.0 - 1.0


def flops_per_path(N, K):
    return K * (40 + 18 * mean_rotated_points(N)) + 30 * N - 14


def traffic_from_profiles():
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/), or None."""
    p = os.path.join(ROOT, "profiles", "forward_c2_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def valu_issue_from_profiles():
    """VALU issue utilisation of the forward kernel (SQ_INSTS_VALU x 4 cycles /
    (kernel cycles x SIMDs), rocprofv3 --pmc, profiles/), or None.  The kernel's
    bound is instruction issue, so this is how close it runs to it."""
    p = os.path.join(ROOT, "profiles", "forward_c2_valu_counters.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("valu_issue_utilisation")


def cpu_baseline(table, pos, w, target_seconds):
    """Reference CPU path (oracle/, FP32 positions + FP64 weights) on the host."""
    from oracle import pyoracle

    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    threads = os.cpu_count() or 1

    class T:
        pass

    t = T()
    t.table, t.ds, t.minK, t.maxK, t.step = table.table, table.ds, table.minK, table.maxK, table.step

    def run(n, th):
        r = pyoracle.simulate(pos.astype("float64"), w["N"], w["L"], w["K"], n, t, w["absorption"], seed=w["seed"],
                              precision=32, threads=th, want_w=False)
        return r["seconds"]

    # Sample: paths 0..n-1 of the same workload, n a power of two sized so that
    # three repeats take about target_seconds (at most 4x the workload's
    # paths); the median repeat is reported.  Threads are pinned one per core.
    n0 = 1 << 12
    s0 = run(n0, threads)
    n = n0
    while n < 4 * w["paths"] and s0 * (n / n0) * 2 * 3 <= target_seconds * 1.5:
        n *= 2
    secs = statistics.median(run(n, threads) for _ in range(3))
    jw = n * (w["N"] - 1) / secs
    out = {"value": jw, "unit": "joint-weights/s", "cores": threads, "kind": "port",
           "sample": f"{n} paths of the same workload (paths 0..{n - 1}), reference CPU path restated in oracle/ "
                     f"(FP32 positions, FP64 weights, same Philox draws, -O3), {threads} pinned std::threads, "
                     f"median of 3 runs {secs:.2f} s"}
    # The reference's own split: 1e6-path granules per thread, so at 2^22 paths
    # only ceil(2^22/1e6) = 5 threads work (FullExperimentRunnerOptimalPerturb.cpp:255-277).
    ref_threads = min(threads, math.ceil(w["paths"] / 1_000_000))
    n_ref = max(n0, n * ref_threads // threads)
    secs_ref = statistics.median(run(n_ref, ref_threads) for _ in range(3))
    ref_part = {"value": n_ref * (w["N"] - 1) / secs_ref, "unit": "joint-weights/s", "cores": ref_threads,
                "kind": "port", "sample": f"{n_ref} paths on {ref_threads} pinned threads (median of 3, "
                                          f"{secs_ref:.2f} s)"}
    return out, ref_part


def measure_adjoint(ctx, w, bc=None, reps=3, roulette=True):
    """Forward + adjoint device time at the same workload (HIP events): the
    table gradient alone; the full gradient (table + reverse sweep) with the
    default sweep (Russian roulette, kappa = 1) and with the exact sweep over
    every path.  With the boundary conditions bc, the exit-angle gradient
    (the sweep's dE/d(initial segments) along d(initial segments)/d(theta,
    phi)) is reported with the roulette's added standard deviation and the
    Monte Carlo standard error, for both sweeps."""
    import numpy as np

    from twisty_amd import host

    if bc is not None:
        jt, jp = host.initial_arc_exit_jacobian(bc, w["N"])
        ctx.set_adjoint_directions(np.stack([np.diff(jt, axis=0), np.diff(jp, axis=0)]))
    out = {}
    exact = None
    legs = [(False, "table_only", 0.0), (True, "full_exact", 0.0)] + ([(True, "full", 1.0)] if roulette else [])
    for seg, key, kappa in legs:
        ctx.set_adjoint_roulette(kappa)
        r = ctx.run_adjoint(w["paths"], w["K"], seed=w["seed"], segments=seg)
        tot, adj = [], []
        for _ in range(reps):
            ctx.run_adjoint(w["paths"], w["K"], seed=w["seed"], segments=seg)
            tot.append(ctx.last_timing()[1])
            adj.append(ctx.last_adjoint_ms())
        out[key] = {"fwd_adjoint_ms": statistics.fmean(tot), "adjoint_ms": statistics.fmean(adj),
                    "paths_swept": ctx.last_adjoint_paths() if seg else 0, "roulette_kappa": kappa if seg else None}
        if seg and bc is not None:
            st = ctx.last_adjoint_direction_stats()
            out[key]["exit_angle_gradient"] = {k: [float(v) for v in st[k]] for k in ("gradient", "roulette_std",
                                                                                        "mc_std")}
        if key == "full_exact":
            exact = r
        if key == "full":
            out[key]["d_seg_rel_to_exact"] = float(np.linalg.norm(r["d_seg"] - exact["d_seg"]) /
                                                   np.linalg.norm(exact["d_seg"]))
            out[key]["d_table_rel_to_exact"] = float(np.linalg.norm(r["d_table"] - exact["d_table"]) /
                                                     np.linalg.norm(exact["d_table"]))
    ctx.set_adjoint_roulette(1.0)
    ctx.set_adjoint_directions(None)
    return out


def measure_c4(capi, host, reps=5):
    """Differentiation workload: N=100, arclength 1.3, 2^20 paths, K=100,
    sigma=0.08; the plain forward and forward + adjoint (table only, exact
    full gradient, roulette) device times, HIP events."""
    import numpy as np

    w = dict(WORKLOAD, N=100, L=1.3, K=100, paths=1 << 20, sigma=0.08)
    ds = float(np.float32(w["L"]) / np.float32(w["N"]))
    table = host.build_rt_table(ds, w["mu"], float(np.float32(w["sigma"]) ** 2), w["scatter"], w["absorption"],
                                numStepsInt=w["numStepsInt"], numCurvatureSteps=w["bins"])
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), w["L"])
    ctx = capi.Context(0)
    ctx.set_tables(table.table, table.minK, table.maxK, table.step, table.ds, w["absorption"])
    ctx.set_curve(host.initial_arc(bc, w["N"]), bc)
    ctx.run(w["paths"], w["K"], seed=w["seed"])
    fwd = []
    for _ in range(reps):
        ctx.run(w["paths"], w["K"], seed=w["seed"])
        fwd.append(ctx.last_timing()[0])
    fwd_ms = statistics.fmean(fwd)
    adj = measure_adjoint(ctx, w, bc)
    ctx.close()
    return {"workload": "C4: N=100, arclength 1.3, K=100, 2^20 paths, 4096-bin RT table sigma=0.08",
            "forward_ms": fwd_ms, **adj,
            "ratio_full": adj["full"]["fwd_adjoint_ms"] / fwd_ms,
            "ratio_full_exact": adj["full_exact"]["fwd_adjoint_ms"] / fwd_ms,
            "ratio_table_only": adj["table_only"]["fwd_adjoint_ms"] / fwd_ms}


def measure_generated(capi, host, cpu_seconds=None, N=66, L=1.65, paths=1 << 22, reps=5):
    """Independent-path estimator (paths built by recursive sphere intersection,
    ExperimentBase.cpp:1246-1386) at N=66 (the count nearest 64 its recursion
    allows), 2^22 paths, sigma=0.05; with the oracle timed on a CPU sample."""
    import numpy as np

    ds = float(np.float32(L) / np.float32(N))
    table = host.build_rt_table(ds, 1.0, float(np.float32(0.05) ** 2), 0.1, 0.004, numStepsInt=20000,
                                numCurvatureSteps=4096)
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    ctx = capi.Context(0)
    ctx.set_tables(table.table, table.minK, table.maxK, table.step, table.ds, 0.004)
    ctx.set_curve(host.initial_arc(bc, N), bc)
    ctx.run_generated(paths, seed=1)
    ms = []
    for _ in range(reps):
        r = ctx.run_generated(paths, seed=1)
        ms.append(ctx.last_timing()[0])
    # the two-table Metropolis-Hastings form (ExperimentBase.cpp:1599-1818) on
    # the same boundary: 2^16 chains x 64 proposals, ds from the minimum
    # arclength up by half the nominal ds, 8 cached tables over that range
    lo = np.float32(host.minimum_arclength(bc, N) / N)
    hi = np.float32(0.5 * ds)
    with tempfile.TemporaryDirectory() as cache_root:  # the tables' .cwt cache, not kept
        tabs, _, _ = host.cached_rt_tables(float(lo), float(lo + hi), 8, 1.0, float(np.float32(0.05) ** 2), 0.1,
                                           0.004, numStepsInt=20000, numCurvatureSteps=4096, cache_root=cache_root)
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    ctx.set_multi_tables(np.stack([t.table for t in tabs]),
                         np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32), ext, N)
    chains, length = 1 << 16, 64
    ctx.run_generated_chains(chains, length, float(lo), float(hi), ext, seed=1)
    cms = []
    for _ in range(reps):
        rc = ctx.run_generated_chains(chains, length, float(lo), float(hi), ext, seed=1)
        cms.append(ctx.last_timing()[0])
    ctx.close()
    ccms = statistics.fmean(cms)
    kms = statistics.fmean(ms)
    out = {"workload": f"independent paths: N={N}, arclength {L}, 2^22 paths, 4096-bin RT table sigma=0.05",
           "kernel": "generate_kernel", "kernel_ms": kms, "joint_weights_per_s": paths * (N - 1) / (kms / 1e3),
           "num_valid": r["num_valid"], "log10_sum_w": r["log10_sum_w"], "cpu_baseline": None,
           "chains": {"workload": f"two-table Metropolis-Hastings form: {chains} chains x {length} proposals, ds in "
                                  f"[{float(lo):.5f}, {float(lo + hi):.5f}], 8 cached tables, absorption + scatter "
                                  f"{ext:.4f}", "kernel_ms": ccms,
                      "proposals_per_s": chains * length / (ccms / 1e3),
                      "joint_weights_per_s": chains * length * (N - 1) / (ccms / 1e3),
                      "num_valid": rc["num_valid"], "log10_sum_w": rc["log10_sum_w"]}}
    if cpu_seconds:
        from oracle import pyoracle

        if not os.path.exists(pyoracle.LIB_PATH):
            pyoracle.build()
        threads = os.cpu_count() or 1
        n = 1 << 14
        secs = pyoracle.generate(bc[:12].astype(np.float64), L, N, n, table, 0.004, seed=1, threads=threads,
                                 want_w=False)["seconds"]
        while secs * 2 <= cpu_seconds and n < paths:
            n *= 2
            secs = pyoracle.generate(bc[:12].astype(np.float64), L, N, n, table, 0.004, seed=1, threads=threads,
                                     want_w=False)["seconds"]
        out["cpu_baseline"] = {"value": n * (N - 1) / secs, "unit": "joint-weights/s", "cores": threads,
                               "kind": "port", "sample": f"{n} paths of the same workload, oracle/ FP64 restatement "
                                                         f"on {threads} std::threads, {secs:.1f} s"}
        tabs_np = np.stack([t.table for t in tabs])
        meta_np = np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32)
        nc = 1 << 10
        run_c = lambda k: pyoracle.generate_chains(bc[:12].astype(np.float64), N, k, length, tabs_np, meta_np, ext,
                                                   float(lo), float(hi), seed=1, threads=threads)["seconds"]
        secs = run_c(nc)
        while secs * 2 <= cpu_seconds and nc < chains:
            nc *= 2
            secs = run_c(nc)
        out["chains"]["cpu_baseline"] = {"value": nc * length / secs, "unit": "proposals/s", "cores": threads,
                                         "kind": "port", "sample": f"{nc} chains x {length} proposals of the same "
                                                                   f"workload, oracle/ FP64 restatement on {threads} "
                                                                   f"std::threads, {secs:.1f} s"}
    return out


def measure_c5(capi, host, reps=2):
    """Large beam spread: N=1024, arclength 2.0 between endpoints 1.0 apart,
    K=200, 2^24 paths, the sum of 3 tables (sigma 0.02/0.05/0.1), 256x256
    exit-direction histogram."""
    import numpy as np

    N, L, K, P = 1024, 2.0, 200, 1 << 24
    ds = float(np.float32(L) / np.float32(N))
    tabs = [host.build_rt_table(ds, 1.0, float(np.float32(s) ** 2), 0.1, 0.004, numStepsInt=20000,
                                numCurvatureSteps=4096) for s in (0.02, 0.05, 0.1)]
    if len({(t.minK, t.maxK, t.step) for t in tabs}) != 1:
        raise RuntimeError("C5 tables do not share one curvature grid")
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    ctx = capi.Context(0)
    t0 = tabs[0]
    ctx.set_tables(np.stack([t.table for t in tabs]), t0.minK, t0.maxK, t0.step, t0.ds, 0.004)
    ctx.set_curve(host.initial_arc(bc, N), bc)
    ctx.set_exit_histogram(256, 256)
    ctx.run(P, K, seed=1)
    fwd, tot = [], []
    for _ in range(reps):
        r = ctx.run(P, K, seed=1)
        f, t = ctx.last_timing()
        fwd.append(f)
        tot.append(t)
    h, _ = ctx.exit_histogram()
    sched = ctx.get_schedule()
    ctx.close()
    return {"workload": "C5: N=1024, arclength 2.0, endpoints 1.0 apart, K=200, 2^24 paths, 3 summed tables "
                        "sigma 0.02/0.05/0.1, 256x256 exit histogram",
            "forward_ms": statistics.fmean(fwd), "forward_plus_histogram_ms": statistics.fmean(tot),
            "joint_weights_per_s": P * (N - 1) / (statistics.fmean(tot) / 1e3), "num_valid": r["num_valid"],
            "histogram_sum": float(h.sum()), "schedule": sched}


def measure_batched(ctx, capi, n_exp=4096, autotune=True, schedule=None):
    """The batched workload in one graph-replayed launch vs one plain run per experiment."""
    import numpy as np

    from twisty_amd import workloads

    t0 = time.perf_counter()
    wl = workloads.batch_experiments(n_exp=n_exp, seed=1, K=100)
    setup_s = time.perf_counter() - t0
    exps = wl["experiments"]
    jw = sum(x["paths"] * (x["N"] - 1) for x in exps)
    b = capi.Batch(ctx, exps, wl["tables"], wl["meta"])
    tuned, tried = None, []
    if schedule:
        b.set_schedule(*schedule)
    elif autotune:
        tuned, tried = b.autotune(reps=1)
    b.run()
    ms = min(b.run()[2] for _ in range(3))
    sched = b.get_schedule()
    b.close()
    t0 = time.perf_counter()
    plain_dev_ms = 0.0
    for x in exps:
        m = wl["meta"][x["table"]]
        ctx.set_tables(wl["tables"][x["table"]], float(m[0]), float(m[1]), float(m[2]),
                       float(np.float32(x["arclength"]) / np.float32(x["N"])), float(m[4]))
        ctx.set_curve(x["positions"], x["boundary"])
        ctx.run(x["paths"], x["K"], seed=x["seed"])
        plain_dev_ms += ctx.last_timing()[1]
    plain_s = time.perf_counter() - t0
    return {"experiments": len(exps), "paths": sum(x["paths"] for x in exps), "joint_weights": jw,
            "batched_ms": ms, "batched_joint_weights_per_s": jw / (ms / 1e3), "schedule": sched,
            "autotune_candidates": len(tried),
            "plain_loop_ms": plain_s * 1e3, "plain_joint_weights_per_s": jw / plain_s,
            "plain_device_ms": plain_dev_ms,
            "batched_vs_plain": plain_s * 1e3 / ms, "batched_vs_plain_device": plain_dev_ms / ms,
            "tables": len(wl["tables"]), "setup_s": setup_s,
            "workload": "4096 experiments: N log-uniform [16,256], L/D ~ U[1.05,3], end direction uniform on "
                        "the sphere, paths log-uniform [2^10,2^16], K=100, weighting sigma ~ U[0.03,0.2] (8 values), "
                        "each experiment's table built at its own ds"}


def measure_batched_c2(ctx, capi, w, table, pos, bc):
    """The C2 workload (one experiment, 2^22 paths) through the batched kernel."""
    ds = float(table.ds)
    exp = [{"N": w["N"], "paths": w["paths"], "K": w["K"], "seed": w["seed"], "table": 0, "arclength": w["L"],
            "positions": pos}]
    meta = [[table.minK, table.maxK, table.step, ds, w["absorption"]]]
    b = capi.Batch(ctx, exp, [table.table], meta)
    best, _ = b.autotune(reps=1)
    b.run()
    ms = min(b.run()[2] for _ in range(3))
    res = b.run()[0][0]
    b.close()
    return {"ms": ms, "joint_weights_per_s": w["paths"] * (w["N"] - 1) / (ms / 1e3), "schedule": best,
            "log10_sum_w": res["log10_sum_w"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=WORKLOAD["paths"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-generated", action="store_true")
    ap.add_argument("--no-roulette", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched leg: run the library's default schedule instead of the autotuner's")
    ap.add_argument("--batched-schedule", default=None,
                    help="batched leg: run this schedule (block,blocks_per_cu,segments_per_lane,tile,lean) "
                         "instead of the autotuner's")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the hot path on a single MI355X "
              f"(--gpus {a.gpus} requested); rerun with --gpus 1", file=sys.stderr)
        return 2
    import numpy as np

    from twisty_amd import capi, host

    w = dict(WORKLOAD)
    w["paths"] = a.paths
    N, L, K = w["N"], w["L"], w["K"]
    ds = float(np.float32(L) / np.float32(N))  # Curve.cpp:24 computes ds in float
    eps = float(np.float32(w["sigma"]) ** 2)
    table = host.build_rt_table(ds, w["mu"], eps, w["scatter"], w["absorption"], numStepsInt=w["numStepsInt"],
                                numCurvatureSteps=w["bins"])
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    pos = host.initial_arc(bc, N)
    ctx = capi.Context(0)
    ctx.set_tables(table.table, table.minK, table.maxK, table.step, table.ds, w["absorption"])
    ctx.set_curve(pos, bc)
    sched = ctx.get_schedule()
    for _ in range(a.warmup):
        ctx.run(w["paths"], K, seed=w["seed"])
    fwd = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = ctx.run(w["paths"], K, seed=w["seed"])
        fwd.append(ctx.last_timing()[0])
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    value = w["paths"] * (N - 1) / (ms_per_step / 1e3)
    fwd_ms = statistics.fmean(fwd)
    progress(f"forward: {a.steps} steps, {fwd_ms:.3f} ms per launch")
    F = flops_per_path(N, K)
    achieved = F * w["paths"] / (fwd_ms / 1e3) / 1e12
    mean_l, rel_se = capi.estimator(res)
    out = {
        "metric": METRIC, "value": value, "unit": "joint-weights/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "C2 plain forward: N=64 segments, arclength 1.5, K=100 perturbations/path, "
                               "2^22 paths, 4096-bin RT table sigma=0.05", "N": N, "arclength": L,
                   "perturbations_per_path": K, "paths": w["paths"], "table_bins": w["bins"], "sigma": w["sigma"],
                   "mu": w["mu"], "scatter": w["scatter"], "absorption": w["absorption"], "seed": w["seed"],
                   "schedule": sched},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic_from_profiles(),
                     "kernel": "forward_kernel", "kernel_ms": fwd_ms, "flops_per_path": F,
                     "valu_issue_utilisation": valu_issue_from_profiles()},
        "estimate": {"log10_avg_path_weight": mean_l, "rel_standard_error": rel_se,
                     "num_valid": res["num_valid"], "num_invalid": res["num_invalid"]},
        "fwd_adjoint_ms": None,
        "batched": None,
    }
    if not a.no_adjoint:
        adj = measure_adjoint(ctx, w, bc, roulette=not a.no_roulette)
        dflt = adj["full"] if "full" in adj else adj["full_exact"]
        progress(f"adjoint: full {dflt['fwd_adjoint_ms']:.3f} ms (exact sweep "
                 f"{adj['full_exact']['fwd_adjoint_ms']:.3f} ms)")
        out["fwd_adjoint_ms"] = dflt["fwd_adjoint_ms"]
        out["adjoint"] = {"forward_ms": fwd_ms, **adj,
                          "ratio_full": dflt["fwd_adjoint_ms"] / fwd_ms,
                          "ratio_full_exact": adj["full_exact"]["fwd_adjoint_ms"] / fwd_ms,
                          "ratio_table_only": adj["table_only"]["fwd_adjoint_ms"] / fwd_ms,
                          "gradients": "table_only: dE/dT (4096 entries, -> sigma), every path; full: dE/dT over "
                                       "every path + dE/d(initial segments) (-> exit angles) from the default "
                                       "reverse sweep, paths kept with probability min(1, omega n / kappa), "
                                       "kappa = 1, and reweighted (unbiased; its added standard deviation is "
                                       "reported next to the Monte Carlo error); full_exact: the same with the "
                                       "sweep over every path"}
    if not a.no_c4:
        out["c4"] = measure_c4(capi, host)
        progress(f"C4: forward {out['c4']['forward_ms']:.3f} ms, full gradient "
                 f"{out['c4']['full']['fwd_adjoint_ms']:.3f} ms (roulette sweep), "
                 f"{out['c4']['full_exact']['fwd_adjoint_ms']:.3f} ms (exact sweep)")
    if not a.no_batched:
        sched = tuple(int(v) for v in a.batched_schedule.split(",")) if a.batched_schedule else None
        out["batched"] = measure_batched(ctx, capi, autotune=not a.no_autotune, schedule=sched)
        progress(f"batched: {out['batched']['batched_ms']:.3f} ms")
        out["batched"]["c2_one_experiment"] = measure_batched_c2(ctx, capi, w, table, pos, bc)
    if not a.no_c5:
        out["c5"] = measure_c5(capi, host)
        progress(f"C5: {out['c5']['forward_plus_histogram_ms']:.2f} ms")
    if not a.no_generated:
        out["generated"] = measure_generated(capi, host, None if a.no_cpu_baseline else a.cpu_seconds / 4)
        progress(f"independent paths: {out['generated']['kernel_ms']:.3f} ms; chains "
                 f"{out['generated']['chains']['kernel_ms']:.3f} ms")
    if not a.no_cpu_baseline:
        cb, refpart = cpu_baseline(table, pos, w, a.cpu_seconds)
        progress(f"cpu baseline: {cb['value']:.3e} joint-weights/s on {cb['cores']} threads")
        out["cpu_baseline"] = cb
        out["cpu_baseline_reference_partitioning"] = refpart
        out["speedup_vs_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
