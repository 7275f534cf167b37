"""GPU tests of the C++ host layer: JSON config -> FullExperimentRunnerOptimalPerturb_HIP.

The runner builds its own table and initial curve; the oracle is fed the same
two inputs (the product's table is bit-identical to the oracle's, see
test_host_cpu) so the estimates compare path for path: the sum of weights may
differ by at most the worst per-path linear error the forward tests allow
(1e-5 relative on log10 weights of about -40 is a linear error below 1e-3).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from tests.lerp import joint_log10w_spread
from twisty_amd import capi, host

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def config(N=20, L=1.2, paths=1 << 16, K=50, sigma=0.1, chain=1, seed=1, bins=4096, write=False, outdir="/tmp",
           skip=0):
    return {"experiment": {
        "experimentParams": {
            "pathsToGenerate": paths, "pathsToSkip": skip, "numSegments": N, "name": "hostgpu",
            "experimentDir": outdir, "maxPerturbThreads": 0, "maxWeightThreads": 0,
            "outputBigFloatWeights": False, "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
            "random": {"bootstrapSeed": 1, "perturbSeed": seed},
            "weighting": {"weightFunction": 0, "mu": 1.0, "eps": float(np.float32(sigma) ** 2),
                          "numStepsInt": 20000, "numCurvatureSteps": bins, "absorption": 0.004, "scatter": 0.1},
            "twisty_hip": {"perturbationsPerPath": K, "writeFiles": write,
                           **({} if chain is None else {"chainLength": chain})}},
        "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1],
                            "endDir": [0, 0, 1], "arclength": L}}}


def oracle_same_inputs(oracle, N, L, K, paths, sigma, chain=1, seed=1, bins=4096, skip=0):
    ds = float(np.float32(L) / np.float32(N))
    eps = float(np.float32(sigma) ** 2)
    t = oracle.build_table(ds, mu=1.0, eps=eps, scatter=0.1, absorption=0.004, numStepsInt=20000,
                           numCurvatureSteps=bins)
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    pos = host.initial_arc(bc, N).astype(np.float64)
    return oracle.simulate(pos, N, L, K, paths, t, 0.004, seed=seed, chain_length=chain, burn_in=skip)


@pytest.mark.parametrize("N,L,K,sigma,chain,skip", [(20, 1.2, 50, 0.1, 1, 0), (64, 1.5, 100, 0.05, 1, 0),
                                                    (64, 1.5, 1, 0.05, 256, 0), (64, 1.5, 1, 0.05, 256, 100)])
def test_runner_matches_oracle(hip_device, oracle, N, L, K, sigma, chain, skip):
    paths = 1 << 15
    r = host.run_config(config(N=N, L=L, paths=paths, K=K, sigma=sigma, chain=chain, skip=skip))
    ref = oracle_same_inputs(oracle, N, L, K, paths, sigma, chain=chain, skip=skip)
    ref_avg = ref["log10_sum_w_exact"] - math.log10(paths)
    print(f"[runner N={N} K={K} L={chain}] gpu log10 avg={r['log10_avg_path_weight']:.9f} "
          f"oracle={ref_avg:.9f} kernel_ms={r['kernel_ms']:.3f} setup_ms={r['setup_ms']}", flush=True)
    assert r["total_paths"] == paths and r["failed_paths"] == 0
    assert abs(10 ** (r["log10_avg_path_weight"] - ref_avg) - 1.0) < 1e-3


def test_runner_default_chain_is_the_reference_chain(hip_device, oracle):
    """No chainLength in the config: one chain of 1e6 paths as the reference's
    CPU runner, with pathsToSkip burn-in, against the oracle on the same draws.
    Over 32768 chained perturbations FP32 and FP64 states drift apart (a
    near-folded sub-chain has a short chord, and its axis amplifies rounding),
    so the two chains agree statistically: within 3 standard errors of the
    oracle chain's batch means."""
    N, L, K, sigma, paths = 20, 1.2, 1, 0.1, 1 << 15
    r = host.run_config(config(N=N, L=L, paths=paths, K=K, sigma=sigma, chain=None, skip=100))
    ref = oracle_same_inputs(oracle, N, L, K, paths, sigma, chain=1_000_000, skip=100)
    ref_avg = ref["log10_sum_w_exact"] - math.log10(paths)
    w = ref["log10w"]
    m = np.nanmax(w)
    batch = np.nanmean(10.0 ** (w - m).reshape(32, -1), axis=1)
    rel_se = batch.std(ddof=1) / math.sqrt(len(batch)) / batch.mean()
    diff = abs(10 ** (r["log10_avg_path_weight"] - ref_avg) - 1.0)
    print(f"[runner default chain] gpu {r['log10_avg_path_weight']:.9f} oracle {ref_avg:.9f} rel diff {diff:.2e} "
          f"oracle rel SE {rel_se:.2e}", flush=True)
    assert r["total_paths"] == paths and diff < 3 * rel_se + 1e-3


def test_full_experiment_driver(hip_device, tmp_path):
    cfg = config(paths=1 << 14, write=True, outdir=str(tmp_path))
    p = tmp_path / "cfg.json"
    p.write_text(json.dumps(cfg))
    exe = os.path.join(ROOT, "twisty_amd", "lib", "twisty_full_experiment")
    out = subprocess.run([exe, str(p)], capture_output=True, text=True, timeout=120)
    print(out.stdout[-800:], flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Avg path weight" in out.stdout
    results = list(tmp_path.rglob("Results.dat"))
    tables = list(tmp_path.rglob("*_RT.csv"))
    assert len(results) == 1 and len(tables) == 1
    assert "Avg path weight" in results[0].read_text()
    assert len(tables[0].read_text().splitlines()) == 4096
    hist = list(tmp_path.rglob("histogram.txt"))
    assert len(hist) == 1
    lines = hist[0].read_text().splitlines()
    assert lines[0] == "Histogram" and len(lines) == 501
    counts = [float(l.split()[2]) for l in lines[1:]]
    assert sum(counts) == cfg["experiment"]["experimentParams"]["pathsToGenerate"]


def _histogram_counts(root):
    hist = list(root.rglob("histogram.txt"))
    assert len(hist) == 1
    return [float(l.split()[2]) for l in hist[0].read_text().splitlines()[1:]]


def test_runner_path_batch_export(hip_device, tmp_path):
    """outputPathBatches: the chunked run (two chunks of whole chains) gives the
    estimate of the plain run; the exported paths keep their endpoints, segment
    lengths and boundary directions, and their curvatures re-weighted on the
    host give the exported weights; sum of exported weights = experiment weight."""
    N, L, K, sigma, paths = 5, 1.2, 20, 0.1, 300000
    plain_cfg = config(N=N, L=L, paths=paths, K=K, sigma=sigma, write=True, outdir=str(tmp_path / "plain"))
    exp_cfg = config(N=N, L=L, paths=paths, K=K, sigma=sigma, write=True, outdir=str(tmp_path / "export"))
    exp_cfg["experiment"]["experimentParams"]["outputPathBatches"] = True
    a = host.run_config(plain_cfg)
    b = host.run_config(exp_cfg)
    print(f"[export] plain {a['log10_experiment_weight']:.12f} chunked {b['log10_experiment_weight']:.12f}", flush=True)
    assert b["total_paths"] == paths and b["failed_paths"] == a["failed_paths"]
    assert abs(a["log10_experiment_weight"] - b["log10_experiment_weight"]) <= 1e-9
    assert _histogram_counts(tmp_path / "plain") == _histogram_counts(tmp_path / "export")

    gen = list((tmp_path / "export").rglob("GeneratedCurves"))
    assert len(gen) == 1
    r = host.read_path_batches(gen[0].parent)
    pos = r["positions"].astype(np.float64)
    n = paths - b["failed_paths"]
    assert pos.shape == (n, N + 1, 3) and r["batches"] == [(0, 0, n)] and r["N"] == N
    assert r["index"]["experiment_name"] == "hostgpu"
    ds = np.float32(L) / np.float32(N)
    seg = np.diff(pos, axis=1)
    assert np.array_equal(r["positions"][:, 0], np.zeros((n, 3), np.float32))
    assert np.abs(pos[:, -1] - [0, 0, 1]).max() < 1e-4
    assert np.abs(np.linalg.norm(seg, axis=2) / ds - 1).max() < 1e-4
    tan = seg / np.linalg.norm(seg, axis=2, keepdims=True)
    assert np.abs(tan[:, 0] - [0, 0, 1]).max() < 1e-4 and np.abs(tan[:, -1] - [0, 0, 1]).max() < 1e-4
    # weights from the exported curves (FullExperimentRunnerOptimalPerturb.cpp:711-722)
    table = host.build_rt_table(float(ds), 1.0, float(np.float32(sigma) ** 2), 0.1, 0.004, numStepsInt=20000,
                                numCurvatureSteps=4096)
    curv = np.linalg.norm(np.diff(tan, axis=1), axis=2) * (N / np.float32(L))
    # The exported points are the kernel's FP32 segments summed in FP64 and
    # rounded once to float, so each segment taken back from them is within
    # one ulp of the largest coordinate per component, each unit tangent within
    # twice that over ds, and the curvature within (N/L) times the two
    # tangents' error; the kernel's own FP32 curvature adds a few ulps.  delta
    # is three times that bound.  Every path's exported log10 weight must
    # then lie within 1e-5 relative of the re-weighted one plus the spread of
    # its joints' log10 weights over curvature +- delta (which holds the lerp
    # jump where a bin edge lies that close).
    ulp = float(np.spacing(np.float32(np.abs(pos).max())))
    dseg = math.sqrt(3.0) * ulp
    delta = 3.0 * ((N / L) * 2.0 * (2.0 * dseg / float(ds)) + 1e-6 * (1.0 + curv))
    lw, spread = joint_log10w_spread(curv, table, 0.004, delta)
    norm = host.norm_log10(N, float(ds), host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L))
    got = np.log10(r["weights"]) - norm
    allow = spread.sum(axis=1)
    excess = np.abs(lw.sum(axis=1) - got) - (1e-5 * np.abs(got) + allow)
    loose = allow > 1e-5 * np.abs(got)
    print(f"[export] re-weighted curves: delta {float(np.max(delta)):.2e}, worst excess {excess.max():.2e}, "
          f"largest allowance {allow.max():.2e}, {int(loose.sum())} of {n} paths above 1e-5 relative", flush=True)
    assert excess.max() <= 0.0
    # The allowance comes from the lerp jumps of joints within delta of a bin
    # edge (about 2 delta / step of them) and stays far below one joint's
    # log10 weight, so a joint dropped, doubled or weighed with another table
    # cannot hide in it.
    assert allow.max() < 1e-2 * np.abs(lw).min()
    total = math.log10(np.sum(r["weights"]))
    assert abs(total - b["log10_experiment_weight"]) < 1e-5
    assert np.abs(r["five_segment"][:, 0] - np.arccos(tan[:, 1, 2])).max() < 1e-3


def test_run_experiments_match_single_runs(hip_device):
    """BatchedExperimentRunner: each experiment of one batched launch gives the
    estimate of run_config on that experiment alone (same seed, chain length,
    burn-in; lanes per path agree at N=20, so the paths are bit-identical and
    only the merge order of the sums differs)."""
    N, K, chain, skip = 20, 2, 64, 10
    geoms = [((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 1.2), ((0, 0, 0), (0, 0, 1), (0.3, 0, 1), (0.6, 0, 0.8), 1.5),
             ((0, 0, 0), (1, 0, 0), (1, 0.2, 0), (1, 0, 0), 1.2)]
    cfg = config(N=N, paths=1 << 14, K=K, chain=chain, skip=skip)
    bcs = np.stack([host.boundary13(*g) for g in geoms])
    seeds = [3, 4, 5]
    res, ms = host.run_experiments(cfg, bcs, seeds)
    print(f"[run_experiments] {len(res)} experiments, device {ms:.3f} ms", flush=True)
    for g, s, r in zip(geoms, seeds, res):
        one = config(N=N, paths=1 << 14, K=K, chain=chain, skip=skip, seed=s)
        one["experiment"]["basicExperiment"].update(startPos=list(g[0]), startDir=list(g[1]), endPos=list(g[2]),
                                                   endDir=list(g[3]), arclength=g[4])
        ref = host.run_config(one)
        assert r["total_paths"] == ref["total_paths"] and r["failed_paths"] == ref["failed_paths"]
        assert abs(r["log10_experiment_weight"] - ref["log10_experiment_weight"]) <= 1e-9 * max(
            1.0, abs(ref["log10_experiment_weight"]))


def _beam_cfg(outdir):
    cfg = config(N=40, paths=1 << 13, K=20, chain=1)
    cfg["experiment"]["experimentParams"]["random"] = {"bootstrapSeed": 5, "perturbSeed": 7}
    cfg["experiment"]["beamSpreadExperiment"] = {
        "targetSeed": 1, "sphereRadius": 3.0, "maxArclength": 9.0, "startPos": [0.0, 0.0, 0.0],
        "startDir": [1.0, 0.0, 0.0], "numInitialCurves": 2, "numPerInitialCurve": 3, "numArclengths": 2}
    return cfg


def test_beam_spread_driver(hip_device, tmp_path):
    """The batched BeamSpread driver: beamSpreadValues.dat holds one line per
    experiment and the value = mean over arclengths and initial curves of the
    max over per-curve seeds; each experiment equals its run alone."""
    cfg = _beam_cfg(tmp_path)
    p = tmp_path / "beam.json"
    p.write_text(json.dumps(cfg))
    exe = os.path.join(ROOT, "twisty_amd", "lib", "twisty_beam_spread")
    out = subprocess.run([exe, str(p)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(out.stdout[-600:], flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    text = (tmp_path / "hostgpu" / "beamSpreadValues.dat").read_text().splitlines()
    value = float(text[1].split()[1])
    rows = [l.split() for l in text[3:]]
    assert len(rows) == 12
    best = {}
    for r in rows:
        k = (int(r[0]), int(r[1]))
        best[k] = max(best.get(k, -np.inf), float(r[5]))
    want = np.log10(np.sum(10.0 ** (np.array(list(best.values())) - max(best.values())))) + max(best.values()) \
        - math.log10(len(best))
    assert abs(value - want) < 1e-9
    # one experiment alone (its 64-bit seed does not fit the config's 32-bit
    # perturbSeed, so through run_experiments, which matches run_config above)
    r = rows[4]
    plan = subprocess.run([exe, str(p), "--plan"], capture_output=True, text=True, timeout=60, cwd=tmp_path).stdout
    end_dir = [float(t) for t in next(l for l in plan.splitlines() if l.startswith("endDir:")).split()[1:]]
    end_pos = [float(t) for t in next(l for l in plan.splitlines() if l.startswith("endPos:")).split()[1:]]
    bc = host.boundary13((0, 0, 0), (1, 0, 0), end_pos, end_dir, float(r[3]))
    ref, _ = host.run_experiments(config(N=40, paths=1 << 13, K=20, chain=1), bc[None], [int(r[4])])
    assert abs(ref[0]["log10_experiment_weight"] - float(r[5])) < 1e-9 * abs(float(r[5]))


def test_ring_benchmark_driver(hip_device, tmp_path):
    """The batched RingBenchmark driver: one row per height in ringBenchmark.dat;
    the reference's single receiver (the last height) equals its run alone."""
    cfg = config(N=24, paths=1 << 13, K=30, chain=1, write=True, outdir=str(tmp_path))
    cfg["experiment"]["ringBenchmark"] = {"startPos": [0.0, 0.0, 0.0], "startDir": [1.0, 0.0, 0.0], "zMin": 3.0,
                                          "zMax": 10.0, "numZValues": 4, "ringRadius": 5}
    p = tmp_path / "ring.json"
    p.write_text(json.dumps(cfg))
    exe = os.path.join(ROOT, "twisty_amd", "lib", "twisty_ring_benchmark")
    out = subprocess.run([exe, str(p)], capture_output=True, text=True, timeout=120)
    print(out.stdout[-600:], flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    dat = list(tmp_path.rglob("ringBenchmark.dat"))
    assert len(dat) == 1
    rows = [l.split() for l in dat[0].read_text().splitlines()[1:]]
    assert len(rows) == 4
    assert np.allclose([float(r[0]) for r in rows], [3.0, 3.0 + 7 / 3, 3.0 + 14 / 3, 10.0], rtol=1e-6)
    z, L = float(rows[-1][0]), float(rows[-1][1])
    d = np.array([5.0, 0.0, z], np.float32)
    one = config(N=24, paths=1 << 13, K=30, chain=1)
    one["experiment"]["basicExperiment"].update(startPos=[0, 0, 0], startDir=[1, 0, 0], endPos=d.tolist(),
                                               endDir=(d / np.linalg.norm(d)).tolist(), arclength=L)
    ref = host.run_config(one)
    assert abs(ref["log10_experiment_weight"] - float(rows[-1][2])) < 1e-6
    assert "Avg path weight" in out.stdout


def _convergence_lines(path):
    out = []
    for line in path.read_text().splitlines():
        n, w = line.split(",")
        out.append((int(n), float(w)))
    return out


def test_runner_convergence_weights(hip_device, tmp_path):
    """ConvergenceWeights/: 2.5e6 paths = three accumulators of 1e6 consecutive
    paths, dealt to maxPerturbThreads=2 threads in blocks of ceil(3/2) = 2
    (FullExperimentRunnerOptimalPerturb.cpp:221-266, 342-367).  Thread 0's
    file has two cumulative lines, thread 1's one; the global file (written
    with outputBigFloatWeights) has three.  Its first line equals a run of the
    first 1e6 paths alone (independent chains, same draws) and its last the
    experiment's average, to the FP32 rounding of the per-path log10 weights, and thread sums add up to the global ones."""
    N, L, K, sigma, paths = 20, 1.2, 20, 0.1, 2_500_000
    cfg = config(N=N, L=L, paths=paths, K=K, sigma=sigma, write=True, outdir=str(tmp_path / "full"))
    cfg["experiment"]["experimentParams"]["maxPerturbThreads"] = 2
    cfg["experiment"]["experimentParams"]["outputBigFloatWeights"] = True
    r = host.run_config(cfg)
    first = host.run_config(config(N=N, L=L, paths=1_000_000, K=K, sigma=sigma))
    d = list((tmp_path / "full").rglob("ConvergenceWeights"))
    assert len(d) == 1
    t0 = _convergence_lines(d[0] / "ConvergenceWeights_0.txt")
    t1 = _convergence_lines(d[0] / "ConvergenceWeights_1.txt")
    g = _convergence_lines(d[0] / "ConvergenceWeights.txt")
    assert sorted(p.name for p in d[0].iterdir()) == ["ConvergenceWeights.txt", "ConvergenceWeights_0.txt",
                                                        "ConvergenceWeights_1.txt"]
    print(f"[convergence] thread0 {t0} thread1 {t1} global {g}", flush=True)
    valid = paths - r["failed_paths"]
    assert [n for n, _ in g][-1] == valid and len(t0) == 2 and len(t1) == 1 and len(g) == 3
    assert g[0][0] == 1_000_000 - first["failed_paths"]
    # the files sum the per-path FP32 log10 weights, the run its device partials
    assert abs(math.log10(g[0][1] * g[0][0]) - first["log10_experiment_weight"]) < 1e-5
    assert abs(math.log10(g[-1][1] * valid) - r["log10_experiment_weight"]) < 1e-5
    assert t0[0] == g[0] and t0[1] == g[1]
    assert t1[0][0] == g[2][0] - g[1][0]
    assert abs(t1[0][1] * t1[0][0] + g[1][1] * g[1][0] - g[2][1] * g[2][0]) < 1e-9 * g[2][1] * g[2][0]
