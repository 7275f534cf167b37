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

from twisty_amd import capi, host

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def config(N=20, L=1.2, paths=1 << 16, K=50, sigma=0.1, chain=1, seed=1, bins=4096, write=False, outdir="/tmp"):
    return {"experiment": {
        "experimentParams": {
            "pathsToGenerate": paths, "pathsToSkip": 0, "numSegments": N, "name": "hostgpu",
            "experimentDir": outdir, "maxPerturbThreads": 0, "maxWeightThreads": 0,
            "outputBigFloatWeights": False, "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
            "random": {"bootstrapSeed": 1, "perturbSeed": seed},
            "weighting": {"weightFunction": 0, "mu": 1.0, "eps": float(np.float32(sigma) ** 2),
                          "numStepsInt": 20000, "numCurvatureSteps": bins, "absorption": 0.004, "scatter": 0.1},
            "twisty_hip": {"perturbationsPerPath": K, "chainLength": chain, "writeFiles": write}},
        "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1],
                            "endDir": [0, 0, 1], "arclength": L}}}


def oracle_same_inputs(oracle, N, L, K, paths, sigma, chain=1, seed=1, bins=4096):
    ds = float(np.float32(L) / np.float32(N))
    eps = float(np.float32(sigma) ** 2)
    t = oracle.build_table(ds, mu=1.0, eps=eps, scatter=0.1, absorption=0.004, numStepsInt=20000,
                           numCurvatureSteps=bins)
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), L)
    pos = host.initial_arc(bc, N).astype(np.float64)
    return oracle.simulate(pos, N, L, K, paths, t, 0.004, seed=seed, chain_length=chain)


@pytest.mark.parametrize("N,L,K,sigma,chain", [(20, 1.2, 50, 0.1, 1), (64, 1.5, 100, 0.05, 1), (64, 1.5, 1, 0.05, 256)])
def test_runner_matches_oracle(hip_device, oracle, N, L, K, sigma, chain):
    paths = 1 << 15
    r = host.run_config(config(N=N, L=L, paths=paths, K=K, sigma=sigma, chain=chain))
    ref = oracle_same_inputs(oracle, N, L, K, paths, sigma, chain=chain)
    ref_avg = ref["log10_sum_w_exact"] - math.log10(paths)
    print(f"[runner N={N} K={K} L={chain}] gpu log10 avg={r['log10_avg_path_weight']:.9f} "
          f"oracle={ref_avg:.9f} kernel_ms={r['kernel_ms']:.3f} setup_ms={r['setup_ms']}", flush=True)
    assert r["total_paths"] == paths and r["failed_paths"] == 0
    assert abs(10 ** (r["log10_avg_path_weight"] - ref_avg) - 1.0) < 1e-3


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
