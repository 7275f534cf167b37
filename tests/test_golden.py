"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit (a change to the oracle
that moves these numbers is caught here), Philox4x32-10 against the published
known-answer vectors, the host table builder against the stored table.
GPU: the HIP path against the fixtures with the forward and adjoint tolerances
(per-path 1e-5 relative on paths whose joints stay in the FP64 bin).
"""
import json
import os

import numpy as np
import pytest

from tests import cases
from tests.lerp import joint_log10w
from twisty_amd import host

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PATH_FIXTURES = [("paths_c1.npz", cases.C1), ("paths_c2.npz", cases.C2), ("paths_n33.npz", cases.MID)]
ADJ_FIXTURES = [("adjoint_n4.npz", cases.SMALL), ("adjoint_n33.npz", cases.MID)]


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_philox_known_answers(oracle):
    with open(os.path.join(GOLD, "philox_kat.json")) as f:
        kat = json.load(f)
    assert len(kat) == 3
    for v in kat:
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"]


@pytest.mark.parametrize("name,case", PATH_FIXTURES)
def test_oracle_reproduces_path_fixture(oracle, name, case):
    g = load(name)
    assert list(g["case"]) == [case.N, case.K, case.bins, case.numStepsInt, case.seed]
    table, pos32 = cases.setup(case, oracle)
    assert np.array_equal(g["positions"], pos32)
    assert np.array_equal(g["table_head"], table.table[:64]) and np.array_equal(g["table_tail"], table.table[-64:])
    n = len(g["log10w"])
    r = oracle.simulate(pos32.astype(np.float64), case.N, case.L, case.K, n, table, case.absorption, seed=case.seed,
                        want_curv=True, want_exit=True)
    assert np.array_equal(r["log10w"], g["log10w"]) and np.array_equal(r["curv"], g["curv"])
    assert np.array_equal(r["exit"], g["exit"])
    # the sum is merged per thread, so the thread count moves the last bits
    assert abs(r["log10_sum_w_exact"] - g["log10_sum_w"]) <= 1e-12 * abs(g["log10_sum_w"])


@pytest.mark.parametrize("name,case", PATH_FIXTURES)
def test_host_table_matches_fixture(name, case):
    g = load(name)
    ds = cases.ds_of(case)
    t = host.build_rt_table(ds, case.mu, float(np.float32(case.sigma) ** 2), case.scatter, case.absorption,
                            numStepsInt=case.numStepsInt, numCurvatureSteps=case.bins)
    assert np.array_equal(t.table[:64], g["table_head"]) and np.array_equal(t.table[-64:], g["table_tail"])
    assert t.table.astype(np.float64).sum() == g["table_sum"]
    assert np.array_equal(np.array([t.ds, t.minK, t.maxK, t.step], np.float32), g["grid"])


@pytest.mark.parametrize("name,case", ADJ_FIXTURES)
def test_oracle_reproduces_adjoint_fixture(oracle, name, case):
    g = load(name)
    table, pos32 = cases.setup(case, oracle)
    r = oracle.adjoint(pos32.astype(np.float64), case.N, case.L, case.K, int(g["paths"]), table, case.absorption,
                       seed=case.seed, lerp_mode=0, threads=1)
    # per-thread partial sums are added in thread order: one thread vs many
    # changes only the last bits
    assert np.allclose(r["d_table"], g["d_table"], rtol=1e-12, atol=1e-15)
    assert np.allclose(r["d_seg"], g["d_seg"], rtol=1e-10, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("name,case", PATH_FIXTURES)
def test_gpu_paths_match_fixture(hip_device, oracle, name, case):
    g = load(name)
    n = len(g["log10w"])
    table, _ = cases.setup(case, oracle)
    with cases.hip_context(case, oracle) as ctx:
        r = ctx.run(n, case.K, seed=case.seed, per_path=True, per_joint=True)
    _, bg = joint_log10w(r["curv"], table, case.absorption)
    _, br = joint_log10w(g["curv"], table, case.absorption)
    clean = ~(bg != br).any(axis=1) & g["valid"]
    rel = np.abs(r["log10w"] - g["log10w"]) / np.abs(g["log10w"])
    print(f"[golden {name}] same-bin paths {clean.sum()}/{n}, max rel {rel[clean].max():.2e}", flush=True)
    assert clean.sum() > n // 2 and rel[clean].max() <= 1e-5
    assert np.array_equal(np.isnan(r["log10w"]), ~g["valid"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,case", ADJ_FIXTURES)
def test_gpu_adjoint_matches_fixture(hip_device, oracle, name, case):
    g = load(name)
    with cases.hip_context(case, oracle) as ctx:
        r = ctx.run_adjoint(int(g["paths"]), case.K, seed=case.seed, kappa=0.0)  # exact sweep: the fixture's draws
    e_seg = np.linalg.norm(r["d_seg"] - g["d_seg"]) / np.linalg.norm(g["d_seg"])
    e_run = np.abs(np.cumsum(r["d_table"]) - np.cumsum(g["d_table"])).max() / np.abs(g["d_table"]).sum()
    print(f"[golden {name}] dSeg rel {e_seg:.2e}, dT running-sum rel {e_run:.2e}", flush=True)
    assert e_seg <= 5e-3 and e_run <= 1e-3
