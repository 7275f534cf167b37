"""CPU tests of the product's host layer and C ABI surface (no device needed)."""
import ctypes as C
import json
import math
import os
import re

import numpy as np
import pytest

from tests import cases
from twisty_amd import capi, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(twisty_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,libpath,listed", [
    ("twisty_hip.h", capi.LIB_PATH, capi.EXPORTED),
    ("twisty_host.h", host.LIB_PATH, host.EXPORTED),
])
def test_library_exports_every_declared_symbol(header, libpath, listed):
    names = declared_functions(header)
    assert names and sorted(listed) == names
    lib = C.CDLL(libpath)
    for n in names:
        assert hasattr(lib, n), n


def test_no_device_paths_fail_cleanly():
    n = C.c_int(-1)
    assert capi.lib().twisty_hip_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        assert capi.lib().twisty_hip_create(0, C.byref(h)) == capi.ERR_ARG
        assert not h.value
    assert capi.lib().twisty_hip_create(-1, None) == capi.ERR_ARG
    assert capi.lib().twisty_hip_destroy(None) == capi.ERR_ARG
    assert capi.lib().twisty_hip_last_error(None) == b"null context"
    assert capi.lib().twisty_hip_version() == 100


@pytest.mark.parametrize("N,L,sigma,mu,bins,steps", [
    (64, 1.5, 0.05, 1.0, 4096, 20000),
    (20, 1.2, 0.1, 1.0, 4096, 20000),
    (100, 1.3, 0.08, 1.0, 4096, 20000),
    (66, 15.0, None, 0.1, 1000, 20000),   # the reference's NoisyCircle settings (eps 0.075)
    (33, 1.4, 0.07, 0.5, 1024, 2000),
])
def test_table_builder_bit_exact_vs_oracle(oracle, N, L, sigma, mu, bins, steps):
    ds = float(np.float32(L) / np.float32(N))
    eps = float(np.float32(sigma) ** 2) if sigma else 0.075
    a = oracle.build_table(ds, mu=mu, eps=eps, scatter=0.1, absorption=0.004, numStepsInt=steps,
                           numCurvatureSteps=bins)
    b = host.build_rt_table(ds, mu, eps, 0.1, 0.004, numStepsInt=steps, numCurvatureSteps=bins)
    assert np.array_equal(a.table.view(np.uint32), b.table.view(np.uint32))
    assert (a.ds, a.minK, a.maxK, a.step, a.minW, a.maxW) == (b.ds, b.minK, b.maxK, b.step, b.minW, b.maxW)


def test_table_builder_rejects_odd_steps():
    with pytest.raises(ValueError):
        host.build_rt_table(0.02, 1.0, 0.0025, 0.1, 0.004, numStepsInt=2001, numCurvatureSteps=64)


@pytest.mark.parametrize("N,L,edir", [(64, 1.5, (0, 0, 1)), (20, 1.2, (0, 0, 1)), (30, 2.2, (0.6, 0.0, 0.8)),
                                      (1024, 2.0, (0, 0, 1))])
def test_initial_arc_vs_oracle(oracle, N, L, edir):
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), edir, L)
    got = host.initial_arc(bc, N)
    ref, _ = oracle.initial_arc(N, L, oracle.bc12(edir=edir))
    assert np.abs(got - ref).max() < 4e-7
    ds = float(np.float32(L) / np.float32(N))
    seg = np.linalg.norm(np.diff(got.astype(np.float64), axis=0), axis=1)
    # float positions of size ~1 carry ~6e-8 rounding, so segment lengths do too
    assert np.abs(seg - ds).max() < 3e-7


def test_minimum_arclength_and_infeasible_arc(oracle):
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0.3, 0.0, 1.0), (0.6, 0.0, 0.8), 1.0)
    lmin = host.minimum_arclength(bc, 24)
    assert lmin == pytest.approx(oracle.lib().oracle_minimum_arclength(oracle.bc12(end=(0.3, 0, 1), edir=(0.6, 0, 0.8)), 24))
    bc[12] = lmin * 1.01
    host.initial_arc(bc, 24)
    bc[12] = lmin * 0.99
    with pytest.raises(ValueError):
        host.initial_arc(bc, 24)


def test_norm(oracle):
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 1.6)
    assert host.norm_log10(64, 1.5 / 64, bc) == 0.0
    ref = oracle.lib().oracle_norm(4, 0.4, oracle.bc12(), 1.6, 2, 2)
    assert math.isclose(host.norm_log10(4, 0.4, bc), math.log10(ref), rel_tol=1e-9)


def test_norm_nested_n5_n6_against_oracle(oracle):
    """The host's N = 5, 6 nested quadrature (thread-split outer rule) equals the
    oracle's at reduced node counts, and K5 = K6 as in the reference's K(N)
    macro (PathWeightUtils.cpp:689-711)."""
    ds = 0.4
    for D in (1.0, 1.8, 2.6):
        bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, D), (0, 0, 1), 2.0)
        bc12 = np.array([0, 0, 0, 0, 0, 1, 0, 0, D, 0, 0, 1], dtype=np.float64)
        for outer, inner in ((60, 60), (97, 41)):
            ref = math.log10(oracle.lib().oracle_norm(5, ds, bc12, 2.0, outer, inner))
            for N in (5, 6):
                assert abs(host.norm_log10(N, ds, bc, outer, inner) - ref) < 1e-13, (D, N, outer, inner)
    assert math.isnan(host.norm_log10(5, ds, bc, 0, 10))


def test_norm_n5_reference_node_counts_near_closed_form():
    """At the reference's 10000 x 1000 x 1000 nodes the N = 5 normaliser is
    (4 pi)^5 p_5(z) / ds^3 (Treloar's 5-step random-flight density) up to the
    rule's first-order error (0.46% here)."""
    from tests.test_oracle import random_flight_density

    ds, D = 0.4, 1.4
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, D), (0, 0, 1), 2.0)
    f = np.float32
    z = float(abs(f(f(D) * (f(1.0) / f(ds))) - f(2.0)))
    closed = math.log10((4 * math.pi) ** 5 * random_flight_density(5, z) / ds ** 3)
    v = host.norm_log10(5, ds, bc)
    assert -0.01 < 10 ** (v - closed) - 1 < 0.0, (v, closed)


def _config(**over):
    cfg = {"experiment": {
        "experimentParams": {
            "pathsToGenerate": 4096, "pathsToSkip": 0, "numSegments": 20, "name": "t", "experimentDir": "/tmp/tw",
            "maxPerturbThreads": 0, "maxWeightThreads": 0, "outputBigFloatWeights": False,
            "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
            "random": {"bootstrapSeed": 1, "perturbSeed": 1},
            "weighting": {"weightFunction": 0, "mu": 1.0, "eps": 0.01, "numStepsInt": 2000,
                          "numCurvatureSteps": 512, "absorption": 0.004, "scatter": 0.1},
            "twisty_hip": {"perturbationsPerPath": 10, "chainLength": 1, "writeFiles": False}},
        "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1],
                            "endDir": [0, 0, 1], "arclength": 1.2}}}
    for k, v in over.items():
        cfg["experiment"]["experimentParams"][k] = v
    return cfg


def test_run_config_reports_errors_without_gpu():
    if capi.device_count() > 0:
        pytest.skip("a device is visible; the GPU tests cover the successful run")
    with pytest.raises(RuntimeError, match="twisty_hip_create"):
        host.run_config(_config())


def test_run_config_rejects_bad_configs():
    with pytest.raises(RuntimeError, match="JSON parse error"):
        host.run_config("{ not json")
    cfg = _config()
    del cfg["experiment"]["experimentParams"]["numSegments"]
    with pytest.raises(RuntimeError, match="numSegments"):
        host.run_config(cfg)
    with pytest.raises(RuntimeError, match="weightFunction 0"):
        cfg = _config()
        cfg["experiment"]["experimentParams"]["weighting"]["weightFunction"] = 1
        host.run_config(cfg)
    with pytest.raises(RuntimeError, match="minimum"):
        cfg = _config()
        cfg["experiment"]["basicExperiment"]["arclength"] = 0.5
        host.run_config(cfg)


def test_reference_config_schema_parses():
    """A config from the reference's corpus layout parses (run stops at the device)."""
    cfg = _config()
    text = json.dumps(cfg)
    try:
        host.run_config(text)
    except RuntimeError as e:
        assert "twisty_hip_create" in str(e) or capi.device_count() > 0


def test_sample_config_fields_parse():
    """The reference's sample configs (configs/Samples/RingBenchmark.json) give
    "scatterValues": [0.1] instead of "scatter" and pathsToSkip 100: both parse."""
    cfg = _config(pathsToSkip=100)
    w = cfg["experiment"]["experimentParams"]["weighting"]
    del w["scatter"]
    w["scatterValues"] = [0.1]
    try:
        host.run_config(cfg)
    except RuntimeError as e:
        assert "twisty_hip_create" in str(e) or capi.device_count() > 0
    del w["scatterValues"]
    with pytest.raises(RuntimeError, match="scatter"):
        host.run_config(cfg)


def test_cached_multi_arclength_tables_round_trip(tmp_path, oracle):
    """CachedMultiArclengthWeightLookupTable: compute + write, then reload from
    the .cwt file (the reference's Tests/WeightTableTest.cpp checks the reload
    to 1e-4; here it is bit-exact), each table equal to a directly built one."""
    nseg, lo, hi, n = 66, 10.0, 20.0, 4
    min_ds, max_ds = float(np.float32(lo) / np.float32(nseg)), float(np.float32(hi) / np.float32(nseg))
    kw = dict(mu=0.1, eps=0.075, scatter=0.1, absorption=0.04, numStepsInt=2000, numCurvatureSteps=1000,
              cache_root=tmp_path)
    computed, loaded1, path = host.cached_rt_tables(min_ds, max_ds, n, **kw)
    assert not loaded1 and os.path.exists(path) and path.endswith(".cwt")
    # file = n records of (table floats, min weight, max weight)
    assert os.path.getsize(path) == n * (1000 + 2) * 4
    reloaded, loaded2, path2 = host.cached_rt_tables(min_ds, max_ds, n, **kw)
    assert loaded2 and path2 == path
    step = np.float32(np.float32(max_ds) - np.float32(min_ds)) / np.float32(n - 1)
    for i, (a, b) in enumerate(zip(computed, reloaded)):
        assert np.array_equal(a.table, b.table) and a.minW == b.minW and a.maxW == b.maxW
        ds = float(np.float32(min_ds) + np.float32(i) * step)
        assert a.ds == ds == b.ds
        d = host.build_rt_table(ds, 0.1, 0.075, 0.1, 0.04, numStepsInt=2000, numCurvatureSteps=1000)
        assert np.array_equal(a.table, d.table)
    csvs = list((tmp_path / "CachedWeightTables").rglob("*_RT.csv"))
    assert len(csvs) == n


def test_cached_tables_reject_bad_grid(tmp_path):
    with pytest.raises(ValueError):
        host.cached_rt_tables(0.2, 0.1, 4, 1.0, 0.01, 0.1, 0.004, numStepsInt=200, numCurvatureSteps=64,
                              cache_root=tmp_path)
    with pytest.raises(ValueError):
        host.cached_rt_tables(0.1, 0.2, 1, 1.0, 0.01, 0.1, 0.004, numStepsInt=200, numCurvatureSteps=64,
                              cache_root=tmp_path)


def _five_segment_angles(p):
    """phi1, theta1, theta2 of FullExperimentRunnerOptimalPerturb.cpp:756-793, in FP64."""
    unit = lambda v: v / np.linalg.norm(v, axis=-1, keepdims=True)
    s1 = unit(p[:, 2] - p[:, 1])
    z = unit(p[:, 4] - p[:, 2])
    other = np.where((np.abs(z[:, 0]) >= 0.99)[:, None], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0])
    x = unit(np.cross(z, other))
    y = np.cross(z, x)
    v = unit(p[:, 3] - p[:, 2])
    return np.stack([np.arccos(s1[:, 2]), np.arctan2(s1[:, 1], s1[:, 0]),
                     np.arctan2((v * y).sum(1), (v * x).sum(1))], axis=1)


@pytest.mark.parametrize("N,n", [(5, 25000), (12, 3000)])
def test_path_batch_export_round_trip(tmp_path, N, n):
    """Files of the path-batch export read back: layout, skipped invalid paths,
    linear weights 10^(log10 w + normaliser), every 10000th weight as text, the
    five-segment angles (N = 5; zeros otherwise)."""
    rng = np.random.default_rng(N)
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0.3, 0, 1), (0.6, 0, 0.8), 1.7)
    arc = host.initial_arc(bc, N)
    pos = (arc[None] + 0.05 * rng.standard_normal((n, N + 1, 3))).astype(np.float32)
    w = (-40 + rng.standard_normal(n)).astype(np.float32)
    w[rng.random(n) < 0.1] = np.nan
    norm = 37.25
    host.write_path_batches(tmp_path, bc, pos, w, norm, prepend="pre_", experiment_name='exp "q"')
    r = host.read_path_batches(tmp_path, prepend="pre_")
    ok = ~np.isnan(w)
    m = int(ok.sum())
    assert r["N"] == N and r["arclength"] == np.float32(1.7)
    assert np.array_equal(r["start"], bc[0:3]) and np.array_equal(r["end_dir"], bc[9:12])
    assert r["batches"] == [(0, 0, m)] and r["index"] == {"experiment_name": 'exp "q"', "path_batch_links": []}
    assert np.array_equal(r["positions"], pos[ok])
    want = 10.0 ** (w[ok].astype(np.float64) + norm)
    assert np.allclose(r["weights"], want, rtol=1e-14, atol=0)
    assert np.allclose(r["text_weights"], want[::10000], rtol=1e-5)
    if N == 5:
        assert np.abs(r["five_segment"] - _five_segment_angles(pos[ok].astype(np.float64))).max() < 1e-4
    else:
        assert r["five_segment"].shape == (m, 3) and not r["five_segment"].any()


def test_path_batch_export_splits_batches(tmp_path):
    """More than 1e6 valid paths: a second batch line and contiguous data."""
    n = 1_000_000 + 4321
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 1.2)
    pos = np.broadcast_to(host.initial_arc(bc, 4)[None], (n, 5, 3)).copy()
    pos[:, 2, 0] = np.arange(n, dtype=np.float32) * 1e-7
    w = np.full(n, -30.0, np.float32)
    host.write_path_batches(tmp_path, bc, pos, w, 0.0)
    r = host.read_path_batches(tmp_path)
    assert r["batches"] == [(0, 0, 1_000_000), (0, 1, 4321)]
    assert np.array_equal(r["positions"][:, 2, 0], pos[:, 2, 0]) and len(r["weights"]) == n
    assert len(r["text_weights"]) == 100 + 1


def test_path_batch_export_reports_unwritable_dir(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 1.2)
    with pytest.raises(RuntimeError, match="write_path_batches failed"):
        host.write_path_batches(blocker, bc, np.zeros((1, 5, 3), np.float32), np.zeros(1, np.float32), 0.0)


def _mt19937_raw(seed, n):
    """std::mt19937(seed) outputs (init_genrand seeding, as NumPy's legacy seeding)."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    return [int(x) for x in bg.random_raw(n)]


def _nonzero_stream(seed, n):
    out, raw = [], iter(_mt19937_raw(seed, 4 * n + 8))
    while len(out) < n:
        v = next(raw)
        if v:
            out.append(v)
    return out


def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995
    assert _mt19937_raw(5489, 10000)[-1] == 4123659995


def _beam_config(tmp_path, **over):
    cfg = _config(numSegments=40, pathsToGenerate=2000)
    cfg["experiment"]["experimentParams"]["random"] = {"bootstrapSeed": 5, "perturbSeed": 7}
    cfg["experiment"]["beamSpreadExperiment"] = {
        "targetSeed": 1, "sphereRadius": 3.0, "maxArclength": 9.0, "startPos": [0.0, 0.0, 0.0],
        "startDir": [1.0, 0.0, 0.0], "numInitialCurves": 2, "numPerInitialCurve": 3, **over}
    p = tmp_path / "beam.json"
    p.write_text(json.dumps(cfg))
    return cfg, p


def _run_driver(name, cfg_path, *args, cwd=None):
    import subprocess
    exe = os.path.join(os.path.dirname(host.LIB_PATH), name)
    return subprocess.run([exe, str(cfg_path), *args], capture_output=True, text=True, timeout=120, cwd=cwd)


def test_beam_spread_plan(tmp_path):
    """BeamSpread geometry and seeds (BeamSpreadExperiment.cpp:183-230): receiver
    direction from mt19937(targetSeed) uniform floats, nonzero seed streams,
    midpoint arclengths above the minimum arclength."""
    cfg, p = _beam_config(tmp_path, numArclengths=2)
    out = _run_driver("twisty_beam_spread", p, "--plan", cwd=tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    u = [np.float32(x) / np.float32(2.0 ** 32) for x in _mt19937_raw(1, 2)]
    n1 = u[0] * np.float32(2) - np.float32(1)
    rxy = np.sqrt(np.float32(1) - n1 * n1)
    ang = np.float32(2) * np.float32(3.141592653589793) * u[1]
    end_dir = np.array([n1, np.cos(ang) * rxy, np.sin(ang) * rxy], np.float64)
    got = np.array([float(t) for t in next(l for l in lines if l.startswith("endDir:")).split()[1:]])
    assert np.abs(got - end_dir / np.linalg.norm(end_dir)).max() < 1e-5
    bc = host.boundary13((0, 0, 0), (1, 0, 0), 3.0 * end_dir, end_dir / np.linalg.norm(end_dir), 0.0)
    mn = host.minimum_arclength(bc, 40)
    exps = [l.split() for l in lines if l.startswith("experiment ")]
    assert len(exps) == 2 * 2 * 3
    ic = _nonzero_stream(5, 2)
    pc = _nonzero_stream(7, 3)
    for e in exps:
        a, i, j = int(e[1]), int(e[2]), int(e[3])
        assert float(e[5]) == pytest.approx(mn + (9.0 - mn) / 2 * (a + 0.5), rel=1e-5)
        assert int(e[7]) == (ic[i] << 32) | pc[j]


def test_ring_benchmark_plan(tmp_path):
    """RingBenchmark geometry (RingBenchmark.cpp:186-197) for every height."""
    cfg = _config()
    cfg["experiment"]["ringBenchmark"] = {"startPos": [0.0, 0.0, 0.0], "startDir": [1.0, 0.0, 0.0], "zMin": 3.0,
                                          "zMax": 10.0, "numZValues": 8, "ringRadius": 5}
    p = tmp_path / "ring.json"
    p.write_text(json.dumps(cfg))
    out = _run_driver("twisty_ring_benchmark", p, "--plan")
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.startswith("experiment ")]
    assert len(rows) == 8
    for z, r in enumerate(rows):
        zz = 3.0 + 1.0 * z
        assert float(r[5]) == pytest.approx(zz, rel=1e-6) and float(r[3]) == 5.0
        assert float(r[7]) == pytest.approx(max(math.hypot(5.0, zz) * 1.1, 3.0), rel=1e-6)


def test_run_experiments_errors(tmp_path):
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 1.2)
    short = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1), 0.5)
    with pytest.raises(RuntimeError, match="experiment 1: no initial curve"):
        host.run_experiments(_config(), np.stack([bc, short]), [1, 2])
    with pytest.raises(ValueError, match="one seed per experiment"):
        host.run_experiments(_config(), np.stack([bc, bc]), [1])
    if capi.device_count() == 0:
        with pytest.raises(RuntimeError, match="twisty_hip_create"):
            host.run_experiments(_config(), np.stack([bc, bc]), [1, 2])


def test_path_generation_mc_without_gpu():
    if capi.device_count() > 0:
        pytest.skip("a device is visible; the GPU tests cover the successful run")
    cfg = _config(numSegments=26)
    with pytest.raises(RuntimeError, match="twisty_hip_create"):
        host.path_generation_mc(cfg)


def test_path_generation_mc_multi_minimum_ds(tmp_path):
    """Multi-arclength overload: its minimum arclength (the quadratic of
    ExperimentBase.cpp:1437-1462) is the end distance for a straight geometry,
    so maxDs just below D / N gives the reference's empty result
    (:1466-1469) and just above reaches the device."""
    N, D = 26, 1.3
    cfg = _config(numSegments=N, pathsToGenerate=1000)
    cfg["experiment"]["basicExperiment"]["endPos"] = [0, 0, D]
    empty = host.path_generation_mc_multi(cfg, 0.03, 0.07, 5, D / N * 0.999, cache_root=tmp_path)
    assert empty["num_total"] == 0 and empty["num_valid"] == 0
    if capi.device_count() == 0:
        with pytest.raises(RuntimeError, match="twisty_hip_create"):
            host.path_generation_mc_multi(cfg, 0.03, 0.07, 5, D / N * 1.001, cache_root=tmp_path)


def test_path_generation_mc_chains_minimum_ds(tmp_path):
    """Two-table overload: its minimum ds is CalculateMinimumArclength / N (the
    end distance for a straight geometry), so maxDs below it gives the
    reference's empty result (ExperimentBase.cpp:1639-1642) and above it
    reaches the device."""
    N, D = 26, 1.3
    cfg = _config(numSegments=N, pathsToGenerate=1000)
    cfg["experiment"]["basicExperiment"]["endPos"] = [0, 0, D]
    empty = host.path_generation_mc_chains(cfg, 0.03, 0.12, 5, D / N * 0.999, cache_root=tmp_path)
    assert empty["num_total"] == 0 and empty["num_valid"] == 0
    if capi.device_count() == 0:
        with pytest.raises(RuntimeError, match="twisty_hip_create"):
            host.path_generation_mc_chains(cfg, 0.03, 0.12, 5, D / N * 1.001, cache_root=tmp_path)


def test_experiment_gradients_errors():
    with pytest.raises(RuntimeError, match="dTableLen"):
        host.experiment_gradients(_config(), table_bins=7)
    if capi.device_count() == 0:
        with pytest.raises(RuntimeError, match="twisty_hip_create"):
            host.experiment_gradients(_config())
