"""ctypes binding of libtwisty_host.so (include/twisty_host.h): weight tables,
initial curve and the ExperimentRunner flow of the C++ host layer."""
import ctypes as C
import json
import os

import numpy as np

from . import capi

LIB_PATH = os.path.join(capi.HERE, "lib", "libtwisty_host.so")
EXPORTED = ["twisty_host_build_rt_table", "twisty_host_build_rt_tables", "twisty_host_initial_arc", "twisty_host_minimum_arclength",
            "twisty_host_norm_log10", "twisty_host_norm_log10_steps", "twisty_host_run_config", "twisty_host_rt_table_eps_derivative",
            "twisty_host_initial_arc_exit_jacobian", "twisty_host_cached_rt_tables",
            "twisty_host_write_path_batches", "twisty_host_run_experiments", "twisty_host_path_generation_mc",
            "twisty_host_path_generation_mc_multi", "twisty_host_path_generation_mc_chains",
            "twisty_host_experiment_gradients"]

_lib = None


class HostResult(C.Structure):
    _fields_ = [
        ("log10_experiment_weight", C.c_double),
        ("log10_avg_path_weight", C.c_double),
        ("log10_sum_squares", C.c_double),
        ("total_paths", C.c_uint64),
        ("failed_paths", C.c_uint64),
        ("clamped_joints", C.c_uint64),
        ("setup_ms", C.c_double),
        ("perturb_ms", C.c_double),
        ("weighting_ms", C.c_double),
        ("total_ms", C.c_double),
        ("kernel_ms", C.c_double),
    ]


class Gradients(C.Structure):
    _fields_ = [
        ("log10_estimate", C.c_double),
        ("d_eps", C.c_double),
        ("d_theta", C.c_double),
        ("d_phi", C.c_double),
        ("kernel_ms", C.c_double),
        ("d_theta_roulette_std", C.c_double),
        ("d_theta_mc_std", C.c_double),
        ("d_phi_roulette_std", C.c_double),
        ("d_phi_mc_std", C.c_double),
        ("paths_swept", C.c_uint64),
    ]


class GenerationResult(C.Structure):
    _fields_ = [
        ("log10_total_weight", C.c_double),
        ("min_log10_w", C.c_double),
        ("max_log10_w", C.c_double),
        ("num_valid", C.c_uint64),
        ("num_total", C.c_uint64),
        ("kernel_ms", C.c_double),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m twisty_amd.build`")
        L = C.CDLL(LIB_PATH)
        f32, u32, f64, vp = C.c_float, C.c_uint32, C.c_double, C.c_void_p
        L.twisty_host_build_rt_table.argtypes = [f32, f32, f32, f32, f32, u32, u32, f32, f32, f32, C.c_int, vp, vp]
        L.twisty_host_build_rt_tables.argtypes = [u32, vp, vp, vp, f32, f32, f32, f32, u32, u32, f32, C.c_int, vp,
                                                  vp]
        L.twisty_host_initial_arc.argtypes = [vp, u32, vp]
        L.twisty_host_minimum_arclength.argtypes = [vp, u32]
        L.twisty_host_minimum_arclength.restype = f32
        L.twisty_host_norm_log10.argtypes = [u32, f32, vp]
        L.twisty_host_norm_log10.restype = f64
        L.twisty_host_norm_log10_steps.argtypes = [u32, f32, vp, u32, u32]
        L.twisty_host_norm_log10_steps.restype = f64
        L.twisty_host_run_config.argtypes = [C.c_char_p, C.POINTER(HostResult), C.c_char_p, u32]
        L.twisty_host_rt_table_eps_derivative.argtypes = [f32, f32, f32, f32, u32, u32, f32, f32, C.c_int, f32,
                                                          C.c_int, vp]
        L.twisty_host_initial_arc_exit_jacobian.argtypes = [vp, u32, vp, vp]
        L.twisty_host_cached_rt_tables.argtypes = [C.c_char_p, f32, f32, f32, f32, f32, u32, u32, f32, f32, f32, f32,
                                                   u32, C.c_int, vp, vp, C.POINTER(C.c_int), C.c_char_p, u32]
        L.twisty_host_run_experiments.argtypes = [C.c_char_p, vp, vp, u32, C.POINTER(HostResult),
                                                  C.POINTER(C.c_double), C.c_char_p, u32]
        L.twisty_host_path_generation_mc.argtypes = [C.c_char_p, C.c_uint64, f64, C.POINTER(GenerationResult),
                                                     C.c_char_p, u32]
        L.twisty_host_path_generation_mc_multi.argtypes = [C.c_char_p, C.c_uint64, f64, f32, f32, u32, f32,
                                                           C.c_char_p, C.POINTER(GenerationResult), C.c_char_p, u32]
        L.twisty_host_path_generation_mc_chains.argtypes = [C.c_char_p, C.c_uint64, f64, f32, f32, u32, f32,
                                                            C.c_uint64, C.c_char_p, C.POINTER(GenerationResult),
                                                            C.c_char_p, u32]
        L.twisty_host_experiment_gradients.argtypes = [C.c_char_p, C.POINTER(Gradients), vp, u32, C.c_char_p, u32]
        L.twisty_host_write_path_batches.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, vp, u32, vp, vp,
                                                     C.c_uint64, f64, C.c_char_p, u32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class WeightTable:
    """A built WeightLookupTableIntegral: table values and its curvature grid."""

    def __init__(self, table, meta):
        self.table = table
        self.ds, self.minK, self.maxK, self.step, self.minW, self.maxW = [float(v) for v in meta]


def build_rt_table(ds, mu, eps, scatter, absorption, bias=1.0, numStepsInt=20000, numCurvatureSteps=4096,
                   minBound=0.0, maxBound=None, threads=0):
    """WeightLookupTableIntegral(params, ds).Initialize(); maxBound defaults to
    10/eps as ExperimentRunner.cpp:168-171 sets it."""
    if maxBound is None:
        maxBound = float(np.float32(10.0) / np.float32(eps))
    t = np.zeros(numCurvatureSteps, np.float32)
    meta = np.zeros(6, np.float32)
    rc = lib().twisty_host_build_rt_table(mu, eps, scatter, absorption, bias, numStepsInt, numCurvatureSteps,
                                          minBound, maxBound, ds, threads, _p(t), _p(meta))
    if rc != 0:
        raise ValueError("twisty_host_build_rt_table rejected its arguments")
    return WeightTable(t, meta)


def build_rt_tables(ds, eps, mu, scatter, absorption, bias=1.0, numStepsInt=20000, numCurvatureSteps=4096,
                    minBound=0.0, threads=0):
    """One table per (ds[k], eps[k]) pair, each as build_rt_table builds it
    (maxBound 10/eps[k]); tables built side by side, one per host thread.
    Returns (tables [count, numCurvatureSteps], meta [count, 6])."""
    ds = np.ascontiguousarray(ds, np.float32)
    eps = np.ascontiguousarray(eps, np.float32)
    if ds.shape != eps.shape or ds.ndim != 1:
        raise ValueError("ds and eps must be 1-D arrays of one length")
    n = len(ds)
    t = np.zeros((n, numCurvatureSteps), np.float32)
    meta = np.zeros((n, 6), np.float32)
    rc = lib().twisty_host_build_rt_tables(n, _p(eps), _p(ds), None, mu, scatter, absorption, bias, numStepsInt,
                                           numCurvatureSteps, minBound, threads, _p(t), _p(meta))
    if rc != 0:
        raise ValueError("twisty_host_build_rt_tables rejected its arguments")
    return t, meta


def cached_rt_tables(min_ds, max_ds, num_ds, mu, eps, scatter, absorption, bias=1.0, numStepsInt=20000,
                     numCurvatureSteps=4096, minBound=0.0, maxBound=None, cache_root=None, threads=0):
    """CachedMultiArclengthWeightLookupTable: (list of WeightTable, loaded_from_cache, cache path)."""
    if maxBound is None:
        maxBound = float(np.float32(10.0) / np.float32(eps))
    t = np.zeros((num_ds, numCurvatureSteps), np.float32)
    meta = np.zeros((num_ds, 6), np.float32)
    loaded = C.c_int()
    msg = C.create_string_buffer(4096)
    root = None if cache_root is None else str(cache_root).encode()
    rc = lib().twisty_host_cached_rt_tables(root, mu, eps, scatter, absorption, bias, numStepsInt, numCurvatureSteps,
                                            minBound, maxBound, min_ds, max_ds, num_ds, threads, _p(t), _p(meta),
                                            C.byref(loaded), msg, 4096)
    if rc != 0:
        raise ValueError(f"twisty_host_cached_rt_tables: {msg.value.decode()}")
    return [WeightTable(t[i], meta[i]) for i in range(num_ds)], bool(loaded.value), msg.value.decode()


def rt_table_eps_derivative(ds, mu, eps, scatter, bias=1.0, numStepsInt=20000, numCurvatureSteps=4096,
                            minBound=0.0, maxBound=None, threads=0):
    """d T / d eps (FP64) of build_rt_table's table; maxBound = None means 10/eps
    and is differentiated along with eps."""
    follow = maxBound is None
    if maxBound is None:
        maxBound = float(np.float32(10.0) / np.float32(eps))
    d = np.zeros(numCurvatureSteps, np.float64)
    rc = lib().twisty_host_rt_table_eps_derivative(mu, eps, scatter, bias, numStepsInt, numCurvatureSteps, minBound,
                                                   maxBound, int(follow), ds, threads, _p(d))
    if rc != 0:
        raise ValueError("twisty_host_rt_table_eps_derivative rejected its arguments")
    return d


def initial_arc_exit_jacobian(boundary, N):
    """(d positions / d theta, d positions / d phi), each (N+1, 3) FP64, for the
    end direction (sin theta cos phi, sin theta sin phi, cos theta)."""
    b = np.ascontiguousarray(boundary, np.float32)
    a = np.zeros((N + 1) * 3, np.float64)
    c = np.zeros((N + 1) * 3, np.float64)
    if lib().twisty_host_initial_arc_exit_jacobian(_p(b), N, _p(a), _p(c)) != 0:
        raise ValueError("no initial arc for these boundary conditions")
    return a.reshape(N + 1, 3), c.reshape(N + 1, 3)


def boundary13(start, start_dir, end, end_dir, arclength):
    return np.array(list(start) + list(start_dir) + list(end) + list(end_dir) + [arclength], np.float32)


def initial_arc(boundary, N):
    b = np.ascontiguousarray(boundary, np.float32)
    out = np.zeros((N + 1) * 3, np.float32)
    if lib().twisty_host_initial_arc(_p(b), N, _p(out)) != 0:
        raise ValueError("no initial arc: arclength below the minimum for these boundary conditions")
    return out.reshape(N + 1, 3)


def minimum_arclength(boundary, N):
    b = np.ascontiguousarray(boundary, np.float32)
    return lib().twisty_host_minimum_arclength(_p(b), N)


def norm_log10(N, ds, boundary, outer_steps=None, inner_steps=None):
    """log10 of the path normaliser; outer_steps / inner_steps override the
    node counts of the N = 5, 6 nested quadrature (reference: 10000 / 1000)."""
    b = np.ascontiguousarray(boundary, np.float32)
    if outer_steps is None and inner_steps is None:
        return lib().twisty_host_norm_log10(N, ds, _p(b))
    return lib().twisty_host_norm_log10_steps(N, ds, _p(b), 10000 if outer_steps is None else outer_steps,
                                              1000 if inner_steps is None else inner_steps)


def write_path_batches(experiment_dir, boundary, positions, log10w, norm_log10, prepend="",
                       experiment_name="Default_Experiment"):
    """Path-batch export files of (P, N+1, 3) float32 paths and their log10 weights (NaN = invalid)."""
    b = np.ascontiguousarray(boundary, np.float32)
    pos = np.ascontiguousarray(positions, np.float32)
    w = np.ascontiguousarray(log10w, np.float32)
    n, P, _ = pos.shape
    if w.shape != (n,):
        raise ValueError("one log10 weight per path")
    err = C.create_string_buffer(1024)
    if lib().twisty_host_write_path_batches(str(experiment_dir).encode(), prepend.encode(), experiment_name.encode(),
                                            _p(b), P - 1, _p(pos), _p(w), n, float(norm_log10), err, 1024) != 0:
        raise RuntimeError(f"write_path_batches failed: {err.value.decode()}")


def read_path_batches(experiment_dir, prepend=""):
    """Reads the files of write_path_batches (or of a runner with outputPathBatches) back."""
    d = os.path.join(str(experiment_dir), "GeneratedCurves")
    raw = np.fromfile(os.path.join(d, "BoundaryConditions.bcf"), np.uint8)
    if raw.size != 56:
        raise ValueError(f"BoundaryConditions.bcf has {raw.size} bytes, expected 56")
    vecs = raw[:48].view(np.float32).reshape(4, 3)
    arclength = float(raw[48:52].view(np.float32)[0])
    N = int(raw[52:56].view(np.uint32)[0])
    with open(os.path.join(d, prepend + "Paths_Metadata.pmd")) as f:
        batches = [tuple(int(t) for t in line.split()) for line in f if line.strip()]
    with open(os.path.join(d, "index.json")) as f:
        index = json.load(f)
    return {
        "start": vecs[0], "start_dir": vecs[1], "end": vecs[2], "end_dir": vecs[3], "arclength": arclength, "N": N,
        "batches": batches, "index": index,
        "positions": np.fromfile(os.path.join(d, prepend + "Paths_Binary.pbd"), np.float32).reshape(-1, N + 1, 3),
        "weights": np.fromfile(os.path.join(d, prepend + "Log10PathWeights_Binary.bdt"), np.float64),
        "text_weights": np.loadtxt(os.path.join(d, prepend + "Log10PathWeights_Text.pwt"), ndmin=1),
        "five_segment": np.fromfile(os.path.join(d, prepend + "FiveSegment_Binary.bdt"), np.float32).reshape(-1, 3),
    }


def run_experiments(config, boundaries, seeds):
    """Experiments sharing a config's experimentParams, one batched launch.

    boundaries: (E, 13) BoundaryConditions_CudaSafe rows with the arclength;
    seeds: E seeds (0 = clock).  Returns (list of result dicts, device ms)."""
    text = config if isinstance(config, str) else json.dumps(config)
    b = np.ascontiguousarray(boundaries, np.float32).reshape(-1, 13)
    s = np.ascontiguousarray(seeds, np.uint64)
    n = b.shape[0]
    if s.shape != (n,):
        raise ValueError("one seed per experiment")
    res = (HostResult * n)()
    ms = C.c_double()
    err = C.create_string_buffer(1024)
    if lib().twisty_host_run_experiments(text.encode(), _p(b), _p(s), n, res, C.byref(ms), err, 1024) != 0:
        raise RuntimeError(f"run_experiments failed: {err.value.decode()}")
    return [{k: getattr(r, k) for k, _ in r._fields_} for r in res], ms.value


def path_generation_mc(config, seed=1, norm_log10=0.0):
    """ExperimentBase::MSegmentPathGenerationMC on a reference-schema config (independent paths)."""
    text = config if isinstance(config, str) else json.dumps(config)
    res = GenerationResult()
    err = C.create_string_buffer(1024)
    if lib().twisty_host_path_generation_mc(text.encode(), seed, norm_log10, C.byref(res), err, 1024) != 0:
        raise RuntimeError(f"path_generation_mc failed: {err.value.decode()}")
    return {k: getattr(res, k) for k, _ in res._fields_}


def path_generation_mc_multi(config, min_ds_tables, max_ds_tables, num_ds_steps, max_ds, seed=1, norm_log10=0.0,
                             cache_root=None):
    """The multi-arclength overload: ds drawn per path, nearest cached table."""
    text = config if isinstance(config, str) else json.dumps(config)
    res = GenerationResult()
    err = C.create_string_buffer(1024)
    root = None if cache_root is None else str(cache_root).encode()
    if lib().twisty_host_path_generation_mc_multi(text.encode(), seed, norm_log10, min_ds_tables, max_ds_tables,
                                                  num_ds_steps, max_ds, root, C.byref(res), err, 1024) != 0:
        raise RuntimeError(f"path_generation_mc_multi failed: {err.value.decode()}")
    return {k: getattr(res, k) for k, _ in res._fields_}


def path_generation_mc_chains(config, min_ds_tables, max_ds_tables, num_ds_steps, max_ds, chain_length=64, seed=1,
                              norm_log10=0.0, cache_root=None):
    """The two-table overload (the NoisyCircle experiments' estimator):
    Metropolis-Hastings chains over the cached tables with absorption + scatter."""
    text = config if isinstance(config, str) else json.dumps(config)
    res = GenerationResult()
    err = C.create_string_buffer(1024)
    root = None if cache_root is None else str(cache_root).encode()
    if lib().twisty_host_path_generation_mc_chains(text.encode(), seed, norm_log10, min_ds_tables, max_ds_tables,
                                                   num_ds_steps, max_ds, chain_length, root, C.byref(res), err,
                                                   1024) != 0:
        raise RuntimeError(f"path_generation_mc_chains failed: {err.value.decode()}")
    return {k: getattr(res, k) for k, _ in res._fields_}


def experiment_gradients(config, table_bins=None):
    """dE/d eps, dE/d(theta, phi) and dE/dT of a config's experiment (C++ chain of the adjoint)."""
    text = config if isinstance(config, str) else json.dumps(config)
    res = Gradients()
    err = C.create_string_buffer(1024)
    dT = np.zeros(table_bins, np.float64) if table_bins else None
    if lib().twisty_host_experiment_gradients(text.encode(), C.byref(res), None if dT is None else _p(dT),
                                              table_bins or 0, err, 1024) != 0:
        raise RuntimeError(f"experiment_gradients failed: {err.value.decode()}")
    out = {k: getattr(res, k) for k, _ in res._fields_}
    if dT is not None:
        out["d_table"] = dT
    return out


def run_config(config):
    """Runs a reference-schema JSON config (dict or text) on the HIP runner."""
    text = config if isinstance(config, str) else json.dumps(config)
    res = HostResult()
    err = C.create_string_buffer(1024)
    if lib().twisty_host_run_config(text.encode(), C.byref(res), err, 1024) != 0:
        raise RuntimeError(f"run_config failed: {err.value.decode()}")
    return {k: getattr(res, k) for k, _ in res._fields_}
