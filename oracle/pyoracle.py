"""ctypes binding of oracle/_build/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(dt):
    return np.ctypeslib.ndpointer(dtype=dt, flags="C_CONTIGUOUS")


def _declare(L):
    u32, u64, f32, f64, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_double, C.c_int32
    L.oracle_philox4x32_10.argtypes = [_p(np.uint32), _p(np.uint32), _p(np.uint32)]
    L.oracle_draw.argtypes = [u64, u32, u64, u32, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(f32)]
    L.oracle_rotate_vector_by_quaternion.argtypes = [_p(np.float64), _p(np.float64)]
    L.oracle_rotate_quat_by_quat.argtypes = [_p(np.float64), _p(np.float64), _p(np.float64)]
    L.oracle_apply_perturbation_f64.argtypes = [_p(np.float64), u32, u32, f32]
    L.oracle_tangents_curvatures_f64.argtypes = [_p(np.float64), u32, f32, _p(np.float64), _p(np.float64)]
    L.oracle_inv_ds.argtypes = [u32, f32]
    L.oracle_inv_ds.restype = f32
    L.oracle_initial_arc.argtypes = [_p(np.float64), f32, u32, _p(np.float64), C.POINTER(f64)]
    L.oracle_initial_arc.restype = C.c_int
    L.oracle_initial_arc_bulge.argtypes = [_p(np.float64), f32, u32, _p(np.float64), _p(np.float64)]
    L.oracle_initial_arc_bulge.restype = C.c_int
    L.oracle_minimum_arclength.argtypes = [_p(np.float64), u32]
    L.oracle_minimum_arclength.restype = f32
    L.oracle_norm.argtypes = [u32, f32, _p(np.float64), f32, u32, u32]
    L.oracle_norm.restype = f64
    L.oracle_gaussian_phase.argtypes = [f32, f32]
    L.oracle_gaussian_phase.restype = f32
    L.oracle_integrand_rt.argtypes = [f64, f64, f64, f64, f64]
    L.oracle_integrand_rt.restype = f64
    L.oracle_table_range.argtypes = [u32, f32, C.POINTER(f32), C.POINTER(f32), C.POINTER(f32)]
    L.oracle_build_table.argtypes = [f32, f32, f32, f32, f32, u32, u32, f32, f32, f32, C.c_int,
                                     _p(np.float32), _p(np.float32)]
    L.oracle_build_table.restype = C.c_int
    L.oracle_integrate.argtypes = [f32, f32, f32, f32, f32, u32, f32, f32, f32]
    L.oracle_integrate.restype = f32
    L.oracle_weight_curvatures.argtypes = [_p(np.float32), u32, _p(np.float32), i32, f32, f32, f32, f32, f32,
                                           C.POINTER(f64)]
    L.oracle_weight_curvatures.restype = C.c_int
    L.oracle_accumulate_log10.argtypes = [_p(np.float64), u64]
    L.oracle_accumulate_log10.restype = f64
    L.oracle_accumulate_combine_log10.argtypes = [_p(np.float64), u64, u64]
    L.oracle_accumulate_combine_log10.restype = f64
    L.oracle_simulate.argtypes = [C.c_int, _p(np.float64), u32, f32, u32, u32, u32, u64, u32, u64, u64,
                                  _p(np.float32), i32, f32, f32, f32, f32, f32, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _p(np.float64)]
    L.oracle_simulate.restype = C.c_int
    L.oracle_exit_bin.argtypes = [_p(np.float64), _p(np.float64), u32, u32, C.POINTER(u32)]
    L.oracle_end_frame.argtypes = [_p(np.float64), _p(np.float64), _p(np.float64)]
    L.oracle_adjoint.argtypes = [_p(np.float64), u32, f32, u32, u64, u64, _p(np.float32), i32, f32, f32, f32, f32,
                                 f32, C.c_int, C.c_int, C.POINTER(f64), C.c_void_p, C.c_void_p, C.c_void_p]
    L.oracle_adjoint.restype = C.c_int
    L.oracle_generate_levels_c.argtypes = [u32, C.POINTER(u32)]
    L.oracle_generate_levels_c.restype = C.c_int
    L.oracle_generate.argtypes = [_p(np.float64), f32, u32, u64, u64, u64, _p(np.float32), i32, f32, f32, f32, f32,
                                  f32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  _p(np.float64)]
    L.oracle_generate.restype = C.c_int
    L.oracle_generate_multi.argtypes = [_p(np.float64), u32, u64, u64, u64, _p(np.float32), i32, _p(np.float32), u32,
                                        f32, f32, f32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, _p(np.float64)]
    L.oracle_generate_multi.restype = C.c_int
    L.oracle_generate_chains.argtypes = [_p(np.float64), u32, u64, u64, u64, u64, _p(np.float32), i32,
                                         _p(np.float32), u32, f32, f32, f32, C.c_double, C.c_int] + \
        [C.c_void_p] * 9 + [_p(np.float64)]
    L.oracle_generate_chains.restype = C.c_int
    L.oracle_bd_log.argtypes = [C.c_double]
    L.oracle_bd_log.restype = C.c_double
    L.oracle_chain_ds.argtypes = [C.c_double, f32, f32, f32]
    L.oracle_chain_ds.restype = C.c_float
    L.oracle_chain_log10_sampling_weight.argtypes = [f32, f32, f32, u32]
    L.oracle_chain_log10_sampling_weight.restype = C.c_double


# ---------------------------------------------------------------------------
# Thin Python helpers
# ---------------------------------------------------------------------------

def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def draw(seed, stream, path, k, N):
    d, l, x = C.c_uint32(), C.c_uint32(), C.c_float()
    lib().oracle_draw(seed, stream, path, k, N, C.byref(d), C.byref(l), C.byref(x))
    return d.value, l.value, x.value


def bc12(start=(0, 0, 0), sdir=(0, 0, 1), end=(0, 0, 1), edir=(0, 0, 1)):
    return np.array(list(start) + list(sdir) + list(end) + list(edir), np.float64)


def initial_arc(N, arclength, bc=None):
    bc = bc12() if bc is None else np.asarray(bc, np.float64)
    pos = np.zeros((N + 1) * 3, np.float64)
    alpha = C.c_double()
    rc = lib().oracle_initial_arc(bc, arclength, N, pos, C.byref(alpha))
    if rc != 0:
        raise ValueError("infeasible initial arc")
    return pos.reshape(N + 1, 3), alpha.value


def initial_arc_bulge(N, arclength, bc, bulge):
    pos = np.zeros((N + 1) * 3, np.float64)
    if lib().oracle_initial_arc_bulge(np.asarray(bc, np.float64), arclength, N, np.asarray(bulge, np.float64), pos):
        raise ValueError("infeasible initial arc")
    return pos.reshape(N + 1, 3)


def table_range(nbins, ds):
    a, b, c = C.c_float(), C.c_float(), C.c_float()
    lib().oracle_table_range(nbins, ds, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


class Table:
    def __init__(self, table, meta):
        self.table = table
        self.ds, self.minK, self.maxK, self.step, self.minW, self.maxW = [float(v) for v in meta]


def build_table(ds, mu=1.0, eps=0.0025, scatter=0.1, absorption=0.004, bias=1.0, numStepsInt=20000,
                numCurvatureSteps=4096, minBound=0.0, maxBound=None, threads=None):
    if maxBound is None:
        maxBound = np.float32(10.0) / np.float32(eps)  # ExperimentRunner.cpp:168-171
    threads = threads or os.cpu_count() or 1
    t = np.zeros(numCurvatureSteps, np.float32)
    meta = np.zeros(6, np.float32)
    rc = lib().oracle_build_table(mu, eps, scatter, absorption, bias, numStepsInt, numCurvatureSteps,
                                  minBound, float(maxBound), ds, threads, t, meta)
    if rc != 0:
        raise ValueError("bad table parameters")
    return Table(t, meta)


def weight_curvatures(curv, table, absorption):
    curv = np.ascontiguousarray(curv, np.float32)
    w = C.c_double()
    ok = lib().oracle_weight_curvatures(curv, len(curv), table.table, len(table.table), table.ds, table.minK,
                                        table.maxK, table.step, absorption, C.byref(w))
    return bool(ok), w.value


def simulate(init_pos, N, arclength, K, n_paths, table, absorption, seed=1, stream=0, path_offset=0,
             chain_length=1, precision=64, threads=None, want_curv=False, want_exit=False, want_w=True, burn_in=0):
    threads = threads or os.cpu_count() or 1
    init = np.ascontiguousarray(init_pos, np.float64).reshape(-1)
    res = np.zeros(8, np.float64)
    w = np.zeros(n_paths, np.float64) if want_w else None
    v = np.zeros(n_paths, np.uint8) if want_w else None
    cv = np.zeros(n_paths * (N - 1), np.float32) if want_curv else None
    ex = np.zeros(n_paths * 3, np.float32) if want_exit else None
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    rc = lib().oracle_simulate(precision, init, N, arclength, K, chain_length, burn_in, seed, stream, path_offset,
                               n_paths, table.table, len(table.table), table.ds, table.minK, table.maxK,
                               table.step, absorption, threads, ptr(w), ptr(v), ptr(cv), ptr(ex), res)
    if rc != 0:
        raise ValueError("oracle_simulate rejected its arguments")
    out = {"log10_sum_w": res[0], "log10_sum_w2": res[1], "max_log10_w": res[2], "n_valid": int(res[3]),
           "n_invalid": int(res[4]), "n_clamped": int(res[5]), "seconds": res[6], "log10_sum_w_exact": res[7]}
    if want_w:
        out["log10w"] = w
        out["valid"] = v.astype(bool)
    if want_curv:
        out["curv"] = cv.reshape(n_paths, N - 1)
    if want_exit:
        out["exit"] = ex.reshape(n_paths, 3)
    return out


def adjoint(init_pos, N, arclength, K, n_paths, table, absorption, seed=1, lerp_mode=0, want_table=True,
            want_seg=True, want_w=False, threads=None, table_values=None):
    """FP64 estimator E = log10(sum w / n_paths) and its derivatives (oracle_adjoint.cpp).

    table_values overrides table.table (for finite differences in the entries).
    Returns dict: log10_sum_w, d_table (bins), d_seg (N, 3), log10w (paths).
    """
    threads = threads or os.cpu_count() or 1
    init = np.ascontiguousarray(init_pos, np.float64).reshape(-1)
    tv = np.ascontiguousarray(table.table if table_values is None else table_values, np.float32)
    out = C.c_double()
    dT = np.zeros(len(tv), np.float64) if want_table else None
    dS = np.zeros(3 * N, np.float64) if want_seg else None
    w = np.zeros(n_paths, np.float64) if want_w else None
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    rc = lib().oracle_adjoint(init, N, arclength, K, seed, n_paths, tv, len(tv), table.ds, table.minK, table.maxK,
                              table.step, absorption, lerp_mode, threads, C.byref(out), ptr(dT), ptr(dS), ptr(w))
    if rc != 0:
        raise ValueError("oracle_adjoint rejected its arguments")
    res = {"log10_sum_w": out.value}
    if want_table:
        res["d_table"] = dT
    if want_seg:
        res["d_seg"] = dS.reshape(N, 3)
    if want_w:
        res["log10w"] = w
    return res


def generate_levels(N):
    """(levels, leaf span) of the independent-path recursion, or None when N is not 2 + 2^k {2, 3}."""
    leaf = C.c_uint32()
    k = lib().oracle_generate_levels_c(N, C.byref(leaf))
    return None if k < 0 else (k, leaf.value)


def generate(bc, arclength, N, n_paths, table, absorption, seed=1, path_offset=0, threads=None, want_w=True,
             want_pos=False, want_joints=False):
    """Independent-path estimator (oracle_generate.cpp): per-path log10 weights and the sums;
    want_joints adds the float joint curvatures ("curv", NaN rows for invalid paths) and each
    path's smallest branch-decision margin ("margin")."""
    threads = threads or os.cpu_count() or 1
    b = np.ascontiguousarray(bc, np.float64)
    res = np.zeros(6, np.float64)
    w = np.zeros(n_paths, np.float64) if want_w else None
    v = np.zeros(n_paths, np.uint8) if want_w else None
    pos = np.zeros(n_paths * (N + 1) * 3, np.float64) if want_pos else None
    curv = np.full(n_paths * (N - 1), np.nan, np.float32) if want_joints else None
    margin = np.zeros(n_paths, np.float64) if want_joints else None
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    rc = lib().oracle_generate(b, arclength, N, seed, path_offset, n_paths, table.table, len(table.table), table.ds,
                               table.minK, table.maxK, table.step, absorption, threads, ptr(w), ptr(v), ptr(pos),
                               ptr(curv), ptr(margin), res)
    if rc != 0:
        raise ValueError("oracle_generate rejected its arguments")
    out = {"log10_sum_w": res[0], "n_valid": int(res[1]), "n_invalid": int(res[2]), "min_log10_w": res[3],
           "max_log10_w": res[4], "seconds": res[5]}
    if want_w:
        out["log10w"] = w
        out["valid"] = v.astype(bool)
    if want_pos:
        out["positions"] = pos.reshape(n_paths, N + 1, 3)
    if want_joints:
        out["curv"] = curv.reshape(n_paths, N - 1)
        out["margin"] = margin
    return out


def generate_multi(bc, N, n_paths, tables, meta, absorption, min_ds, max_ds, seed=1, path_offset=0, threads=None):
    """Multi-arclength independent-path estimator (oracle_generate.cpp): tables (T, bins), meta (T, 4).
    Also returns the float joint curvatures ("curv"), each path's smallest branch-decision margin
    ("margin") and the table each path is weighted with ("table")."""
    threads = threads or os.cpu_count() or 1
    t = np.ascontiguousarray(tables, np.float32)
    m = np.ascontiguousarray(meta, np.float32)
    res = np.zeros(6, np.float64)
    w = np.zeros(n_paths, np.float64)
    v = np.zeros(n_paths, np.uint8)
    curv = np.full(n_paths * (N - 1), np.nan, np.float32)
    margin = np.zeros(n_paths, np.float64)
    tab = np.zeros(n_paths, np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib().oracle_generate_multi(np.ascontiguousarray(bc, np.float64), N, seed, path_offset, n_paths, t,
                                     t.shape[1], m, t.shape[0], absorption, min_ds, max_ds, threads, p(w), p(v),
                                     p(curv), p(margin), p(tab), res)
    if rc != 0:
        raise ValueError("oracle_generate_multi rejected its arguments")
    return {"log10_sum_w": res[0], "n_valid": int(res[1]), "n_invalid": int(res[2]), "min_log10_w": res[3],
            "max_log10_w": res[4], "seconds": res[5], "log10w": w, "valid": v.astype(bool),
            "curv": curv.reshape(n_paths, N - 1), "margin": margin, "table": tab}


def generate_chains(bc, N, n_chains, chain_length, tables, meta, extinction, min_ds, max_ds, norm_log10=0.0, seed=1,
                    chain_offset=0, threads=None):
    """Metropolis-Hastings chains of the two-table independent-path estimator
    (oracle_generate_chains): tables (T, bins), meta (T, 4).  Per proposal
    (chain-major): log10 weight, validity, the chain's weight after it, joint
    curvatures, branch margin, acceptance margin, u, ds and table."""
    threads = threads or os.cpu_count() or 1
    t = np.ascontiguousarray(tables, np.float32)
    m = np.ascontiguousarray(meta, np.float32)
    n = n_chains * chain_length
    res = np.zeros(6, np.float64)
    w = np.zeros(n, np.float64)
    v = np.zeros(n, np.uint8)
    ch = np.zeros(n, np.float64)
    curv = np.full(n * (N - 1), np.nan, np.float32)
    margin = np.zeros(n, np.float64)
    am = np.zeros(n, np.float64)
    u = np.zeros(n, np.float64)
    ds = np.zeros(n, np.float32)
    tab = np.zeros(n, np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib().oracle_generate_chains(np.ascontiguousarray(bc, np.float64), N, seed, chain_offset, n_chains,
                                      chain_length, t, t.shape[1], m, t.shape[0], extinction, min_ds, max_ds,
                                      norm_log10, threads, p(w), p(v), p(ch), p(curv), p(margin), p(am), p(u), p(ds),
                                      p(tab), res)
    if rc != 0:
        raise ValueError("oracle_generate_chains rejected its arguments")
    return {"log10_sum_w": res[0], "n_valid": int(res[1]), "n_invalid": int(res[2]), "min_log10_w": res[3],
            "max_log10_w": res[4], "seconds": res[5], "log10w": w, "valid": v.astype(bool), "chain": ch,
            "curv": curv.reshape(n, N - 1), "margin": margin, "accept_margin": am, "u": u, "ds": ds,
            "table": tab}


def bd_log(y):
    return lib().oracle_bd_log(float(y))


def chain_ds(u, extinction, min_ds, max_ds):
    return lib().oracle_chain_ds(float(u), extinction, min_ds, max_ds)


def chain_log10_sampling_weight(extinction, ds, min_ds, N):
    return lib().oracle_chain_log10_sampling_weight(extinction, ds, min_ds, N)


def exit_bins(tangents, e, n_theta, n_phi):
    """Exit-direction bin of each FP64 tangent (rows of `tangents`)."""
    e = np.ascontiguousarray(e, np.float64)
    out = np.zeros(len(tangents), np.uint32)
    b = C.c_uint32()
    for i, t in enumerate(np.ascontiguousarray(tangents, np.float64)):
        lib().oracle_exit_bin(np.ascontiguousarray(t), e, n_theta, n_phi, C.byref(b))
        out[i] = b.value
    return out
