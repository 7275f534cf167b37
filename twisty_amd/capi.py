"""ctypes binding of libtwisty_hip.so (include/twisty_hip.h).

This is the Python face of the C ABI; it does no computation of its own.  If
the library is missing the import of `lib()` raises: there is no CPU fallback.
"""
import ctypes as C
import math
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtwisty_hip.so")

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED = 0, 1, 2, 3, 4

# Every symbol include/twisty_hip.h declares.
EXPORTED = [
    "twisty_hip_version", "twisty_hip_device_count", "twisty_hip_create", "twisty_hip_destroy",
    "twisty_hip_last_error", "twisty_hip_set_tables", "twisty_hip_set_curve", "twisty_hip_run",
    "twisty_hip_set_schedule", "twisty_hip_get_schedule", "twisty_hip_last_timing",
    "twisty_hip_set_weight_histogram", "twisty_hip_get_weight_histogram",
    "twisty_hip_set_keep_final_state", "twisty_hip_get_final_positions", "twisty_hip_set_burn_in",
    "twisty_hip_run_generated", "twisty_hip_set_adjoint_roulette", "twisty_hip_last_adjoint_paths",
    "twisty_hip_set_multi_tables", "twisty_hip_run_generated_multi", "twisty_hip_run_generated_chains",
    "twisty_hip_set_exit_histogram", "twisty_hip_get_exit_histogram", "twisty_hip_run_adjoint",
    "twisty_hip_set_adjoint_schedule", "twisty_hip_last_adjoint_timing",
    "twisty_hip_batch_create", "twisty_hip_batch_destroy", "twisty_hip_batch_set_schedule",
    "twisty_hip_batch_get_schedule", "twisty_hip_batch_run", "twisty_hip_batch_autotune",
    "twisty_hip_run_noalloc", "twisty_hip_set_adjoint_directions", "twisty_hip_last_adjoint_direction_stats",
]


class Experiment(C.Structure):
    _fields_ = [
        ("num_paths", C.c_uint64),
        ("seed", C.c_uint64),
        ("num_segments", C.c_uint32),
        ("perturbations_per_path", C.c_uint32),
        ("table_index", C.c_uint32),
        ("positions_offset", C.c_uint32),
        ("arclength", C.c_float),
        ("chain_length", C.c_uint32),
        ("burn_in", C.c_uint32),
    ]


class BatchSchedule(C.Structure):
    _fields_ = [
        ("block_size", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("segments_per_lane", C.c_uint32),
        ("tile_paths", C.c_uint32),
        ("lean", C.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Partials(C.Structure):
    _fields_ = [
        ("log10_sum_w", C.c_double),
        ("log10_sum_w2", C.c_double),
        ("max_log10_w", C.c_double),
        ("num_paths", C.c_uint64),
        ("num_valid", C.c_uint64),
        ("num_invalid", C.c_uint64),
        ("num_clamped", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TwistyHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"twisty_hip error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m twisty_amd.build`")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, u32, u64, f32, i = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_int
    L.twisty_hip_version.restype = i
    L.twisty_hip_device_count.argtypes = [C.POINTER(i)]
    L.twisty_hip_create.argtypes = [i, C.POINTER(vp)]
    L.twisty_hip_destroy.argtypes = [vp]
    L.twisty_hip_last_error.argtypes = [vp]
    L.twisty_hip_last_error.restype = C.c_char_p
    L.twisty_hip_set_tables.argtypes = [vp, vp, u32, u32, f32, f32, f32, f32, f32]
    L.twisty_hip_set_curve.argtypes = [vp, vp, u32, vp]
    L.twisty_hip_run.argtypes = [vp, u64, u64, u32, u32, u64, C.POINTER(Partials), vp, vp]
    L.twisty_hip_run_noalloc.argtypes = [vp, u64, u64, u32, u32, u64, C.POINTER(Partials)]
    L.twisty_hip_set_schedule.argtypes = [vp, u32, u32, u32]
    L.twisty_hip_get_schedule.argtypes = [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    L.twisty_hip_last_timing.argtypes = [vp, C.POINTER(f32), C.POINTER(f32)]
    L.twisty_hip_set_weight_histogram.argtypes = [vp, u32, C.c_double, C.c_double, C.c_double]
    L.twisty_hip_set_keep_final_state.argtypes = [vp, C.c_int]
    L.twisty_hip_set_burn_in.argtypes = [vp, u32]
    L.twisty_hip_set_adjoint_roulette.argtypes = [vp, f32]
    L.twisty_hip_last_adjoint_paths.argtypes = [vp, C.POINTER(u64)]
    L.twisty_hip_set_multi_tables.argtypes = [vp, vp, u32, u32, vp, f32, u32]
    L.twisty_hip_run_generated_multi.argtypes = [vp, u64, u64, u64, f32, f32, vp, C.POINTER(C.c_double), vp, vp]
    L.twisty_hip_run_generated_chains.argtypes = [vp, u64, u64, u64, u64, f32, f32, f32, C.c_double, vp,
                                                  C.POINTER(C.c_double), vp, vp, vp]
    L.twisty_hip_run_generated.argtypes = [vp, u64, u64, u64, vp, C.POINTER(C.c_double), vp, vp]
    L.twisty_hip_get_final_positions.argtypes = [vp, vp, u64]
    L.twisty_hip_get_weight_histogram.argtypes = [vp, vp]
    L.twisty_hip_set_exit_histogram.argtypes = [vp, u32, u32]
    L.twisty_hip_get_exit_histogram.argtypes = [vp, vp, vp, u64]
    L.twisty_hip_run_adjoint.argtypes = [vp, u64, u64, u32, u64, C.POINTER(Partials), vp, vp]
    L.twisty_hip_set_adjoint_schedule.argtypes = [vp, u32, u32, u32]
    L.twisty_hip_set_adjoint_directions.argtypes = [vp, vp, u32]
    L.twisty_hip_last_adjoint_direction_stats.argtypes = [vp, vp, vp, vp]
    L.twisty_hip_last_adjoint_timing.argtypes = [vp, C.POINTER(f32)]
    L.twisty_hip_batch_create.argtypes = [vp, vp, u32, vp, u32, u32, vp, vp, u64, C.POINTER(vp)]
    L.twisty_hip_batch_destroy.argtypes = [vp]
    L.twisty_hip_batch_set_schedule.argtypes = [vp, C.POINTER(BatchSchedule)]
    L.twisty_hip_batch_get_schedule.argtypes = [vp, C.POINTER(BatchSchedule)]
    L.twisty_hip_batch_run.argtypes = [vp, vp, vp, C.POINTER(f32)]
    L.twisty_hip_batch_autotune.argtypes = [vp, u32, C.POINTER(BatchSchedule), vp, vp, u32, C.POINTER(u32)]


def device_count():
    n = C.c_int(0)
    lib().twisty_hip_device_count(C.byref(n))
    return n.value


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context: stream, device buffers, tables and initial curve."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().twisty_hip_create(device, C.byref(self._h))
        if rc != OK:
            raise TwistyHipError(rc, f"twisty_hip_create(device={device}) failed")

    def close(self):
        for b in list(getattr(self, "_batches", [])):
            b.close()
        if self._h:
            lib().twisty_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise TwistyHipError(rc, lib().twisty_hip_last_error(self._h).decode())

    def set_tables(self, tables, minK, maxK, step, ds, absorption):
        t = np.ascontiguousarray(tables, np.float32)
        if t.ndim == 1:
            t = t[None, :]
        self._check(lib().twisty_hip_set_tables(self._h, _ptr(t), t.shape[1], t.shape[0], minK, maxK, step, ds,
                                                absorption))
        self.table_bins = t.shape[1]

    def set_curve(self, positions, boundary):
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        bc = np.ascontiguousarray(boundary, np.float32).reshape(13)
        n = pos.size // 3 - 1
        self._check(lib().twisty_hip_set_curve(self._h, _ptr(pos), n, _ptr(bc)))
        self.N = n

    def set_schedule(self, lanes_per_path=0, block_size=0, blocks_per_cu=0):
        self._check(lib().twisty_hip_set_schedule(self._h, lanes_per_path, block_size, blocks_per_cu))

    def get_schedule(self):
        g, b, c, l = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(lib().twisty_hip_get_schedule(self._h, C.byref(g), C.byref(b), C.byref(c), C.byref(l)))
        return {"lanes_per_path": g.value, "block_size": b.value, "blocks_per_cu": c.value, "lds_bytes": l.value}

    def run(self, num_paths, K, seed=1, path_offset=0, chain_length=1, per_path=False, per_joint=False):
        out = Partials()
        pp = np.empty(num_paths, np.float32) if per_path else None
        pj = np.empty(num_paths * (self.N - 1), np.float32) if per_joint else None
        self._check(lib().twisty_hip_run(self._h, num_paths, path_offset, K, chain_length, seed, C.byref(out),
                                         _ptr(pp), _ptr(pj)))
        res = out.as_dict()
        if per_path:
            res["log10w"] = pp
        if per_joint:
            res["curv"] = pj.reshape(num_paths, self.N - 1)
        return res

    def run_noalloc(self, num_paths, K, seed=1, path_offset=0, chain_length=1):
        """twisty_hip_run_noalloc: the estimate of run() with no device
        allocation (partials in the library's static device scratch)."""
        out = Partials()
        self._check(lib().twisty_hip_run_noalloc(self._h, num_paths, path_offset, K, chain_length, seed,
                                                 C.byref(out)))
        return out.as_dict()

    def set_adjoint_roulette(self, kappa):
        """Russian roulette over paths before the adjoint (0 = exact, every path)."""
        self._check(lib().twisty_hip_set_adjoint_roulette(self._h, kappa))

    def last_adjoint_paths(self):
        v = C.c_uint64()
        self._check(lib().twisty_hip_last_adjoint_paths(self._h, C.byref(v)))
        return v.value

    def run_generated(self, num_paths, seed=1, path_offset=0, per_path=False, per_joint=False):
        """Independent-path estimator (paths built by recursive sphere intersection)."""
        out = Partials()
        mn = C.c_double()
        pp = np.empty(num_paths, np.float32) if per_path else None
        pj = np.empty(num_paths * (self.N - 1), np.float32) if per_joint else None
        self._check(lib().twisty_hip_run_generated(self._h, num_paths, path_offset, seed, C.byref(out), C.byref(mn),
                                                   _ptr(pp), _ptr(pj)))
        res = out.as_dict()
        res["min_log10_w"] = mn.value
        if per_path:
            res["log10w"] = pp
        if per_joint:
            res["curv"] = pj.reshape(num_paths, self.N - 1)
        return res

    def set_multi_tables(self, tables, meta, absorption, num_segments):
        """Multi-ds tables: tables (T, bins), meta (T, 4) = ds, minK, maxK, step."""
        t = np.ascontiguousarray(tables, np.float32)
        m = np.ascontiguousarray(meta, np.float32)
        self._check(lib().twisty_hip_set_multi_tables(self._h, _ptr(t), t.shape[1], t.shape[0], _ptr(m), absorption,
                                                      num_segments))

    def run_generated_multi(self, num_paths, min_ds, max_ds, seed=1, path_offset=0, per_path=False,
                            per_joint=False):
        """Independent paths with ds drawn per path and the nearest cached table."""
        out = Partials()
        mn = C.c_double()
        pp = np.empty(num_paths, np.float32) if per_path else None
        pj = np.empty(num_paths * (self.N - 1), np.float32) if per_joint else None
        self._check(lib().twisty_hip_run_generated_multi(self._h, num_paths, path_offset, seed, min_ds, max_ds,
                                                         C.byref(out), C.byref(mn), _ptr(pp), _ptr(pj)))
        res = out.as_dict()
        res["min_log10_w"] = mn.value
        if per_path:
            res["log10w"] = pp
        if per_joint:
            res["curv"] = pj.reshape(num_paths, self.N - 1)
        return res

    def run_generated_chains(self, num_chains, chain_length, min_ds, max_ds, extinction, norm_log10=0.0, seed=1,
                             chain_offset=0, per_path=False, per_joint=False):
        """Metropolis-Hastings chains of the two-table independent-path estimator
        (twisty_hip_run_generated_chains): per proposal (chain-major) its log10
        weight ("log10w") and the chain's weight after it ("chain")."""
        n = num_chains * chain_length
        out = Partials()
        mn = C.c_double()
        pp = np.empty(n, np.float32) if per_path else None
        ch = np.empty(n, np.float32) if per_path else None
        pj = np.empty(n * (self.N - 1), np.float32) if per_joint else None
        self._check(lib().twisty_hip_run_generated_chains(self._h, num_chains, chain_length, chain_offset, seed, min_ds,
                                                          max_ds, extinction, norm_log10, C.byref(out), C.byref(mn),
                                                          _ptr(pp), _ptr(ch), _ptr(pj)))
        res = out.as_dict()
        res["min_log10_w"] = mn.value
        if per_path:
            res["log10w"] = pp
            res["chain"] = ch
        if per_joint:
            res["curv"] = pj.reshape(n, self.N - 1)
        return res

    def last_timing(self):
        a, b = C.c_float(), C.c_float()
        self._check(lib().twisty_hip_last_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_weight_histogram(self, num_bins=500, min_weight=4.69653e-16, max_weight=1.17517e-13, norm_log10=0.0):
        """Weight-magnitude histogram (defaults: the reference's 500 bins and range)."""
        self._check(lib().twisty_hip_set_weight_histogram(self._h, num_bins, min_weight, max_weight, norm_log10))
        self.weight_bins = num_bins

    def weight_histogram(self):
        h = np.empty(self.weight_bins, np.uint64)
        self._check(lib().twisty_hip_get_weight_histogram(self._h, _ptr(h)))
        return h

    def set_burn_in(self, paths_per_chain):
        """Unweighted paths at the start of every chain of the following runs (pathsToSkip)."""
        self._check(lib().twisty_hip_set_burn_in(self._h, paths_per_chain))

    def set_keep_final_state(self, enable=True):
        self._check(lib().twisty_hip_set_keep_final_state(self._h, int(enable)))

    def final_positions(self, num_paths):
        """(num_paths, N+1, 3) float32 final points of the last run."""
        out = np.empty((num_paths, self.N + 1, 3), np.float32)
        self._check(lib().twisty_hip_get_final_positions(self._h, _ptr(out), num_paths))
        return out

    def set_exit_histogram(self, n_theta, n_phi):
        self._check(lib().twisty_hip_set_exit_histogram(self._h, n_theta, n_phi))
        self.hist_shape = (n_theta, n_phi)

    def exit_histogram(self, num_paths=None):
        """(histogram (n_theta, n_phi) float64, per-path bins or None) of the last run."""
        nt, nph = self.hist_shape
        h = np.empty(nt * nph, np.float64)
        bins = np.empty(num_paths, np.uint32) if num_paths else None
        self._check(lib().twisty_hip_get_exit_histogram(self._h, _ptr(h), _ptr(bins), num_paths or 0))
        return h.reshape(nt, nph), bins

    def set_adjoint_schedule(self, lanes_per_path=0, block_size=0, blocks_per_cu=0):
        self._check(lib().twisty_hip_set_adjoint_schedule(self._h, lanes_per_path, block_size, blocks_per_cu))

    def run_adjoint(self, num_paths, K, seed=1, path_offset=0, segments=True, table_bins=None, kappa=None):
        """Forward + adjoint: partials, dE/dT (bins) and dE/d(initial segments) (N, 3) or None.
        kappa, when given, sets the roulette of the reverse sweep first (0 = exact)."""
        if kappa is not None:
            self.set_adjoint_roulette(kappa)
        out = Partials()
        nb = table_bins if table_bins is not None else self.table_bins
        dT = np.empty(nb, np.float64)
        dS = np.empty(3 * self.N, np.float64) if segments else None
        self._check(lib().twisty_hip_run_adjoint(self._h, num_paths, path_offset, K, seed, C.byref(out), _ptr(dT),
                                                 _ptr(dS)))
        res = out.as_dict()
        res["d_table"] = dT
        res["d_seg"] = None if dS is None else dS.reshape(self.N, 3)
        return res

    def set_adjoint_directions(self, directions):
        """Directions (D, N, 3) of the initial segments for per-result statistics
        of the reverse sweep's gradient (None or D = 0 switches it off)."""
        if directions is None:
            self._check(lib().twisty_hip_set_adjoint_directions(self._h, None, 0))
            return
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3 * self.N)
        self._check(lib().twisty_hip_set_adjoint_directions(self._h, _ptr(d), d.shape[0]))
        self._num_dirs = d.shape[0]

    def last_adjoint_direction_stats(self):
        """dE/dv, the roulette's added standard deviation and the Monte Carlo
        standard error of dE/dv, per direction v of set_adjoint_directions."""
        D = getattr(self, "_num_dirs", 0)
        g, r, m = (np.empty(D, np.float64) for _ in range(3))
        self._check(lib().twisty_hip_last_adjoint_direction_stats(self._h, _ptr(g), _ptr(r), _ptr(m)))
        return {"gradient": g, "roulette_std": r, "mc_std": m}

    def last_adjoint_ms(self):
        a = C.c_float()
        self._check(lib().twisty_hip_last_adjoint_timing(self._h, C.byref(a)))
        return a.value


class Batch:
    """A batch of experiments on one context (twisty_hip_batch_*).

    experiments: list of dicts with N, paths, K, seed, table, arclength,
    positions ((N+1, 3)) and optionally chain (paths per chain, default 1) and
    burn_in (default 0); tables: (numTables, bins); meta: (numTables, 5) =
    minK, maxK, step, ds, absorption.
    """

    def __init__(self, ctx, experiments, tables, meta):
        self.ctx = ctx
        n = len(experiments)
        arr = (Experiment * n)()
        pos = []
        off = 0
        self.paths = []
        for i, x in enumerate(experiments):
            p = np.ascontiguousarray(x["positions"], np.float32).reshape(-1, 3)
            assert p.shape[0] == x["N"] + 1
            arr[i] = Experiment(x["paths"], x.get("seed", 1), x["N"], x["K"], x["table"], off, x["arclength"],
                                x.get("chain", 1), x.get("burn_in", 0))
            pos.append(p)
            off += p.shape[0]
            self.paths.append(x["paths"])
        self.positions = np.ascontiguousarray(np.concatenate(pos), np.float32)
        t = np.ascontiguousarray(tables, np.float32)
        m = np.ascontiguousarray(meta, np.float32)
        self.n = n
        self._h = C.c_void_p()
        if not hasattr(ctx, "_batches"):
            ctx._batches = weakref.WeakSet()
        ctx._batches.add(self)
        ctx._check(lib().twisty_hip_batch_create(ctx._h, C.cast(arr, C.c_void_p), n, _ptr(t), t.shape[1], t.shape[0],
                                                 _ptr(m), _ptr(self.positions), self.positions.shape[0],
                                                 C.byref(self._h)))

    def close(self):
        if self._h:
            lib().twisty_hip_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, block_size=0, blocks_per_cu=0, segments_per_lane=0, tile_paths=0, lean=0):
        """Launch schedule (twisty_hip_batch_schedule): lean 0 keeps the tables
        and draw records in LDS, 1 reads the tables from global memory and
        exchanges the records with ds_bpermute, 2 reads the tables from global
        memory and keeps the records in LDS (block 256).  Raises and keeps the
        previous schedule when the combination is not compiled."""
        s = BatchSchedule(block_size, blocks_per_cu, segments_per_lane, tile_paths, lean)
        self.ctx._check(lib().twisty_hip_batch_set_schedule(self._h, C.byref(s)))

    def get_schedule(self):
        s = BatchSchedule()
        self.ctx._check(lib().twisty_hip_batch_get_schedule(self._h, C.byref(s)))
        return s.as_dict()

    def run(self, per_path=False):
        """(list of partial dicts, per-path log10 list or None, device ms)."""
        out = (Partials * self.n)()
        pp = np.empty(sum(self.paths), np.float32) if per_path else None
        ms = C.c_float()
        self.ctx._check(lib().twisty_hip_batch_run(self._h, C.cast(out, C.c_void_p), _ptr(pp), C.byref(ms)))
        res = [out[i].as_dict() for i in range(self.n)]
        split = None
        if per_path:
            split = np.split(pp, np.cumsum(self.paths)[:-1])
        return res, split, ms.value

    def autotune(self, reps=2, cap=64):
        best = BatchSchedule()
        tried = (BatchSchedule * cap)()
        ms = np.zeros(cap, np.float32)
        n = C.c_uint32()
        self.ctx._check(lib().twisty_hip_batch_autotune(self._h, reps, C.byref(best), C.cast(tried, C.c_void_p),
                                                        _ptr(ms), cap, C.byref(n)))
        return best.as_dict(), [(tried[i].as_dict(), float(ms[i])) for i in range(min(n.value, cap))]


def estimator(res, norm=1.0):
    """Average path weight and its Monte Carlo standard error, both as log10.

    Mirrors the driver's "Avg path weight" = experimentWeight / paths
    (Experiments/FullExperiment.cpp:244): invalid paths count with weight 0.
    """
    n = res["num_paths"]
    l1, l2 = res["log10_sum_w"], res["log10_sum_w2"]
    mean_l = l1 - math.log10(n) + math.log10(norm)
    # var = E[W^2] - E[W]^2, both scaled by 10^{-2 l1} to stay finite
    e2 = 10.0 ** (l2 - 2 * l1) * n  # E[W^2]/E[W]^2
    rel_var = max(e2 - 1.0, 0.0) / max(n - 1, 1)
    return mean_l, math.sqrt(rel_var)  # relative standard error of the mean
