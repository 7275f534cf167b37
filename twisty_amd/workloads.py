"""Seeded generators of the batched workload (many heterogeneous experiments).

The batch: start (0,0,0) with tangent +z; end direction e uniform on the unit
sphere, end point at distance D = 1 along e with end tangent e (the recipe of
Experiments/BeamSpreadExperiment.cpp:186-197); arclength L = D * U[1.05, 3.0];
N log-uniform in [16, 256]; paths log-uniform in [2^10, 2^16]; K
perturbations per path.  A draw whose arclength is below the minimum for its
boundary conditions (PathGeneration::CalculateMinimumArclength) or admits no
initial arc is rejected and drawn again.  The experiments share a few
radiative-transfer weighting parameter sets, Gaussian widths sigma ~ U[0.03, 0.2];
each experiment takes one at random and, as the reference's runner does
(FullExperimentRunnerOptimalPerturb.cpp:55-65), weighs its paths with the table
of that sigma built at its own ds = L / N (experiments with the same sigma and
the same float ds share one table).
"""
import math

import numpy as np

from . import host


def batch_experiments(n_exp=4096, seed=1, K=100, n_range=(16, 256), paths_range=(1 << 10, 1 << 16),
                      ratio_range=(1.05, 3.0), n_tables=8, sigma_range=(0.03, 0.2), bins=4096, numStepsInt=20000,
                      mu=1.0, scatter=0.1, absorption=0.004):
    rng = np.random.default_rng(seed)
    exps = []
    while len(exps) < n_exp:
        N = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        paths = int(round(math.exp(rng.uniform(math.log(paths_range[0]), math.log(paths_range[1])))))
        n1 = rng.uniform() * 2.0 - 1.0
        e1 = rng.uniform()
        s = math.sqrt(max(0.0, 1.0 - n1 * n1))
        e = (n1, math.cos(2 * math.pi * e1) * s, math.sin(2 * math.pi * e1) * s)
        L = float(rng.uniform(*ratio_range))
        bc = host.boundary13((0, 0, 0), (0, 0, 1), e, e, L)
        if not L > host.minimum_arclength(bc, N) * 1.0001:
            continue
        try:
            pos = host.initial_arc(bc, N)
        except ValueError:
            continue
        exps.append({"N": N, "paths": paths, "K": K, "seed": seed + len(exps), "arclength": L, "boundary": bc,
                     "positions": pos, "end_dir": e})
    sigmas = rng.uniform(*sigma_range, size=n_tables)
    for x in exps:
        x["sigma_index"] = int(rng.integers(n_tables))
    # one table per distinct (sigma, float ds)
    key_of, keys = {}, []
    for x in exps:
        ds = np.float32(x["arclength"]) / np.float32(x["N"])
        k = (x["sigma_index"], int(np.float32(ds).view(np.uint32)))
        if k not in key_of:
            key_of[k] = len(keys)
            keys.append((x["sigma_index"], float(ds)))
        x["table"] = key_of[k]
    eps = np.array([np.float32(sigmas[si]) ** 2 for si, _ in keys], np.float32)
    dss = np.array([d for _, d in keys], np.float32)
    tables, m6 = host.build_rt_tables(dss, eps, mu, scatter, absorption, numStepsInt=numStepsInt,
                                      numCurvatureSteps=bins)
    # Batch table meta: minK, maxK, step, ds, absorption
    meta = np.stack([m6[:, 1], m6[:, 2], m6[:, 3], m6[:, 0], np.full(len(keys), absorption, np.float32)], axis=1)
    return {"experiments": exps, "tables": tables, "meta": meta.astype(np.float32), "sigmas": sigmas,
            "table_sigma_index": np.array([si for si, _ in keys], np.int32)}
