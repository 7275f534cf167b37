"""Seeded generators of the batched workload (many heterogeneous experiments).

The batch: start (0,0,0) with tangent +z; end direction e uniform on the unit
sphere, end point at distance D = 1 along e with end tangent e (the recipe of
Experiments/BeamSpreadExperiment.cpp:186-197); arclength L = D * U[1.05, 3.0];
N log-uniform in [16, 256]; paths log-uniform in [2^10, 2^16]; K
perturbations per path.  A draw whose arclength is below the minimum for its
boundary conditions (PathGeneration::CalculateMinimumArclength) or admits no
initial arc is rejected and drawn again.  The experiments share a few
radiative-transfer tables with Gaussian widths sigma ~ U[0.03, 0.2]; each
experiment takes one at random.  The tables are built on the curvature grid of
the batch's smallest ds, so every experiment's curvatures fall inside.
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
        x["table"] = int(rng.integers(n_tables))
    ds_min = min(float(np.float32(x["arclength"]) / np.float32(x["N"])) for x in exps)
    tables, meta = [], []
    for sg in sigmas:
        t = host.build_rt_table(ds_min, mu, float(np.float32(sg) ** 2), scatter, absorption, numStepsInt=numStepsInt,
                                numCurvatureSteps=bins)
        tables.append(t.table)
        meta.append((t.minK, t.maxK, t.step, t.ds, absorption))
    return {"experiments": exps, "tables": np.stack(tables), "meta": np.array(meta, np.float32),
            "sigmas": sigmas, "ds_table": ds_min}
