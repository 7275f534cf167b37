"""GPU adjoint (twisty_hip_run_adjoint) against the FP64 oracle.

The oracle adjoint (oracle/oracle_adjoint.cpp, pinned by central differences in
test_gradients_cpu) computes the same derivative in FP64 with the reference's
lookup: dE/dT exactly, dE/d(initial segments) with the bin-average slope.
Tolerances, FP32 kernel vs FP64 oracle on the same draws:
  * table gradient: ||g - r|| <= 1e-2 ||r|| and, for the running sum over
    entries, max |G - R| <= 1e-3 sum |r|.  Per-path weights agree to 1e-5 in
    log10 (about 1e-3 relative in weight at log10 w ~ -40), and a joint whose
    FP32 curvature falls into the neighbouring bin moves its share to the next
    entry, which the running sum does not see;
  * exact identities (check_invariants): table homogeneity and the rotation /
    scale invariance of the initial curve;
  * segment gradient: ||g - r|| <= 5e-3 ||r|| (the same, plus the FP32 rewind
    of K rotations);
  * the table gradient also against central differences of the oracle's
    estimator in the reference lookup, where E is exactly a function of the
    table entries;
  * sigma and the exit angles through torch.autograd against the oracle.
"""
import math

import numpy as np
import pytest

from tests import cases
from twisty_amd import capi, host

pytestmark = pytest.mark.gpu


def rel_norm(g, r):
    return np.linalg.norm(g - r) / np.linalg.norm(r)


def oracle_adjoint(case, oracle, paths, table=None, pos=None, K=None, mode=0):
    t, pos32 = cases.setup(case, oracle)
    t = t if table is None else table
    pos = pos32.astype(np.float64) if pos is None else pos
    return oracle.adjoint(pos, case.N, case.L, case.K if K is None else K, paths, t, case.absorption,
                          seed=case.seed, lerp_mode=mode)


@pytest.mark.parametrize("case,paths", [(cases.C4, 1 << 14), (cases.C1, 1 << 14), (cases.MID, 1 << 13),
                                        (cases.SMALL, 1 << 12)])
def test_adjoint_matches_oracle(hip_device, oracle, case, paths):
    print(f"[adjoint {case.name} paths={paths}]", flush=True)
    ctx = cases.hip_context(case, oracle)
    g = ctx.run_adjoint(paths, case.K, seed=case.seed, kappa=0.0)  # exact sweep: same draws as the oracle
    r = oracle_adjoint(case, oracle, paths)
    e_t, e_s = rel_norm(g["d_table"], r["d_table"]), rel_norm(g["d_seg"], r["d_seg"])
    print(f"  |dT| rel {e_t:.2e}  |dSeg| rel {e_s:.2e}  fwd {ctx.last_timing()[0]:.3f} ms "
          f"adj {ctx.last_adjoint_ms():.3f} ms", flush=True)
    assert abs(g["log10_sum_w"] - r["log10_sum_w"]) < 1e-3 * abs(r["log10_sum_w"])
    # A joint whose FP32 curvature lands one bin over moves its share to the
    # neighbouring entry: entrywise this is a 1e-3-level difference, in the
    # running sum over entries it cancels.
    e_c = np.abs(np.cumsum(g["d_table"]) - np.cumsum(r["d_table"])).max() / np.abs(r["d_table"]).sum()
    print(f"  running-sum dT rel {e_c:.2e}", flush=True)
    assert e_t <= 1e-2
    assert e_c <= 1e-3
    assert e_s <= 5e-3
    check_invariants(g, case.N, cases.setup(case, oracle)[0].table, cases.setup(case, oracle)[1])
    ctx.close()


def check_invariants(g, N, table, pos):
    """Exact identities of E = log10(mean w):
    * the lerp is homogeneous in the table, so scaling T by lambda scales every
      path weight by lambda^(N-1): sum_b T[b] dE/dT[b] = (N-1)/ln 10;
    * weights depend on tangents and on N/L only, and the rotation axes are
      directions, so E is unchanged by a rotation or a uniform scaling of the
      initial segments: sum_i d_i x dE/dd_i = 0 and sum_i d_i . dE/dd_i = 0."""
    s = float(np.dot(table.astype(np.float64), g["d_table"]))
    assert abs(s - (N - 1) / math.log(10)) <= 1e-4 * (N - 1) / math.log(10)
    if g.get("d_seg") is not None:
        d = np.diff(pos.astype(np.float64), axis=0)
        lam = g["d_seg"]
        scale = np.linalg.norm(d, axis=1) @ np.linalg.norm(lam, axis=1)
        assert np.abs(np.cross(d, lam).sum(axis=0)).max() <= 1e-4 * scale
        assert abs(np.sum(d * lam)) <= 1e-4 * scale


@pytest.mark.parametrize("h", [1e-4, 1e-3])
def test_adjoint_table_gradient_vs_oracle_fd(hip_device, oracle, h):
    """dE/dT along a random direction v (entries scaled by the table) against
    FP64 central differences of the oracle's estimator at T +- h v, h relative
    (1e-4 is the differentiation workload's step).  The float tables are
    rounded after the step, so the adjoint is applied to the rounded step."""
    case, paths = cases.C4, 1 << 13
    ctx = cases.hip_context(case, oracle)
    g = ctx.run_adjoint(paths, case.K, seed=case.seed, segments=False)
    table, pos32 = cases.setup(case, oracle)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(len(table.table)) * table.table
    tp = (table.table + h * v).astype(np.float32)
    tm = (table.table - h * v).astype(np.float32)

    def est(tv):
        return oracle.adjoint(pos32.astype(np.float64), case.N, case.L, case.K, paths, table, case.absorption,
                              seed=case.seed, lerp_mode=0, want_table=False, want_seg=False,
                              table_values=tv)["log10_sum_w"]

    fd = est(tp) - est(tm)
    an = g["d_table"] @ (tp.astype(np.float64) - tm.astype(np.float64))
    print(f"[adjoint FD h={h:g}] fd={fd:.6e} adjoint={an:.6e}", flush=True)
    assert abs(fd - an) <= 3e-3 * abs(fd)
    ctx.close()


@pytest.mark.parametrize("sched", [(4, 256, 0), (8, 512, 0), (16, 512, 0), (64, 512, 0)])
def test_adjoint_schedules_agree(hip_device, oracle, sched):
    case, paths = cases.C2, 1 << 13
    ctx = cases.hip_context(case, oracle)
    base = ctx.run_adjoint(paths, case.K, seed=case.seed)
    ctx.set_adjoint_schedule(*sched)
    other = ctx.run_adjoint(paths, case.K, seed=case.seed)
    print(f"[adjoint schedule {sched}] dT {rel_norm(other['d_table'], base['d_table']):.2e} "
          f"dSeg {rel_norm(other['d_seg'], base['d_seg']):.2e}", flush=True)
    assert rel_norm(other["d_table"], base["d_table"]) < 1e-5
    assert rel_norm(other["d_seg"], base["d_seg"]) < 1e-4
    no_sweep = ctx.run_adjoint(paths, case.K, seed=case.seed, segments=False)
    assert rel_norm(no_sweep["d_table"], base["d_table"]) < 1e-5
    ctx.close()


def test_adjoint_large_n(hip_device, oracle):
    case, paths = cases.C5, 1 << 10
    print(f"[adjoint {case.name} paths={paths}]", flush=True)
    ctx = cases.hip_context(case, oracle)
    g = ctx.run_adjoint(paths, case.K, seed=case.seed, kappa=0.0)
    r = oracle_adjoint(case, oracle, paths)
    e_t, e_s = rel_norm(g["d_table"], r["d_table"]), rel_norm(g["d_seg"], r["d_seg"])
    print(f"  |dT| rel {e_t:.2e}  |dSeg| rel {e_s:.2e}", flush=True)
    # log10 w ~ -600 here: 1e-5 relative per path is ~1% in weight
    assert e_t <= 5e-2 and e_s <= 5e-2
    ctx.close()


def test_autograd_sigma_and_exit_angles(hip_device, oracle):
    """C4 geometry: d E / d sigma and d E / d(theta, phi) of the exit direction."""
    import torch

    from twisty_amd import autograd as ag

    case, paths = cases.C4, 1 << 13
    N, L, K = case.N, case.L, case.K
    ds = cases.ds_of(case)
    for edir, bulge in (((0.0, 0.0, 1.0), (1.0, 0.0, 0.0)), ((0.6, 0.0, 0.8), None)):
        bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), edir, L)
        th0, ph0 = math.acos(edir[2]), math.atan2(edir[1], edir[0])
        sigma = torch.tensor(case.sigma, dtype=torch.float64, requires_grad=True)
        ang = torch.tensor([th0, ph0], dtype=torch.float64, requires_grad=True)
        hctx = capi.Context(0)
        hctx.set_adjoint_roulette(0.0)  # exact sweep: the same draws as the oracle
        table, grid = ag.rt_table_with_grid(sigma, ds, numStepsInt=case.numStepsInt, bins=case.bins)
        E = ag.log10_estimate(table, ag.exit_curve(ang, bc, N), hctx, grid, bc, paths, K, seed=case.seed)
        E.backward()
        hctx.close()

        def arc(t_, p_):
            e = (math.sin(t_) * math.cos(p_), math.sin(t_) * math.sin(p_), math.cos(t_))
            if bulge is not None:
                return oracle.initial_arc_bulge(N, L, oracle.bc12(edir=e), bulge)
            return oracle.initial_arc(N, L, oracle.bc12(edir=e))[0]

        def otable(s):
            return oracle.build_table(ds, eps=float(np.float32(s) ** 2), numCurvatureSteps=case.bins,
                                      numStepsInt=case.numStepsInt)

        pos0 = host.initial_arc(bc, N).astype(np.float64)
        # sigma: (1) the chain d_table . dT/deps . 2 sigma with the oracle's FP64
        # table gradient; (2) a fourth-order central difference of the oracle
        # estimator over tables rebuilt at sigma +- h, 2h.  The tables are
        # float, so each rebuild carries rounding noise; h = 2% of sigma keeps
        # it below 1e-3 of the difference while the O(h^4) error stays ~1e-5.
        hs = 2e-2 * case.sigma

        def est_sigma(s):
            return oracle.adjoint(pos0, N, L, K, paths, otable(s), case.absorption, seed=case.seed, lerp_mode=0,
                                  want_table=False, want_seg=False)["log10_sum_w"]

        f = {k: est_sigma(case.sigma + k * hs) for k in (-2, -1, 1, 2)}
        fd_sigma = (f[-2] - 8 * f[-1] + 8 * f[1] - f[2]) / (12 * hs)
        # exit angles: the oracle adjoint's dE/d(segments) chained with the FP64
        # Jacobian of the oracle's own arc (central differences)
        r = oracle.adjoint(pos0, N, L, K, paths, otable(case.sigma), case.absorption, seed=case.seed, lerp_mode=0)
        h = 1e-6
        jt = (arc(th0 + h, ph0) - arc(th0 - h, ph0)) / (2 * h)
        jp = (arc(th0, ph0 + h) - arc(th0, ph0 - h)) / (2 * h)
        ref_ang = np.array([np.sum(r["d_seg"] * np.diff(jt, axis=0)), np.sum(r["d_seg"] * np.diff(jp, axis=0))])
        gs, ga = sigma.grad.item(), ang.grad.numpy()
        print(f"[autograd edir={edir}] dE/dsigma gpu={gs:.6e} fd={fd_sigma:.6e}; dE/dangles gpu={ga} ref={ref_ang}",
              flush=True)
        eps = float(np.float32(case.sigma) ** 2)
        dTde = host.rt_table_eps_derivative(ds, case.mu, eps, case.scatter, numStepsInt=case.numStepsInt,
                                            numCurvatureSteps=case.bins)
        chain = float(r["d_table"] @ dTde) * 2 * case.sigma
        print(f"  dE/dsigma oracle chain={chain:.6e}", flush=True)
        assert abs(gs - chain) <= 2e-3 * abs(chain)
        assert abs(gs - fd_sigma) <= 3e-3 * abs(fd_sigma)
        scale = np.abs(ref_ang).max()
        assert np.all(np.abs(ga - ref_ang) <= 5e-3 * scale + 1e-9)


def test_adjoint_roulette_against_exact(hip_device, oracle):
    """The default reverse sweep (roulette, kappa = 1): about a tenth of the
    paths are swept, the table gradient stays exact (every path), and on C4
    the segment gradient stays within the FP32-vs-FP64 tolerances of the exact
    sweep.  Along three fixed directions of the initial segments the reported
    roulette standard deviation covers the difference to the exact sweep, is
    zero for the exact sweep and far below the gradient's Monte Carlo error."""
    case = cases.C4
    paths = 1 << 20
    rng = np.random.default_rng(3)
    with cases.hip_context(case, oracle) as ctx:
        dirs = rng.standard_normal((3, case.N, 3))
        ctx.set_adjoint_directions(dirs)
        ex = ctx.run_adjoint(paths, case.K, seed=case.seed, kappa=0.0)
        assert ctx.last_adjoint_paths() == paths
        sx = ctx.last_adjoint_direction_stats()
        r = ctx.run_adjoint(paths, case.K, seed=case.seed, kappa=1.0)
        visited = ctx.last_adjoint_paths()
        sr = ctx.last_adjoint_direction_stats()
        again = ctx.run_adjoint(paths, case.K, seed=case.seed)  # the default is kappa = 1
        ctx.set_adjoint_directions(None)
    e_seg = rel_norm(r["d_seg"], ex["d_seg"])
    e_tab = rel_norm(r["d_table"], ex["d_table"])
    print(f"[roulette C4] visited {visited}/{paths}, dSeg rel {e_seg:.2e}, dT rel {e_tab:.2e}", flush=True)
    for k in range(3):
        print(f"  dir {k}: exact {sx['gradient'][k]:.6e} (mc std {sx['mc_std'][k]:.3e}), roulette "
              f"{sr['gradient'][k]:.6e} (roulette std {sr['roulette_std'][k]:.3e}, mc std {sr['mc_std'][k]:.3e})",
              flush=True)
    assert 0 < visited < paths // 3
    assert e_seg <= 5e-3 and e_tab <= 1e-5
    for st, g in ((sx, ex), (sr, r)):
        proj = np.array([np.sum(g["d_seg"] * d) for d in dirs])
        assert np.allclose(st["gradient"], proj, rtol=1e-6, atol=1e-12 * np.abs(proj).max())
    assert np.all(sx["roulette_std"] == 0.0) and np.all(sx["mc_std"] > 0.0)
    assert np.all(sr["roulette_std"] > 0.0) and np.all(sr["roulette_std"] <= 0.2 * sx["mc_std"])
    assert np.all(np.abs(sr["gradient"] - sx["gradient"]) <= 5.0 * sr["roulette_std"] + 1e-6 * sx["mc_std"])
    # the MC error estimated from the swept sample agrees with the exact sweep's
    assert np.all(np.abs(sr["mc_std"] / sx["mc_std"] - 1.0) <= 0.25)
    # same seed, same kept set: the default run agrees up to float atomics order
    assert rel_norm(again["d_seg"], r["d_seg"]) < 1e-5


def test_host_experiment_gradients_match_autograd(hip_device, oracle):
    """The C++ chain (twisty_host_experiment_gradients) equals the torch.autograd
    composition of the same pieces: dE/d eps * 2 sigma = dE/d sigma, and the
    exit angles, for e = +z and a tilted exit."""
    import torch

    from twisty_amd import autograd as ag

    case, paths = cases.C4, 1 << 13
    N, L, K = case.N, case.L, case.K
    ds = cases.ds_of(case)
    for edir in ((0.0, 0.0, 1.0), (0.6, 0.0, 0.8)):
        bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), edir, L)
        th0, ph0 = math.acos(edir[2]), math.atan2(edir[1], edir[0])
        sigma = torch.tensor(case.sigma, dtype=torch.float64, requires_grad=True)
        ang = torch.tensor([th0, ph0], dtype=torch.float64, requires_grad=True)
        hctx = capi.Context(0)
        table, grid = ag.rt_table_with_grid(sigma, ds, numStepsInt=case.numStepsInt, bins=case.bins)
        E = ag.log10_estimate(table, ag.exit_curve(ang, bc, N), hctx, grid, bc, paths, K, seed=case.seed)
        E.backward()
        hctx.close()
        cfg = {"experiment": {
            "experimentParams": {
                "pathsToGenerate": paths, "pathsToSkip": 0, "numSegments": N, "name": "grad", "experimentDir": "/tmp",
                "maxPerturbThreads": 0, "maxWeightThreads": 0, "outputBigFloatWeights": False,
                "outputPathBatches": False, "perturbMethod": 0, "useGpu": True,
                "random": {"bootstrapSeed": 1, "perturbSeed": case.seed},
                "weighting": {"weightFunction": 0, "mu": 1.0, "eps": float(np.float32(case.sigma) ** 2),
                              "numStepsInt": case.numStepsInt, "numCurvatureSteps": case.bins,
                              "absorption": case.absorption, "scatter": 0.1},
                "twisty_hip": {"perturbationsPerPath": K, "chainLength": 1, "writeFiles": False}},
            "basicExperiment": {"startPos": [0, 0, 0], "startDir": [0, 0, 1], "endPos": [0, 0, 1],
                                "endDir": list(edir), "arclength": L}}}
        g = host.experiment_gradients(cfg, table_bins=case.bins)
        d_sigma = g["d_eps"] * 2.0 * float(np.float32(case.sigma))
        print(f"[host gradients e={edir}] E {g['log10_estimate']:.9f} vs {float(E.detach()):.9f}; dsigma {d_sigma:.6e} vs "
              f"{float(sigma.grad):.6e}; dtheta {g['d_theta']:.6e} vs {float(ang.grad[0]):.6e}; dphi {g['d_phi']:.6e} "
              f"vs {float(ang.grad[1]):.6e}", flush=True)
        print(f"  exit-angle errors: roulette std ({g['d_theta_roulette_std']:.3e}, {g['d_phi_roulette_std']:.3e}), "
              f"MC std ({g['d_theta_mc_std']:.3e}, {g['d_phi_mc_std']:.3e}); swept {g['paths_swept']}/{paths}",
              flush=True)
        assert 0 < g["paths_swept"] < paths
        assert g["d_theta_mc_std"] > 0.0 and g["d_theta_roulette_std"] <= g["d_theta_mc_std"]
        assert abs(g["log10_estimate"] - float(E.detach())) < 1e-9 * abs(float(E.detach()))
        assert d_sigma == pytest.approx(float(sigma.grad), rel=1e-4)
        scale = max(abs(float(ang.grad[0])), abs(float(ang.grad[1])))
        assert abs(g["d_theta"] - float(ang.grad[0])) < 1e-4 * scale
        assert abs(g["d_phi"] - float(ang.grad[1])) < 1e-4 * scale
