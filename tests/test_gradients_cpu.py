"""CPU checks behind the adjoint: the FP64 oracle adjoint, the table's eps
derivative and the initial curve's exit-direction Jacobian, each against
central finite differences of an independent computation.

The oracle adjoint (oracle/oracle_adjoint.cpp) is the reference the GPU adjoint
is compared with.  It is pinned here by central differences of the oracle's own
estimator: exactly in lerp mode 1 (weights continuous in the curvature), and for
the table entries also in the reference's lerp (mode 0), where the estimator is
linear in each entry.
"""
import math

import numpy as np
import pytest

from twisty_amd import host


def small_case(oracle, N=20, L=1.2, sigma=0.1, bins=1024, steps=2000):
    ds = float(np.float32(L) / np.float32(N))
    eps = float(np.float32(sigma) ** 2)
    t = oracle.build_table(ds, eps=eps, numCurvatureSteps=bins, numStepsInt=steps)
    pos, _ = oracle.initial_arc(N, L)
    return t, pos


def test_oracle_adjoint_forward_matches_simulate(oracle):
    """Mode 0 of the adjoint's forward is the oracle path (segment form)."""
    t, pos = small_case(oracle)
    a = oracle.adjoint(pos, 20, 1.2, 20, 2048, t, 0.004, lerp_mode=0, want_table=False, want_seg=False, want_w=True)
    s = oracle.simulate(pos, 20, 1.2, 20, 2048, t, 0.004)
    assert abs(a["log10_sum_w"] - s["log10_sum_w_exact"]) < 1e-10
    assert np.abs(a["log10w"] - s["log10w"]).max() < 1e-9


@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_adjoint_identities(oracle, mode):
    """Table homogeneity: sum_b T[b] dE/dT[b] = (N-1)/ln 10.  Invariance of E
    under a rotation or uniform scaling of the initial segments:
    sum_i d_i x dE/dd_i = 0, sum_i d_i . dE/dd_i = 0."""
    t, pos = small_case(oracle)
    N = 20
    r = oracle.adjoint(pos, N, 1.2, 20, 4096, t, 0.004, lerp_mode=mode)
    # (mode 0 lerps in float: each joint's value carries one float rounding)
    tol = 1e-10 if mode == 1 else 1e-6
    assert abs(t.table.astype(np.float64) @ r["d_table"] - (N - 1) / math.log(10)) < tol * (N - 1)
    d = np.diff(pos, axis=0)
    lam = r["d_seg"]
    scale = np.linalg.norm(d, axis=1) @ np.linalg.norm(lam, axis=1)
    assert np.abs(np.cross(d, lam).sum(axis=0)).max() < 1e-10 * scale
    assert abs(np.sum(d * lam)) < 1e-10 * scale


@pytest.mark.parametrize("mode,h,tol", [(1, 1e-4, 1e-7), (0, 1e-3, 2e-4)])
def test_oracle_table_gradient_fd(oracle, mode, h, tol):
    t, pos = small_case(oracle)
    N, L, K, P = 20, 1.2, 20, 4096
    r = oracle.adjoint(pos, N, L, K, P, t, 0.004, lerp_mode=mode)
    rng = np.random.default_rng(7)
    v = rng.standard_normal(len(t.table)) * t.table
    tp = (t.table + h * v).astype(np.float32)
    tm = (t.table - h * v).astype(np.float32)
    ep = oracle.adjoint(pos, N, L, K, P, t, 0.004, lerp_mode=mode, want_table=False, want_seg=False,
                        table_values=tp)["log10_sum_w"]
    em = oracle.adjoint(pos, N, L, K, P, t, 0.004, lerp_mode=mode, want_table=False, want_seg=False,
                        table_values=tm)["log10_sum_w"]
    an = r["d_table"] @ (tp.astype(np.float64) - tm.astype(np.float64))
    assert abs((ep - em) - an) <= tol * abs(an)


@pytest.mark.parametrize("K", [1, 20])
def test_oracle_segment_gradient_fd_continuous_lerp(oracle, K):
    """Mode 1: dE/d(initial segments) against central differences in the positions."""
    t, pos = small_case(oracle)
    N, L, P = 20, 1.2, 4096
    r = oracle.adjoint(pos, N, L, K, P, t, 0.004, lerp_mode=1)
    rng = np.random.default_rng(K)
    dp = rng.standard_normal(pos.shape)
    an = float(np.sum(r["d_seg"] * np.diff(dp, axis=0)))
    # Piecewise-linear weights: a kink crossed inside [-h, h] leaves an error of
    # order h; with K = 1 most joints of all paths share the arc's curvature,
    # so h is kept below the distance of those curvatures to a bin edge.
    h = 2e-6
    ep = oracle.adjoint(pos + h * dp, N, L, K, P, t, 0.004, lerp_mode=1, want_table=False, want_seg=False)
    em = oracle.adjoint(pos - h * dp, N, L, K, P, t, 0.004, lerp_mode=1, want_table=False, want_seg=False)
    fd = (ep["log10_sum_w"] - em["log10_sum_w"]) / (2 * h)
    assert abs(fd - an) <= 2e-5 * abs(an)


def test_oracle_segment_gradient_reference_lerp_statistics(oracle):
    """Mode 0: the bin-average slope matches the mean finite difference.

    The reference lerp jumps at every bin edge, so a finite difference of a
    fixed-seed estimate mixes the in-bin slope with the edges crossed.  Over
    many step sizes the differences scatter around the adjoint; their mean
    agrees within 4 standard errors.
    """
    t, pos = small_case(oracle)
    N, L, K, P = 20, 1.2, 20, 4096
    r = oracle.adjoint(pos, N, L, K, P, t, 0.004, lerp_mode=0)
    rng = np.random.default_rng(3)
    dp = rng.standard_normal(pos.shape)
    an = float(np.sum(r["d_seg"] * np.diff(dp, axis=0)))
    fds = []
    for h in np.geomspace(1e-5, 3e-4, 16):
        ep = oracle.adjoint(pos + h * dp, N, L, K, P, t, 0.004, lerp_mode=0, want_table=False, want_seg=False)
        em = oracle.adjoint(pos - h * dp, N, L, K, P, t, 0.004, lerp_mode=0, want_table=False, want_seg=False)
        fds.append((ep["log10_sum_w"] - em["log10_sum_w"]) / (2 * h))
    fds = np.array(fds)
    se = fds.std(ddof=1) / math.sqrt(len(fds))
    assert abs(fds.mean() - an) <= 4 * se + 1e-3 * abs(an)


def _rt_table_f64(ds, mu, eps, scatter, kappas, n=2000, bias=1.0):
    """FP64 restatement of the table integral (double phase, same node set)."""
    m, c = eps, mu
    h = (10.0 / eps) / n
    bds = np.float64(np.float32(scatter * ds))
    A = math.sqrt(math.pi * m * 0.5) / (1.0 - math.exp(-2.0 / m))
    idx = np.concatenate([[0, n], np.arange(1, n // 2) * 2 - 1, np.arange(1, n // 2 - 1) * 2])
    coef = np.concatenate([[1 / 3, 1 / 3], np.full(n // 2 - 1, 4 / 3), np.full(n // 2 - 2, 2 / 3)]) * h
    p = idx * h
    f = p * np.exp(bds * A * np.exp(-m * p * p * 0.5) - c * c * p * p * 0.5)
    out = []
    for k in kappas:
        kds = np.float64(np.float32(np.float32(k) * np.float32(ds)))
        s = np.sin(kds * p) / kds if kds > 0 else p
        out.append(np.sum(coef * f * s) * bias / (2 * math.pi ** 2))
    return np.array(out)


@pytest.mark.parametrize("N,L,sigma,mu", [(100, 1.3, 0.08, 1.0), (64, 1.5, 0.05, 1.0), (20, 1.2, 0.1, 0.5)])
def test_table_eps_derivative_fd(oracle, N, L, sigma, mu):
    ds = float(np.float32(L) / np.float32(N))
    eps = float(np.float32(sigma) ** 2)
    bins = 1024
    d = host.rt_table_eps_derivative(ds, mu, eps, 0.1, numStepsInt=2000, numCurvatureSteps=bins)
    tab = oracle.build_table(ds, mu=mu, eps=eps, numStepsInt=2000, numCurvatureSteps=bins)
    sel = np.array([0, 3, 40, 200, 600, 1000])
    kap = np.where(sel == 0, 0.0, sel.astype(np.float32) * np.float32(tab.step))
    # the FP64 restatement reproduces the oracle's (float-phase) table values
    # (entries <= 0 are replaced by the table minimum, PathWeightUtils.cpp:339-344)
    v = _rt_table_f64(ds, mu, eps, 0.1, kap)
    keep = v > 0
    assert keep.sum() >= 4
    assert np.allclose(v[keep], tab.table[sel][keep], rtol=2e-6, atol=0)
    assert np.all(tab.table[sel][~keep] == np.float32(tab.minW))
    h = 1e-5 * eps
    fd = (_rt_table_f64(ds, mu, eps + h, 0.1, kap) - _rt_table_f64(ds, mu, eps - h, 0.1, kap)) / (2 * h)
    scale = np.abs(fd[keep]).max()
    assert np.allclose(d[sel][keep], fd[keep], rtol=1e-5, atol=1e-7 * scale)


@pytest.mark.parametrize("edir", [(0.0, 0.0, 1.0), (0.6, 0.0, 0.8), (0.3, -0.4, math.sqrt(1 - 0.25))])
def test_initial_arc_exit_jacobian(oracle, edir):
    N, L = 30, 2.2
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), edir, L)
    jt, jp = host.initial_arc_exit_jacobian(bc, N)
    th = math.acos(edir[2])
    ph = math.atan2(edir[1], edir[0])
    h = 1e-6

    # startDir on the chord (edir = +z): the bulge plane is the least-aligned
    # axis at the base point and is held fixed for the derivative.
    bulge = (1.0, 0.0, 0.0) if edir == (0.0, 0.0, 1.0) else None

    def arc(t_, p_):
        e = (math.sin(t_) * math.cos(p_), math.sin(t_) * math.sin(p_), math.cos(t_))
        if bulge is not None:
            return oracle.initial_arc_bulge(N, L, oracle.bc12(edir=e), bulge)
        pos, _ = oracle.initial_arc(N, L, oracle.bc12(edir=e))
        return pos

    assert np.abs(arc(th, ph) - host.initial_arc(bc, N)).max() < 4e-7
    fdt = (arc(th + h, ph) - arc(th - h, ph)) / (2 * h)
    fdp = (arc(th, ph + h) - arc(th, ph - h)) / (2 * h)
    assert np.abs(jt - fdt).max() < 1e-6 * max(1.0, np.abs(fdt).max())
    assert np.abs(jp - fdp).max() < 1e-6 * max(1.0, np.abs(fdp).max())
    assert np.abs(jt[:2]).max() == 0.0 and np.abs(jt[-1]).max() == 0.0


def test_exit_gradient_chain_continuous_lerp(oracle):
    """dE/dtheta = sum_i dE/dd_i . dd_i/dtheta against central differences of
    the oracle estimator over initial arcs rebuilt at theta +- h (mode 1)."""
    N, L, K, P = 24, 1.6, 10, 4096
    edir = (0.6, 0.0, 0.8)
    ds = float(np.float32(L) / np.float32(N))
    t = oracle.build_table(ds, eps=0.01, numCurvatureSteps=1024, numStepsInt=2000)
    bc = host.boundary13((0, 0, 0), (0, 0, 1), (0, 0, 1), edir, L)
    jt, jp = host.initial_arc_exit_jacobian(bc, N)
    th, ph = math.acos(0.8), 0.0

    def arc(t_, p_):
        e = (math.sin(t_) * math.cos(p_), math.sin(t_) * math.sin(p_), math.cos(t_))
        pos, _ = oracle.initial_arc(N, L, oracle.bc12(edir=e))
        return pos

    r = oracle.adjoint(arc(th, ph), N, L, K, P, t, 0.004, lerp_mode=1)
    for jac, (dth, dph) in ((jt, (1, 0)), (jp, (0, 1))):
        an = float(np.sum(r["d_seg"] * np.diff(jac, axis=0)))
        h = 2e-6
        ep = oracle.adjoint(arc(th + dth * h, ph + dph * h), N, L, K, P, t, 0.004, lerp_mode=1, want_table=False,
                            want_seg=False)["log10_sum_w"]
        em = oracle.adjoint(arc(th - dth * h, ph - dph * h), N, L, K, P, t, 0.004, lerp_mode=1, want_table=False,
                            want_seg=False)["log10_sum_w"]
        fd = (ep - em) / (2 * h)
        assert abs(fd - an) <= 2e-4 * abs(an) + 1e-9
