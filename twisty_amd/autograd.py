"""torch.autograd face of the adjoint.

Three Functions compose into the gradient of the log-estimator with respect to
the physical parameters:

* ``rt_table(sigma, ...)`` -- the radiative-transfer table of the host builder
  for eps = sigma^2 (the config's eps slot, PathWeightUtils.cpp:401-420 after
  the reference's (mu, eps) swap); backward uses the FP64 d T / d eps.
* ``exit_curve(angles, boundary, N)`` -- the initial curve for the end direction
  (sin t cos p, sin t sin p, cos t); backward uses the FP64 Jacobian of the arc.
* ``log10_estimate(table, positions, ctx, ...)`` -- E = log10(mean path weight)
  of one run of the GPU path; backward is twisty_hip_run_adjoint's dE/dT and
  dE/d(initial segments), mapped onto the positions.

    sigma = torch.tensor(0.08, dtype=torch.float64, requires_grad=True)
    ang = torch.tensor([0.0, 0.0], dtype=torch.float64, requires_grad=True)
    table, grid = rt_table_with_grid(sigma, ds)
    E = log10_estimate(table, exit_curve(ang, bc, N), ctx, grid, bc, paths, K)
    E.backward()   # sigma.grad, ang.grad

All computation happens in the C++/HIP libraries; tensors are CPU tensors that
carry the values across.
"""
import math

import numpy as np
import torch

from . import host


class _RTTable(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma, ds, mu, scatter, absorption, numStepsInt, bins, grid_out):
        eps = float(np.float32(float(sigma)) ** 2)
        t = host.build_rt_table(ds, mu, eps, scatter, absorption, numStepsInt=numStepsInt, numCurvatureSteps=bins)
        grid_out.append((t.minK, t.maxK, t.step, t.ds, absorption))
        d = host.rt_table_eps_derivative(ds, mu, eps, scatter, numStepsInt=numStepsInt, numCurvatureSteps=bins)
        ctx.save_for_backward(torch.from_numpy(d * (2.0 * float(sigma))))
        return torch.from_numpy(t.table.astype(np.float64))

    @staticmethod
    def backward(ctx, g):
        (d,) = ctx.saved_tensors
        return (g * d).sum().reshape(()), None, None, None, None, None, None, None


def rt_table_with_grid(sigma, ds, mu=1.0, scatter=0.1, absorption=0.004, numStepsInt=20000, bins=4096):
    """(table (bins,) float64 tensor, grid = (minK, maxK, step, ds, absorption))."""
    grid = []
    t = _RTTable.apply(sigma, ds, mu, scatter, absorption, numStepsInt, bins, grid)
    return t, grid[0]


class _ExitCurve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, angles, boundary, N):
        th, ph = float(angles[0]), float(angles[1])
        bc = np.array(boundary, np.float32).copy()
        bc[9:12] = (math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))
        pos = host.initial_arc(bc, N)
        jt, jp = host.initial_arc_exit_jacobian(bc, N)
        ctx.save_for_backward(torch.from_numpy(jt), torch.from_numpy(jp))
        return torch.from_numpy(pos.astype(np.float64))

    @staticmethod
    def backward(ctx, g):
        jt, jp = ctx.saved_tensors
        return torch.stack([(g * jt).sum(), (g * jp).sum()]), None, None


def exit_curve(angles, boundary, N):
    """Initial curve (N+1, 3) for end-direction angles (theta, phi)."""
    return _ExitCurve.apply(angles, boundary, N)


class _LogEstimate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, table, positions, hctx, grid, boundary, num_paths, K, seed):
        minK, maxK, step, ds, absorption = grid
        hctx.set_tables(table.detach().numpy().astype(np.float32), minK, maxK, step, ds, absorption)
        bc = np.array(boundary, np.float32).copy()
        hctx.set_curve(positions.detach().numpy().astype(np.float32), bc)
        want_t, want_p = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if want_t or want_p:
            r = hctx.run_adjoint(num_paths, K, seed=seed, segments=want_p)
            ctx.dT = torch.from_numpy(r["d_table"])
            ctx.dS = torch.from_numpy(r["d_seg"]) if want_p else None
        else:
            r = hctx.run(num_paths, K, seed=seed)
        return torch.tensor(r["log10_sum_w"] - math.log10(num_paths), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gt = gp = None
        if ctx.needs_input_grad[0]:
            gt = g * ctx.dT
        if ctx.needs_input_grad[1]:
            dS = ctx.dS  # dE/dd_i with d_i = p_{i+1} - p_i
            gpos = torch.zeros((dS.shape[0] + 1, 3), dtype=torch.float64)
            gpos[1:] += dS
            gpos[:-1] -= dS
            gp = g * gpos
        return gt, gp, None, None, None, None, None, None


def log10_estimate(table, positions, hctx, grid, boundary, num_paths, K, seed=1):
    """E = log10(mean path weight) over num_paths K-perturbation paths on hctx."""
    return _LogEstimate.apply(table, positions, hctx, grid, boundary, num_paths, K, seed)
