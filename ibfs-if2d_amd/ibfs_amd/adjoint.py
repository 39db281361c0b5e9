"""torch.autograd harness around the discrete adjoint of the IBFS step.

``ForceHistory.apply(u_B, solver, nsteps, ckpt)`` runs nsteps steps of a case
started from rest with the prescribed surface velocity u_B (Nf values: x
components then y components) and returns the (nsteps, rowlen) history rows
[fx(bodies), fy(bodies), circ(levels), slip], the lines of force.dat,
circ.dat and slip.dat.  Its forward pass is only the forward sweep (the
library keeps checkpoints of the state every ckpt steps on the device); its
backward pass is the reverse sweep of the discrete adjoint, driven by the
cotangent of every history row, so any differentiable function of the force
and circulation histories (not the slip, a maximum) has its gradient with
respect to u_B.  ``MeanForce`` is the mean drag/lift objective built on it.
Tensors live on the host (float64); the solver holds the device state; one
forward sweep per solver may be pending at a time.
"""
from __future__ import annotations

import numpy as np
import torch


def _sweep(ctx, solver, nsteps, ckpt, ub):
    """The forward sweep from rest; the solver remembers which sweep its
    device tape holds, so that a backward after other forwards (as
    gradcheck does) sweeps its own forward again first."""
    solver.reset()
    hist = solver.adjoint_forward(nsteps, ckpt, ub)
    token = object()
    solver._ibfs_tape = token
    ctx.solver, ctx.token, ctx.sweep = solver, token, (nsteps, ckpt, ub)
    return hist


def _reverse(ctx, g):
    solver = ctx.solver
    if getattr(solver, "_ibfs_tape", None) is not ctx.token:
        nsteps, ckpt, ub = ctx.sweep
        solver.reset()
        solver.adjoint_forward(nsteps, ckpt, ub)
        solver._ibfs_tape = ctx.token
    return torch.from_numpy(solver.adjoint_reverse(g))


class ForceHistory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ub, solver, nsteps, ckpt=50):
        hist = _sweep(ctx, solver, nsteps, ckpt, ub.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(hist)

    @staticmethod
    def backward(ctx, ghist):
        g = ghist.detach().cpu().numpy().astype(np.float64).copy()
        g[:, -1] = 0.0  # the slip column (a maximum) is not differentiated
        return _reverse(ctx, g), None, None, None


class MeanForce(torch.autograd.Function):
    """J = mean over the last nwin of nspin + nwin steps of cd * (sum of body x
    forces) + cl * (sum of y forces); forward runs the forward sweep only."""

    @staticmethod
    def forward(ctx, ub, solver, nspin, nwin, ckpt=50, cd=1.0, cl=0.0):
        hist = _sweep(ctx, solver, nspin + nwin, ckpt, ub.detach().cpu().numpy().astype(np.float64))
        nb = solver.nbody
        win = hist[nspin:]
        J = (cd * win[:, :nb].sum() + cl * win[:, nb:2 * nb].sum()) / nwin
        ctx.shape, ctx.args = hist.shape, (nspin, nwin, cd, cl, nb)
        return torch.tensor(float(J), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        nspin, nwin, cd, cl, nb = ctx.args
        g = np.zeros(ctx.shape)
        g[nspin:, :nb] = cd / nwin
        g[nspin:, nb:2 * nb] = cl / nwin
        return _reverse(ctx, float(gJ) * g), None, None, None, None, None, None


def objective_only(ub, solver, nspin, nwin, cd=1.0, cl=0.0) -> float:
    """J of the same run without any adjoint bookkeeping (for finite differences)."""
    solver.reset()
    return float(solver.objective(nspin, nwin, np.asarray(ub, dtype=np.float64), cd, cl)[0])
