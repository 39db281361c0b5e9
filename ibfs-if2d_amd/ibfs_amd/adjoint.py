"""torch.autograd harness around the discrete adjoint of the IBFS step.

``MeanForce.apply(u_B, solver, nspin, nwin, ckpt, cd, cl)`` returns the
objective J = mean over the window of (cd * drag + cl * lift) of a case
started from rest, as a differentiable function of the prescribed surface
velocity u_B (Nf values: x components then y components).  The forward pass
runs the HIP library's adjoint driver (forward sweep with checkpoints, then
the reverse sweep), so backward only scales the stored gradient.  Tensors
live on the host (float64); the solver holds the device state.
"""
from __future__ import annotations

import numpy as np
import torch


class MeanForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ub, solver, nspin, nwin, ckpt=50, cd=1.0, cl=0.0):
        solver.reset()
        J, grad = solver.adjoint(nspin, nwin, ckpt, ub.detach().cpu().numpy().astype(np.float64), cd, cl)
        ctx.save_for_backward(torch.from_numpy(grad))
        return torch.tensor(float(J[0]), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None, None, None


def objective_only(ub, solver, nspin, nwin, cd=1.0, cl=0.0) -> float:
    """J of the same run without the reverse sweep (for finite differences)."""
    solver.reset()
    return float(solver.objective(nspin, nwin, np.asarray(ub, dtype=np.float64), cd, cl)[0])
