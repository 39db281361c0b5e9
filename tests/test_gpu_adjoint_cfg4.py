"""The adjoint configuration at full size: 384^2 x 4 levels, cylinder d = 1
(188 IB points), Re 150, dt 0.003; objective = mean drag over a 400-step
window after a 2000-step spin-up; control u_B in R^{188x2} ~ N(0, 0.05^2)
(seed 0), active in every step; checkpoints every 50 steps.  Checks: central
finite differences on 16 random unit directions (seed 1) within 1e-4, and
the FP64 tangent-linear model along one of them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cfg4_gradient_vs_finite_differences_and_tlm():
    from ibfs_amd import cases, hip
    from ibfs_amd.adjoint import objective_only

    s = hip.Solver(**cases.cylinder_384_adjoint())
    assert s.nb == 188 and s.L == 4
    s.build_cholesky()
    ub = np.random.default_rng(0).normal(0.0, 0.05, s.nf)
    s.reset()
    J, g = s.adjoint(2000, 400, 50, ub, 1.0, 0.0)
    rng = np.random.default_rng(1)
    eps = 2e-4  # truncation ~(eps^2) J''' stays well below 1e-4; rounding of J ~1e-13 / eps is negligible
    worst = 0.0
    dirs = []
    for k in range(16):
        d = rng.standard_normal(s.nf)
        d /= np.linalg.norm(d)
        dirs.append(d)
        fd = (objective_only(ub + eps * d, s, 2000, 400) - objective_only(ub - eps * d, s, 2000, 400)) / (2 * eps)
        ad = g @ d
        err = abs(fd - ad) / max(abs(fd), 1e-12)
        worst = max(worst, err)
        print(f"direction {k}: fd {fd:.10e} adjoint {ad:.10e} rel {err:.2e}")
    print("worst relative difference", worst)
    assert worst < 1e-4
    s.reset()
    Jt, dJ = s.tangent(2000, 400, ub, dirs[0], 1.0, 0.0)
    print("tlm", dJ[0], "adjoint", g @ dirs[0])
    assert abs(dJ[0] - g @ dirs[0]) <= 1e-8 * abs(dJ[0])
    assert abs(Jt[0] - J[0]) <= 1e-12 * abs(J[0])


def test_cfg4_nonlinear_force_functional():
    """The same case with a functional that is not linear in the forces: J =
    mean over the 400-step window of F_x^2 (ForceHistory, reverse sweep in
    backward), against central differences on 4 random directions."""
    import torch

    from ibfs_amd import cases, hip
    from ibfs_amd.adjoint import ForceHistory

    s = hip.Solver(**cases.cylinder_384_adjoint())
    s.build_cholesky()
    ub0 = np.random.default_rng(0).normal(0.0, 0.05, s.nf)

    def J(u):
        fx = ForceHistory.apply(u, s, 2400, 50)[2000:, 0]
        return (fx * fx).mean()

    ub = torch.tensor(ub0, requires_grad=True)
    J(ub).backward()
    g = ub.grad.numpy().copy()
    rng = np.random.default_rng(2)
    eps = 2e-4
    worst = 0.0
    with torch.no_grad():
        for k in range(4):
            d = rng.standard_normal(s.nf)
            d /= np.linalg.norm(d)
            fd = (J(torch.tensor(ub0 + eps * d)) - J(torch.tensor(ub0 - eps * d))).item() / (2 * eps)
            err = abs(fd - g @ d) / max(abs(fd), 1e-12)
            worst = max(worst, err)
            print(f"direction {k}: fd {fd:.10e} adjoint {g @ d:.10e} rel {err:.2e}")
    assert worst < 1e-4
