"""Discrete adjoint of the time step: gradients of the mean drag/lift over a
window with respect to the prescribed IB surface velocity u_B.

Checks (FP64):
* adjoint vs the FP64 tangent-linear model: grad . du == dJ_TLM(du) for
  random directions (the dot-product identity; exact up to rounding),
* adjoint vs central finite differences of the forward model on random
  unit directions (within 1e-4 relative),
* the checkpoint interval does not change the gradient,
* torch.autograd harness: gradcheck against finite differences."""
import numpy as np
import pytest

from conftest import case_args

pytestmark = pytest.mark.gpu

NSPIN, NWIN = 12, 8


def make(name="cyl32"):
    from ibfs_amd import hip

    s = hip.Solver(**case_args(name))
    s.build_cholesky()
    return s


def ub_for(s, seed=0, sd=0.05):
    return np.random.default_rng(seed).normal(0.0, sd, s.nf)


@pytest.mark.parametrize("name", ["cyl32", "naca48x40"])
def test_adjoint_matches_tangent_linear(name):
    s = make(name)
    ub = ub_for(s)
    s.reset()
    J, g = s.adjoint(NSPIN, NWIN, 5, ub, cd=1.0, cl=0.5)
    rng = np.random.default_rng(7)
    for _ in range(3):
        du = rng.standard_normal(s.nf)
        s.reset()
        Jt, dJ = s.tangent(NSPIN, NWIN, ub, du, cd=1.0, cl=0.5)
        assert abs(Jt[0] - J[0]) <= 1e-12 * abs(J[0])
        print(name, "adjoint", g @ du, "tlm", dJ[0])
        assert abs(g @ du - dJ[0]) <= 1e-9 * (abs(dJ[0]) + 1e-30) + 1e-14 * np.abs(g).sum() * np.abs(du).max()


def test_adjoint_matches_central_differences():
    from ibfs_amd.adjoint import objective_only

    s = make("cyl32")
    ub = ub_for(s)
    s.reset()
    J, g = s.adjoint(NSPIN, NWIN, 4, ub, cd=1.0, cl=0.0)
    rng = np.random.default_rng(1)
    eps = 1e-4
    for _ in range(4):
        d = rng.standard_normal(s.nf)
        d /= np.linalg.norm(d)
        fd = (objective_only(ub + eps * d, s, NSPIN, NWIN) - objective_only(ub - eps * d, s, NSPIN, NWIN)) / (2 * eps)
        print("fd", fd, "adjoint", g @ d)
        assert abs(fd - g @ d) <= 1e-4 * max(abs(fd), 1e-3 * np.linalg.norm(g))


def test_checkpoint_interval_is_irrelevant():
    s = make("cyl32")
    ub = ub_for(s)
    grads = []
    for ck in (1, 3, 7, NSPIN + NWIN):
        s.reset()
        grads.append(s.adjoint(NSPIN, NWIN, ck, ub)[1])
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


def test_torch_autograd_harness():
    import torch

    from ibfs_amd.adjoint import MeanForce

    s = make("cyl24r3")
    ub = torch.tensor(ub_for(s), dtype=torch.float64, requires_grad=True)
    J = MeanForce.apply(ub, s, 6, 4, 3, 1.0, 0.3)
    J.backward()
    g = ub.grad.clone()
    assert g.shape == ub.shape and torch.isfinite(g).all()
    f = lambda u: MeanForce.apply(u, s, 6, 4, 3, 1.0, 0.3)
    assert torch.autograd.gradcheck(f, (ub.detach().clone().requires_grad_(True),), eps=1e-5, atol=1e-8, rtol=1e-4)
