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


def test_force_history_nonlinear_functional():
    """ForceHistory: the forward pass is the forward sweep only and the
    reverse sweep runs in backward, driven by the cotangent of every history
    row.  A functional that is not linear in the forces, the mean of F_x^2
    over the window plus a circulation term, passes gradcheck and matches
    central differences."""
    import torch

    from ibfs_amd.adjoint import ForceHistory

    s = make("cyl24r3")
    nsp, nwin = 6, 4

    def J(u):
        h = ForceHistory.apply(u, s, nsp + nwin, 3)
        fx = h[nsp:, 0]
        return (fx * fx).mean() + 0.01 * h[nsp:, 2 * s.nbody].mean()

    ub0 = ub_for(s)
    ub = torch.tensor(ub0, dtype=torch.float64, requires_grad=True)
    Jv = J(ub)
    Jv.backward()
    g = ub.grad.numpy().copy()
    assert np.all(np.isfinite(g)) and np.linalg.norm(g) > 0
    rng = np.random.default_rng(3)
    eps = 1e-5
    for _ in range(3):
        d = rng.standard_normal(s.nf)
        d /= np.linalg.norm(d)
        with torch.no_grad():
            fd = (J(torch.tensor(ub0 + eps * d)) - J(torch.tensor(ub0 - eps * d))).item() / (2 * eps)
        print("fd", fd, "adjoint", g @ d)
        assert abs(fd - g @ d) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(g))
    assert torch.autograd.gradcheck(J, (torch.tensor(ub0, requires_grad=True),), eps=1e-5, atol=1e-8, rtol=1e-4)


@pytest.mark.parametrize("itime", [1, 5])
def test_step_vjp_matches_finite_differences(itime):
    """ibfs_step_vjp: the cotangents of one step's outputs (omega, rhs_old, q
    and its history row: body forces and circulations) give those of its
    inputs (omega, rhs_old, q) and of its motion vector.  Checked by central
    differences of phi = <a, step(x)> along random directions of each input,
    from a state reached by a few steps (at itime 1 the step is the Euler
    start, which ignores rhs_old)."""
    s = make("cyl32")
    s.build_cholesky()
    s.advance(3)
    st0 = s.get_state()
    rng = np.random.default_rng(11)
    W, Q = s.nw * s.L, s.nq * s.L
    a = {"omega": rng.standard_normal(W), "rhs_old": rng.standard_normal(W), "q": rng.standard_normal(Q),
         "row": rng.standard_normal(s.ncols)}
    a["row"][-1] = 0.0  # the slip is not differentiated

    def phi(omega, rhs_old, q, motion):
        s.set_state(omega=omega, rhs_old=rhs_old, q=q)
        s.itime = itime - 1
        row = s.advance(1, motion=motion)[0]
        out = s.get_state()
        return (a["omega"] @ out["omega"] + a["rhs_old"] @ out["rhs_old"] + a["q"] @ out["q"]
                + a["row"] @ row)

    g = s.step_vjp(itime, st0["omega"], st0["q"], a["omega"], a["rhs_old"], a["q"], a["row"])
    base = dict(omega=st0["omega"], rhs_old=st0["rhs_old"], q=st0["q"], motion=np.zeros(s.nf))
    for key, scale in (("omega", np.abs(st0["omega"]).max()), ("rhs_old", np.abs(st0["rhs_old"]).max() + 1e-3),
                       ("q", np.abs(st0["q"]).max()), ("motion", s.delta)):
        d = rng.standard_normal(base[key].size)
        d *= scale / np.linalg.norm(d)
        eps = 1e-4
        xp = dict(base); xp[key] = base[key] + eps * d
        xm = dict(base); xm[key] = base[key] - eps * d
        fd = (phi(**xp) - phi(**xm)) / (2 * eps)
        ad = g[key] @ d
        print(itime, key, "fd", fd, "vjp", ad)
        if key == "rhs_old" and itime == 1:
            assert ad == 0.0 and abs(fd) < 1e-9 * np.linalg.norm(d)
        else:
            assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-6)
