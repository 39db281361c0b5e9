"""The stationary = F force solve on the device: cjgr (src/operators.f90:
393-441) on the device a_times, warm-started from the previous fb, against
the oracle's cjgr step by step, and against the factor form."""
import numpy as np
import pytest

import oracle as orc
from conftest import case_args, relerr

pytestmark = pytest.mark.gpu

CGTOL = 1e-9
TOL = 1e-11  # the FP64 parity bar of the factor form


def cg_pair(name):
    from ibfs_amd import hip

    a = case_args(name)
    o = orc.Oracle(**a)
    g = hip.Solver(**a)
    o.set_solver(False, cgtol=CGTOL, cg_max_iter=3000)
    g.set_solver(False, cgtol=CGTOL, cg_max_iter=3000)
    return o, g


def run_oracle(o, nsteps):
    h, its = [], []
    for _ in range(nsteps):
        f, c, s = o.advance()
        h.append(np.concatenate([f, c, [s]]))
        its.append(o.cg_iters)
    return np.array(h), its


def errors(st, o, h, oh, nb2):
    e = {"omega": 0.0, "q": 0.0}
    for k in range(o.L):
        w = slice(k * o.nw, (k + 1) * o.nw)
        qq = slice(k * o.nq, (k + 1) * o.nq)
        e["omega"] = max(e["omega"], relerr(st["omega"][w], o.omega[w]))
        e["q"] = max(e["q"], relerr(st["q"][qq], o.q[qq]))
    e["fb"] = relerr(st["fb"], o.fb)
    e["force"] = np.max(np.abs(h[:, :nb2] - oh[:, :nb2])) / np.max(np.abs(oh[:, :nb2]))
    return e


@pytest.mark.parametrize("name", ["cyl32", "naca48x40"])
def test_cg_step_matches_oracle(name):
    """12 steps from rest, no factor built.  A is slightly nonsymmetric, so
    the CG path near its stopping test depends on rounding: two oracle runs
    that differ only in the rounding of the transforms (the other axis
    order) already differ by more than the 1e-11 bar of the factor form and
    may take different iteration counts.  That CPU-vs-CPU spread is the
    floor.  The device run meets the FP64 bar of the factor form (1e-11),
    or 10x that floor where the floor is larger."""
    from ibfs_amd import hip

    a = case_args(name)
    tol = 1e-12
    oa, ob = orc.Oracle(**a), orc.Oracle(**a)
    ob.set_dst_order(1)
    for o in (oa, ob):
        o.set_solver(False, cgtol=tol, cg_max_iter=3000)
    g = hip.Solver(**a)
    g.set_solver(False, cgtol=tol, cg_max_iter=3000)
    nsteps = 12
    ha, ita = run_oracle(oa, nsteps)
    hb, itb = run_oracle(ob, nsteps)
    hg = g.advance(nsteps)
    itg = g.cg_iterations(nsteps)
    nb2 = 2 * oa.nbody
    sb = {"omega": ob.omega, "q": ob.q, "fb": ob.fb}
    floor = errors(sb, oa, hb, ha, nb2)
    dev = errors(g.get_state(), oa, hg, ha, nb2)
    print("iterations oracle", ita, "oracle/other order", itb, "device", [int(v) for v in itg])
    print("floor", {k: f"{v:.1e}" for k, v in floor.items()}, "device", {k: f"{v:.1e}" for k, v in dev.items()})
    assert all(v > 0 for v in itg)
    # at cgtol 1e-12 the stopping test sits at the rounding level of r.r, so
    # the stopping iteration of a step depends on the summation order: the
    # counts are held in aggregate (and per step loosely), the fields below
    # to the floor
    assert abs(int(np.sum(itg)) - int(np.sum(ita))) <= 0.05 * np.sum(ita)
    assert np.max(np.abs(np.array(ita) - itg)) <= 0.15 * max(ita)
    for k in dev:
        assert dev[k] <= max(10 * floor[k], TOL), k


def test_cg_form_against_factor_form():
    """The CG solve of A and the factor of sym(A) (choldc reads the upper
    triangle) differ by the asymmetry of A, ~1e-5 relative: the force
    histories of the two forms agree to 1e-4 of the force scale."""
    from ibfs_amd import hip

    a = case_args("cyl32")
    gc = hip.Solver(**a)
    gc.build_cholesky()
    gg = hip.Solver(**a)
    gg.set_solver(False, cgtol=CGTOL)
    hc = gc.advance(12)
    hg = gg.advance(12)
    nb2 = 2 * gc.nbody
    assert np.max(np.abs(hc[:, :nb2] - hg[:, :nb2])) < 1e-4 * np.max(np.abs(hc[:, :nb2]))
    it = gg.cg_iterations(12)
    assert it[0] > 0 and it[-1] < it[0]  # the warm start pays after the impulsive start


def test_cg_form_needs_one_fp64_case():
    from ibfs_amd import hip

    a = case_args("cyl32")
    g = hip.Solver(**a, precision=hip.MIXED)
    g.set_solver(False)
    with pytest.raises(hip.IBFSError):
        g.advance(1)


def test_moving_body_matches_oracle():
    """A moving body: the cylinder of move_body's example 3 (src/grid.f90:
    555-567), oscillating in y with y'(t) = 0.5 cos t (here 5 cos 10t, so that
    it moves a few cells in 12 steps), bdyno 1, stationary = F.  Before each
    step both sides get the new point positions (setup_reg) and the point
    velocities times delta as the motion vector (src/grid.f90:573-574); CG
    solves each step's force system.  Checked against the oracle within 10x
    of the CPU-vs-CPU floor where that is above the FP64 bar, as the
    fixed-body CG test."""
    from ibfs_amd import hip

    a = case_args("cyl32")
    nb = len(a["xb"])
    a["bdyno"] = np.ones(nb, dtype=np.int32)
    tol = 1e-12
    oa, ob = orc.Oracle(**a), orc.Oracle(**a)
    ob.set_dst_order(1)
    g = hip.Solver(**a)
    for s in (oa, ob, g):
        s.set_solver(False, cgtol=tol, cg_max_iter=3000)
    xb0, yb0 = np.asarray(a["xb"], float), np.asarray(a["yb"], float)
    dt, delta = a["dt"], oa.delta
    ha, hb, hg = [], [], []
    for it in range(1, 13):
        t, t0 = dt * it, dt * (it - 1)
        yb = yb0 + 0.5 * (np.sin(10 * t) - np.sin(10 * t0))
        yb0 = yb
        motion = np.zeros(2 * nb)
        motion[nb:] = 5.0 * np.cos(10 * t) * delta
        for o, h in ((oa, ha), (ob, hb)):
            o.move(xb0, yb)
            f, c, s = o.advance(motion=motion)
            h.append(np.concatenate([f, c, [s]]))
        g.move_body(xb0, yb)
        hg.append(g.advance(1, motion=motion)[0])
    ha, hb, hg = np.array(ha), np.array(hb), np.array(hg)
    nb2 = 2 * oa.nbody
    floor = errors({"omega": ob.omega, "q": ob.q, "fb": ob.fb}, oa, hb, ha, nb2)
    dev = errors(g.get_state(), oa, hg, ha, nb2)
    print("floor", {k: f"{v:.1e}" for k, v in floor.items()}, "device", {k: f"{v:.1e}" for k, v in dev.items()})
    assert np.max(np.abs(ha[:, :nb2])) > 0
    for k in dev:
        assert dev[k] <= max(10 * floor[k], TOL), k
