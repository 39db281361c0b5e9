"""Actuator slots (src/grid.f90:579-718, src/operators.f90:150-152): a
cylinder with two jets appended as slots with body numbers -1 and -2 (a
three-point slot with a steady plus oscillating jet behind the body, a
one-point slot with a purely oscillating jet above it).  The library
evaluates the jets on the device every step (ibfs_set_actuators); the oracle
gets the motion from its restatement of actuation/u_jet.  Histories, whose
force columns cover body numbers -2..0, and the final state must agree to
the FP64 bar; the jets must keep their timing across split advance calls."""
import numpy as np
import pytest

import oracle as orc
from conftest import case_args, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-11
# (x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j): read_act_parameters
PARAMS = [(0.9, 0.0, 0.3, 0.5, 0.25, 0.0, 3.0), (0.0, 0.9, 0.1, 0.0, 0.8, 90.0, 5.0)]


def actuated_case():
    from ibfs_amd import geometry

    a = case_args("cyl32")
    delta = a["len_"] / a["m"]
    xs, ys, bs, npt = [a["xb"]], [a["yb"]], [a["bdyno"]], []
    for k, p in enumerate(PARAMS):
        px, py, pb = geometry.actuator_slot(p[0], p[1], p[2], delta, p[5], -(k + 1))
        xs.append(px)
        ys.append(py)
        bs.append(pb)
        npt.append(len(px))
    a.update(xb=np.concatenate(xs), yb=np.concatenate(ys), bdyno=np.concatenate(bs).astype(np.int32))
    return a, npt, delta


def oracle_run(a, npt, delta, cm, cv, nsteps):
    o = orc.Oracle(**a)
    o.set_cholesky(cm, cv)
    rows = []
    for _ in range(nsteps):
        mo = orc.actuation(o.itime + 1, o.nb, delta, a["dt"], npt, PARAMS)
        f, c, s = o.advance(motion=mo)
        rows.append(np.concatenate([f, c, [s]]))
    return o, np.array(rows)


def test_actuated_cylinder_matches_oracle():
    from ibfs_amd import hip

    a, npt, delta = actuated_case()
    assert npt == [3, 1]
    g = hip.Solver(**a)
    assert (g.bmin, g.bmax, g.nbody) == (-2, 0, 3)
    g.set_actuators(npt, PARAMS)
    cm, cv = g.build_cholesky()
    o, oh = oracle_run(a, npt, delta, cm, cv, 20)
    gh = np.concatenate([g.advance(7), g.advance(13)])  # the jets keep their timing across calls
    st = g.get_state()
    nbd = g.nbody
    for b in range(nbd):  # body numbers -2, -1, 0, each on its own force scale
        cols = [b, nbd + b]
        scale = np.max(np.abs(oh[:, cols]))
        assert scale > 0, f"body {b - 2} carries no force"
        assert np.max(np.abs(gh[:, cols] - oh[:, cols])) < TOL * scale, f"force of body {b - 2}"
    for k in range(o.L):
        scale = np.sum(np.abs(o.omega[k * o.nw:(k + 1) * o.nw]))
        assert np.max(np.abs(gh[:, 2 * nbd + k] - oh[:, 2 * nbd + k])) < TOL * scale, f"circulation {k + 1}"
    # slip is a small difference of O(delta) terms: compared on the delta scale
    assert np.max(np.abs(gh[:, -1] - oh[:, -1])) < TOL * delta
    for k in range(o.L):
        w = slice(k * o.nw, (k + 1) * o.nw)
        qq = slice(k * o.nq, (k + 1) * o.nq)
        assert relerr(st["omega"][w], o.omega[w]) < TOL, f"omega level {k + 1}"
        assert relerr(st["q"][qq], o.q[qq]) < TOL, f"q level {k + 1}"
    assert relerr(st["fb"], o.fb) < TOL


def test_actuators_removed_and_mixed_form():
    """n_act = 0 removes the jets (the run equals a run with zero motion);
    the mixed form with jets stays within 1e-5 of the oracle."""
    from ibfs_amd import hip

    a, npt, delta = actuated_case()
    g = hip.Solver(**a)
    cm, cv = g.build_cholesky()
    g.set_actuators(npt, PARAMS)
    g.set_actuators([], np.zeros((0, 7)))
    h0 = g.advance(5)
    g2 = hip.Solver(**a)
    g2.set_cholesky(cm, cv)
    assert np.array_equal(h0, g2.advance(5))
    m = hip.Solver(**a, precision=hip.MIXED)
    m.set_cholesky(cm, cv)
    m.set_actuators(npt, PARAMS)
    mh = m.advance(20)
    o, oh = oracle_run(a, npt, delta, cm, cv, 20)
    nb = m.nbody
    assert np.max(np.abs(mh[:, : 2 * nb] - oh[:, : 2 * nb])) < 1e-5 * np.max(np.abs(oh[:, : 2 * nb]))
    st = m.get_state()
    for k in range(o.L):
        w = slice(k * o.nw, (k + 1) * o.nw)
        assert relerr(st["omega"][w], o.omega[w]) < 1e-5


def test_bad_actuator_arguments_are_rejected():
    from ibfs_amd import hip

    a, npt, delta = actuated_case()
    g = hip.Solver(**a)
    with pytest.raises(hip.IBFSError, match=r"\(1\)"):
        g.set_actuators([0], [PARAMS[0]])
    with pytest.raises(hip.IBFSError, match=r"\(1\)"):
        g.set_actuators([g.nb + 1], [PARAMS[0]])
    bad = list(PARAMS[0])
    bad[3] = float("nan")
    with pytest.raises(hip.IBFSError, match=r"\(1\)"):
        g.set_actuators([3], [bad])
    g.set_actuators(npt, PARAMS)
    g.build_cholesky()
    with pytest.raises(hip.IBFSError, match="actuators"):
        g.adjoint(2, 2, 2, np.zeros(g.nf))
