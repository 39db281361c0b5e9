"""The 200^2 x 3 cylinder (the CPU reference case: Re 100, dt 0.004, 157
points) over its full 2000 steps on the device, against the oracle's run
stored in tests/golden/cyl200_history.npz (re-derived for its first 20 steps
by the CPU tests).  The yardstick is the CPU-vs-CPU floor stored beside it:
the same oracle with the other DST axis order, a rounding-level change of
the transforms that any two correct DST-I implementations show.  The device
run uses the oracle's own factor, rebuilt here by the oracle's preprocess,
as a restarted reference run would read it from ibfs.chd."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import relerr

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-11


def test_cylinder_200_full_history():
    from ibfs_amd import cases, hip

    fx = np.load(os.path.join(GOLDEN, "cyl200_history.npz"))
    args = cases.cylinder_200()
    o = orc.Oracle(**args)
    o.preprocess()
    cm, cv = o.cholmat.copy(), o.cholvec.copy()
    assert relerr(np.array([cm.sum(), (cm * cm).sum(), cv.sum()]), fx["chol_sum"]) < 1e-13
    g = hip.Solver(**args)
    g.set_cholesky(cm, cv)
    gh = np.concatenate([g.advance(500) for _ in range(4)])
    st = g.get_state()
    ha = fx["hist_a"]
    L, nw, nq = g.L, g.nw, g.nq
    # fields after step 2000; the oracle's fluxes from the stored vorticity
    _, q_ref = o.vort2flux(fx["omega_a"].copy(), L)
    fscale = np.max(np.abs(ha[:, 0]))
    res = {"force_x": np.max(np.abs(gh[:, 0] - ha[:, 0])) / fscale,
           "force_y": np.max(np.abs(gh[:, 1] - ha[:, 1])) / fscale,
           "fb": relerr(st["fb"], fx["fb_a"])}
    floor = {"force_x": fx["floor_force"][0], "force_y": fx["floor_force"][1], "fb": float(fx["floor_fb"])}
    for k in range(L):
        res[f"omega{k + 1}"] = relerr(st["omega"][k * nw:(k + 1) * nw], fx["omega_a"][k * nw:(k + 1) * nw])
        res[f"q{k + 1}"] = relerr(st["q"][k * nq:(k + 1) * nq], q_ref[k * nq:(k + 1) * nq])
        floor[f"omega{k + 1}"] = fx["floor_omega"][k]
        floor[f"q{k + 1}"] = fx["floor_q"][k]
    # circulation per level, relative to the level's total |omega|
    for k in range(L):
        scale = np.sum(np.abs(fx["omega_a"][k * nw:(k + 1) * nw]))
        res[f"circ{k + 1}"] = np.max(np.abs(gh[:, 2 + k] - ha[:, 2 + k])) / scale
    for key in sorted(res):
        print(f"{key:8s} device {res[key]:.2e}   cpu-vs-cpu floor {floor.get(key, float('nan')):.2e}")
    for key, v in res.items():
        assert v < max(TOL, 10.0 * floor.get(key, 0.0)), key
