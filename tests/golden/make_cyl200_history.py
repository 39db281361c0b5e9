"""Writes cyl200_history.npz: the oracle's full-length run of the 200^2 x 3
cylinder (diameter 1, 157 points, domain [-2, 2]^2, Re 100, dt 0.004, 2000
steps from the impulsive start; the reference's own defaults with
M = N = 200, MGRIDLEV = 3, ISTOP = 2000), and the CPU-vs-CPU noise floor over
the same window.

Run A is the oracle as it is; run B transforms along j before i (another
valid DST-I evaluation order, the rounding difference that any two correct
DST implementations show).  Both use run A's Cholesky factor (built by the
oracle's own preprocess).  Stored:

  hist_a, hist_b : 2000 rows of [fx, fy, circ(1..3), slip] (write_force,
                   write_circ, write_slip; src/operators.f90:184-186)
  omega_a        : run A's vorticity after step 2000, 3 levels
  fb_a           : run A's forces after step 2000
  chol_sum       : sum and sum of squares of the factor (a check that a
                   rebuilt factor is the one used here)
  floor_*        : run A vs run B (the yardstick for a device run)

The test that consumes it rebuilds the factor with the oracle and re-derives
the first 20 rows (tests/test_oracle.py), so the fixture is checked on every
CPU test run.  Takes about three minutes (two cores).
"""
import multiprocessing
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ibfs-if2d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as orc  # noqa: E402
from ibfs_amd import cases  # noqa: E402

NSTEPS = 2000


def run(job):
    args, order, cm, cv, nsteps = job
    o = orc.Oracle(**args)
    o.set_dst_order(order)
    o.set_cholesky(cm, cv)
    rows = []
    for _ in range(nsteps):
        f, c, s = o.advance()
        rows.append(np.concatenate([f, c, [s]]))
    return dict(hist=np.array(rows), omega=o.omega.copy(), q=o.q.copy(), fb=o.fb.copy(), L=o.L, nw=o.nw,
                nq=o.nq)


def rel_levels(a, b, size, L):
    return np.array([np.max(np.abs(a[k * size:(k + 1) * size] - b[k * size:(k + 1) * size]))
                     / np.max(np.abs(a[k * size:(k + 1) * size])) for k in range(L)])


def main():
    args = cases.cylinder_200()
    f0 = orc.Oracle(**args)
    f0.preprocess()  # raises if choldc fails
    cm, cv = f0.cholmat.copy(), f0.cholvec.copy()
    with multiprocessing.Pool(2) as pool:  # the two runs side by side
        ra, rb = pool.map(run, [(args, 0, cm, cv, NSTEPS), (args, 1, cm, cv, NSTEPS)])
    L, nw, nq = ra["L"], ra["nw"], ra["nq"]
    ha, hb = ra["hist"], rb["hist"]
    fscale = np.max(np.abs(ha[:, 0]))
    out = dict(
        hist_a=ha, hist_b=hb, omega_a=ra["omega"], fb_a=ra["fb"],
        chol_sum=np.array([cm.sum(), (cm * cm).sum(), cv.sum()]),
        floor_omega=rel_levels(ra["omega"], rb["omega"], nw, L),
        floor_q=rel_levels(ra["q"], rb["q"], nq, L),
        floor_force=np.max(np.abs(ha[:, :2] - hb[:, :2]), axis=0) / fscale,
        floor_fb=np.max(np.abs(ra["fb"] - rb["fb"])) / np.max(np.abs(ra["fb"])),
    )
    np.savez_compressed(os.path.join(HERE, "cyl200_history.npz"), **out)
    print({k: v for k, v in out.items() if k.startswith("floor")})


if __name__ == "__main__":
    main()
