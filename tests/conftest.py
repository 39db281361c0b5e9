import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ibfs-if2d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


# Small cases the oracle runs in well under a second per step.
#   name: (m, n, mgridlev, dt, Re, len, offx, offy, body)
CASES = {
    "cyl32": (32, 32, 3, 0.02, 100.0, 4.0, 2.0, 2.0, ("circle", 1.0)),
    "naca48x40": (48, 40, 2, 0.01, 200.0, 3.0, 1.0, 1.25, ("naca", 15.0)),
    "cyl40r5": (40, 40, 3, 0.02, 150.0, 5.0, 2.5, 2.5, ("circle", 1.0)),
    "cyl24r3": (24, 24, 2, 0.03, 80.0, 6.0, 3.0, 3.0, ("circle", 1.5)),
}


def case_geometry(name):
    from ibfs_amd import geometry

    m, n, L, dt, re, ln, ox, oy, body = CASES[name]
    delta = ln / m
    if body[0] == "circle":
        xb, yb, bd = geometry.circle(delta, body[1], 0.0, 0.0)
    else:
        xb, yb, bd = geometry.naca4(delta, aoa_deg=body[1], x_le=-0.25)
    return xb, yb, bd


def case_args(name):
    m, n, L, dt, re, ln, ox, oy, _ = CASES[name]
    xb, yb, bd = case_geometry(name)
    return dict(m=m, n=n, mgridlev=L, dt=dt, re=re, len_=ln, offx=ox, offy=oy, xb=xb, yb=yb, bdyno=bd)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)
