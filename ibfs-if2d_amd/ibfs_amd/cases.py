"""The measured configurations as solver arguments (host scaffold).

Each function returns keyword arguments for ``hip.Solver`` / ``hip.Ensemble``
(m, n, mgridlev, dt, re, len_, offx, offy, xb, yb, bdyno): the namelist
values of input/ibfs.inp plus the geom.inp point list.
"""
from __future__ import annotations

import numpy as np

from . import geometry


def cylinder_200():
    """CPU reference case: 200^2 x 3 levels, domain [-2,2]^2, circle d=1
    (157 points), Re 100, dt 0.004."""
    m, ln = 200, 4.0
    xb, yb, bd = geometry.circle(ln / m, 1.0, 0.0, 0.0)
    return dict(m=m, n=m, mgridlev=3, dt=0.004, re=100.0, len_=ln, offx=2.0, offy=2.0, xb=xb, yb=yb, bdyno=bd)


def cylinder_grid(m: int, levels: int = 3):
    """m x m cells, domain [-2,2]^2, circle d = 1 at IB spacing = grid
    spacing, Re 100, dt = 0.25 * grid spacing (size sweeps)."""
    ln = 4.0
    xb, yb, bd = geometry.circle(ln / m, 1.0, 0.0, 0.0)
    return dict(m=m, n=m, mgridlev=levels, dt=0.25 * ln / m, re=100.0, len_=ln, offx=2.0, offy=2.0, xb=xb, yb=yb,
                bdyno=bd)


def naca_512():
    """Headline case: 512^2 x 5 levels, domain [-1,3] x [-2,2], NACA0012,
    chord 1, 20 deg nose-up about the quarter chord, IB spacing = grid
    spacing (261 points), Re 300, dt 0.002."""
    m, ln = 512, 4.0
    xb, yb, bd = geometry.naca4(ln / m, aoa_deg=20.0)
    return dict(m=m, n=m, mgridlev=5, dt=0.002, re=300.0, len_=ln, offx=1.0, offy=2.0, xb=xb, yb=yb, bdyno=bd)


def ensemble_64(seed: int = 0, ncases: int = 64):
    """Uneven ensemble: per case m = n in {256, 320, 384, 512}, 3..5 levels,
    even cases cylinders (d = 1), odd cases NACA0012 at AoA ~ U[0, 30] deg,
    Re ~ U[50, 300]; LEN 4, offsets (2, 2) / (1, 2), IB spacing = grid
    spacing, dt = 0.25 * grid spacing."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(ncases):
        m = int(rng.choice([256, 320, 384, 512]))
        L = int(rng.integers(3, 6))
        re = float(rng.uniform(50.0, 300.0))
        aoa = float(rng.uniform(0.0, 30.0))
        ln = 4.0
        d = ln / m
        if c % 2 == 0:
            xb, yb, bd = geometry.circle(d, 1.0, 0.0, 0.0)
            ox, oy = 2.0, 2.0
        else:
            xb, yb, bd = geometry.naca4(d, aoa_deg=aoa)
            ox, oy = 1.0, 2.0
        out.append(dict(m=m, n=m, mgridlev=L, dt=0.25 * d, re=re, len_=ln, offx=ox, offy=oy, xb=xb, yb=yb,
                        bdyno=bd))
    return out


def cylinder_384_adjoint():
    """Adjoint case: 384^2 x 4 levels, domain [-3.2,3.2]^2 (188 points on the
    d = 1 circle), Re 150, dt 0.003."""
    m, ln = 384, 6.4
    xb, yb, bd = geometry.circle(ln / m, 1.0, 0.0, 0.0)
    return dict(m=m, n=m, mgridlev=4, dt=0.003, re=150.0, len_=ln, offx=3.2, offy=3.2, xb=xb, yb=yb, bdyno=bd)


def tandem_1024():
    """Large case: 1024^2 x 6 levels, dx = 0.01, two d = 1 cylinders at
    (-2, 0) and (2, 0) (628 points, both body 0), Re 200, dt 0.0015."""
    m, ln = 1024, 10.24
    x1, y1, b1 = geometry.circle(ln / m, 1.0, -2.0, 0.0)
    x2, y2, b2 = geometry.circle(ln / m, 1.0, 2.0, 0.0)
    return dict(m=m, n=m, mgridlev=6, dt=0.0015, re=200.0, len_=ln, offx=5.12, offy=5.12,
                xb=np.concatenate([x1, x2]), yb=np.concatenate([y1, y2]), bdyno=np.concatenate([b1, b2]))
