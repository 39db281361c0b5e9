"""Immersed-boundary point sets (host scaffold, not on the device path).

``circle`` restates the reference primitive (src/grid.f90:496-520) with its
operation order, because the point coordinates feed the smearing weights and
every downstream value.  ``naca4`` is this project's generator for the
airfoil cases: the reference relies on MATLAB tooling for airfoils
(matlab/transform.m:1-71 resamples a surface at a given spacing,
matlab/rotate.m:3-15 rotates it about the quarter chord); here the NACA
4-digit closed form is resampled at equal arc length and rotated the same
way.  ``read_geom``/``write_geom`` use the geom.inp format read at
src/grid.f90:270-276.
"""
from __future__ import annotations

import math

import numpy as np

PI_REF = 3.1415926535897932384  # src/parameters.f90:35


def circle(ds: float, dia: float, x: float, y: float, num: int = 0):
    """src/grid.f90:496-520: npt = INT(pi*dia/ds) points, first at angle 0."""
    npt = int(PI_REF * dia / ds)
    i = np.arange(1, npt + 1, dtype=np.float64)
    ang = 2.0 * PI_REF * (i - 1.0) / float(npt)
    xb = x + 0.5 * dia * np.cos(ang)
    yb = y + 0.5 * dia * np.sin(ang)
    return xb, yb, np.full(npt, num, dtype=np.int32)


def _naca_half_thickness(xc: np.ndarray, t: float) -> np.ndarray:
    return 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc**2 + 0.2843 * xc**3 - 0.1015 * xc**4)


def naca4(ds: float, t: float = 0.12, chord: float = 1.0, aoa_deg: float = 0.0, x_le: float = 0.0,
          y_le: float = 0.0, pivot: float = 0.25, num: int = 0, npt: int | None = None):
    """Symmetric NACA 00tt airfoil, leading edge at (x_le, y_le) before the
    rotation, points at equal arc length (spacing ~ds) around the closed
    contour, rotated nose-up by aoa_deg about the point pivot*chord behind
    the leading edge (matlab/rotate.m:3-15 convention)."""
    # dense contour: upper surface TE -> LE, lower surface LE -> TE, closed back to start
    beta = np.linspace(0.0, math.pi, 20001)
    xc = 0.5 * (1.0 - np.cos(beta))  # cosine clustering at LE and TE
    yt = _naca_half_thickness(xc, t)
    xu, yu = xc[::-1], yt[::-1]
    xl, yl = xc[1:], -yt[1:]
    cx = np.concatenate([xu, xl, xu[:1]])
    cy = np.concatenate([yu, yl, yu[:1]])
    seg = np.hypot(np.diff(cx), np.diff(cy))
    s = np.concatenate([[0.0], np.cumsum(seg)])
    perim = s[-1] * chord
    if npt is None:
        npt = int(round(perim / ds))
    # start the walk at the leading edge so the point set is symmetric at aoa 0
    s_le = s[len(xu) - 1]
    st = (s_le + np.arange(npt) * (s[-1] / npt)) % s[-1]
    px = np.interp(st, s, cx) * chord
    py = np.interp(st, s, cy) * chord
    a = -aoa_deg * math.pi / 180.0
    xo = pivot * chord
    rx = xo + math.cos(a) * (px - xo) - math.sin(a) * py
    ry = math.sin(a) * (px - xo) + math.cos(a) * py
    return rx + x_le, ry + y_le, np.full(npt, num, dtype=np.int32)


def actuator_slot(x: float, y: float, hs: float, ds: float, angle: float, num: int):
    """The points of one actuator slot (src/grid.f90:636-718): a segment of
    width hs centred on (x, y), normal to the jet direction angle (degrees),
    as npt + 1 points with npt = INT(hs/ds) (one point at (x, y) when hs <
    ds), all with body number num (negative: -1 for actuator 001, ...).
    Same expressions and order as the reference: the segment direction is
    pi*(angle + 90)/180 and point i sits at x - hs/2 cos + (i-1) hs/npt cos."""
    npt = int(hs / ds)
    if npt == 0:
        return np.array([x]), np.array([y]), np.array([num], dtype=np.int32)
    a = PI_REF * (angle + 90.0) / 180.0
    ca, sa = math.cos(a), math.sin(a)
    seg = [float(i - 1) * hs / float(npt) for i in range(1, npt + 2)]
    xb = np.array([x - hs / 2.0 * ca + sg * ca for sg in seg])
    yb = np.array([y - hs / 2.0 * sa + sg * sa for sg in seg])
    return xb, yb, np.full(npt + 1, num, dtype=np.int32)


def write_geom(path: str, xb, yb, bdyno) -> None:
    """geom.inp: Nb, then 'x y bodyno' per line (src/grid.f90:272-275)."""
    with open(path, "w") as f:
        f.write(f"{len(xb)}\n")
        for x, y, b in zip(xb, yb, bdyno):
            f.write(f"{x!r} {y!r} {int(b)}\n")


def read_geom(path: str):
    with open(path) as f:
        toks = f.read().split()
    nb = int(toks[0])
    vals = toks[1:1 + 3 * nb]
    xb = np.array([float(v.replace("d", "e").replace("D", "E")) for v in vals[0::3]])
    yb = np.array([float(v.replace("d", "e").replace("D", "E")) for v in vals[1::3]])
    bd = np.array([int(v) for v in vals[2::3]], dtype=np.int32)
    return xb, yb, bd
