"""ctypes front end of the CPU restatement in ``ibfs_oracle.c``.

TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
baseline timed beside it.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module.

Arrays follow the reference layouts (Fortran column-major, see
src/variables.f90:13-23 and src/grid.f90:119-134 of the reference):

* vertex fields ``omega``/``rhs_old``: (m-1)*(n-1) per level, i fastest;
* fluxes ``q``/``q0``: Nq per level, all u-faces (i fastest) then all v-faces;
* forces ``fb``: Nf = 2*Nb, all x components then all y components;
* ``cholmat``: Nf x Nf column-major, ``cholvec``: Nf.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libibfs_oracle.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, _D, _D, _I, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_dims.argtypes = [C.c_void_p, _I]
        L.orc_ptr.restype = _D
        L.orc_ptr.argtypes = [C.c_void_p, C.c_int]
        L.orc_ismear.restype = _I
        L.orc_ismear.argtypes = [C.c_void_p]
        L.orc_scalar.restype = C.c_double
        L.orc_scalar.argtypes = [C.c_void_p, C.c_int]
        L.orc_factors.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_nlevmax.restype = C.c_int
        L.orc_nlevmax.argtypes = [C.c_int, C.c_int]
        for nm in ("orc_dst", "orc_ctci", "orc_ainv", "orc_rot", "orc_reg", "orc_regT", "orc_a_times",
                   "orc_cholsl", "orc_force"):
            getattr(L, nm).argtypes = [C.c_void_p, _D, _D]
        L.orc_curl.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_coarsify.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_get_bc.argtypes = [C.c_void_p, _D, _D, C.c_double]
        L.orc_apply_bc.argtypes = [C.c_void_p, _D, _D, C.c_double]
        L.orc_nonlinear.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
        L.orc_vort2flux.argtypes = [C.c_void_p, _D, _D, C.c_int]
        L.orc_choldc.restype = C.c_int
        L.orc_choldc.argtypes = [C.c_void_p]
        L.orc_preprocess.restype = C.c_int
        L.orc_preprocess.argtypes = [C.c_void_p]
        L.orc_a_column.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_circ.argtypes = [C.c_void_p, _D]
        L.orc_advance.restype = C.c_int
        L.orc_advance.argtypes = [C.c_void_p, _I, _D, C.c_int, _D, _D, _D]
        L.orc_step1.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_dst_order.argtypes = [C.c_void_p, C.c_int]
        L.orc_cjgr.restype = C.c_int
        L.orc_cjgr.argtypes = [C.c_void_p, _D, _D]
        L.orc_set_solver.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.orc_cg_iters.restype = C.c_int
        L.orc_cg_iters.argtypes = [C.c_void_p]
        L.orc_move.argtypes = [C.c_void_p, _D, _D]
        L.orc_dst1.argtypes = [_D, _D, C.c_int]
        L.orc_fft.argtypes = [_D, _D, C.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def dst1(x: np.ndarray) -> np.ndarray:
    """1-D unnormalised DST-I (RODFT00) of a vector."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_dst1(_p(x), _p(y), x.size)
    return y


def fft(x: np.ndarray) -> np.ndarray:
    """Forward complex DFT exp(-2 pi i jk/n) of a vector, by the oracle's FFT."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    xr = x.view(np.float64).copy()
    y = np.empty_like(xr)
    lib().orc_fft(_p(xr), _p(y), x.size)
    return y.view(np.complex128)


def actuation(itime, nb, delta, dt, npt, params, motion=None):
    """actuation and u_jet (src/grid.f90:579-633), the motion the reference
    hands to the step when actuators exist (src/operators.f90:150-152): a
    copy of motion (zeros if None, length 2 nb) with, for every point of slot
    a (the last sum(npt) points, in slot order), x and y entries
      (u_j + u_j_osc sin(omg_j dt itime)) (cos, sin)(pi/180 theta_j) * delta,
    params[a] = (x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j)."""
    import math

    pi = 3.1415926535897932384  # src/parameters.f90:35
    out = np.zeros(2 * nb) if motion is None else np.array(motion, dtype=np.float64)
    n_act_pt = int(np.sum(npt))
    count = 0
    for a, n_pt in enumerate(npt):
        _, _, _, u_j, u_j_osc, theta_j, omg_j = (float(v) for v in params[a])
        for _ in range(int(n_pt)):
            count += 1
            i = nb - n_act_pt + count  # 1-based point index
            base = u_j + u_j_osc * math.sin(omg_j * dt * float(itime))
            out[i - 1] = base * math.cos(pi / 180.0 * theta_j) * delta
            out[nb + i - 1] = base * math.sin(pi / 180.0 * theta_j) * delta
    return out


class Oracle:
    """One reference-layout simulation state plus its operators."""

    def __init__(self, m, n, mgridlev, dt, re, len_, offx, offy, xb, yb, bdyno=None, unif_flow=True,
                 clamp_levels=True):
        L = lib()
        if clamp_levels:
            mgridlev = min(mgridlev, L.orc_nlevmax(m, n))
        xb = np.ascontiguousarray(xb, dtype=np.float64)
        yb = np.ascontiguousarray(yb, dtype=np.float64)
        nb = xb.size
        bd = np.zeros(nb, dtype=np.int32) if bdyno is None else np.ascontiguousarray(bdyno, dtype=np.int32)
        self._h = L.orc_create(m, n, mgridlev, dt, re, len_, offx, offy, nb, _p(xb), _p(yb),
                               bd.ctypes.data_as(_I), 1 if unif_flow else 0)
        if not self._h:
            raise ValueError("orc_create rejected the grid (m, n must be multiples of 4, >= 8)")
        d = (C.c_int * 11)()
        L.orc_dims(self._h, d)
        (self.m, self.n, self.L, self.nb, self.nf, self.nq, self.nw, self.nbc, self.nsmear,
         self.bmin, self.bmax) = list(d)
        self.nbody = self.bmax - self.bmin + 1
        self.dt, self.re = dt, re
        self.delta = L.orc_scalar(self._h, 0)
        self.normalize = L.orc_scalar(self._h, 1)
        self.itime = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    # ---- raw state views (live, writable) ----
    def _view(self, which, count):
        ptr = lib().orc_ptr(self._h, which)
        return np.ctypeslib.as_array(ptr, shape=(count,))

    @property
    def omega(self):
        return self._view(0, self.nw * self.L)

    @property
    def rhs_old(self):
        return self._view(1, self.nw * self.L)

    @property
    def q(self):
        return self._view(2, self.nq * self.L)

    @property
    def q0(self):
        return self._view(3, self.nq * self.L)

    @property
    def fb(self):
        return self._view(4, self.nf)

    @property
    def cholmat(self):
        return self._view(5, self.nf * self.nf)

    @property
    def cholvec(self):
        return self._view(6, self.nf)

    @property
    def smear(self):
        return self._view(7, self.nq * max(self.nsmear, 1))

    @property
    def ismear(self):
        ptr = lib().orc_ismear(self._h)
        return np.ctypeslib.as_array(ptr, shape=(self.nq * max(self.nsmear, 1),))

    @property
    def lam(self):
        return self._view(8, self.nw)

    @property
    def lam1(self):
        return self._view(9, self.nw * self.L)

    @property
    def lam1i(self):
        return self._view(10, self.nw * self.L)

    @property
    def laminv(self):
        return self._view(11, self.nw)

    def factors(self):
        c1 = np.zeros(self.L)
        c2 = np.zeros(self.L)
        vf = np.zeros(self.L)
        lib().orc_factors(self._h, _p(c1), _p(c2), _p(vf))
        return c1, c2, vf

    # ---- operators (new arrays out) ----
    def _w(self):
        return np.zeros(self.nw)

    def dst(self, x):
        y = self._w()
        lib().orc_dst(self._h, _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def ctci(self, x):
        y = self._w()
        lib().orc_ctci(self._h, _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def ainv(self, x):
        y = self._w()
        lib().orc_ainv(self._h, _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def rot(self, x):
        y = self._w()
        lib().orc_rot(self._h, _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def curl(self, x, xbc):
        y = np.zeros(self.nq)
        lib().orc_curl(self._h, _p(np.ascontiguousarray(x, np.float64)),
                       _p(np.ascontiguousarray(xbc, np.float64)), _p(y))
        return y

    def coarsify(self, crhs, rhs):
        y = self._w()
        lib().orc_coarsify(self._h, _p(np.ascontiguousarray(crhs, np.float64)),
                           _p(np.ascontiguousarray(rhs, np.float64)), _p(y))
        return y

    def get_bc(self, r, fac):
        y = np.zeros(self.nbc)
        lib().orc_get_bc(self._h, _p(np.ascontiguousarray(r, np.float64)), _p(y), fac)
        return y

    def apply_bc(self, r, rbc, fac):
        r = np.array(r, dtype=np.float64, copy=True)
        lib().orc_apply_bc(self._h, _p(r), _p(np.ascontiguousarray(rbc, np.float64)), fac)
        return r

    def nonlinear(self, omega, q, q0, bc):
        y = np.zeros(self.nq)
        a = [np.ascontiguousarray(v, np.float64) for v in (omega, q, q0, bc)]
        lib().orc_nonlinear(self._h, *[_p(v) for v in a], _p(y))
        return y

    def reg(self, h0):
        y = np.zeros(self.nq)
        lib().orc_reg(self._h, _p(np.ascontiguousarray(h0, np.float64)), _p(y))
        return y

    def regT(self, h):
        y = np.zeros(self.nf)
        lib().orc_regT(self._h, _p(np.ascontiguousarray(h, np.float64)), _p(y))
        return y

    def vort2flux(self, omega, nlev):
        """Returns (omega after in-place coarsify, vel (Nq*nlev))."""
        om = np.array(omega, dtype=np.float64, copy=True)
        vel = np.zeros(self.nq * nlev)
        lib().orc_vort2flux(self._h, _p(om), _p(vel), nlev)
        return om, vel

    def a_times(self, x):
        y = np.zeros(self.nf)
        lib().orc_a_times(self._h, _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def a_column(self, i):
        y = np.zeros(self.nf)
        lib().orc_a_column(self._h, int(i), _p(y))
        return y

    def preprocess(self):
        rc = lib().orc_preprocess(self._h)
        if rc:
            raise RuntimeError(f"choldc failed at row {rc}")

    def choldc(self):
        return lib().orc_choldc(self._h)

    def cholsl(self, b):
        y = np.zeros(self.nf)
        lib().orc_cholsl(self._h, _p(np.ascontiguousarray(b, np.float64)), _p(y))
        return y

    def cjgr(self, b, x0=None):
        """cjgr (src/operators.f90:393-441) from the guess x0 (zero if None).
        Returns (x, iterations)."""
        x = np.zeros(self.nf) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        it = lib().orc_cjgr(self._h, _p(x), _p(np.ascontiguousarray(b, np.float64)))
        return x, it

    def set_solver(self, stationary, cgtol=1e-7, cg_max_iter=3000):
        """stationary: the factor's cholsl; otherwise cjgr warm-started from
        the previous fb (src/operators.f90:166-171)."""
        lib().orc_set_solver(self._h, 1 if stationary else 0, float(cgtol), int(cg_max_iter))

    @property
    def cg_iters(self):
        return lib().orc_cg_iters(self._h)

    def move(self, xb, yb):
        """New body point positions and setup_reg (src/operators.f90:142-144)."""
        lib().orc_move(self._h, _p(np.ascontiguousarray(xb, np.float64)), _p(np.ascontiguousarray(yb, np.float64)))

    def force(self, xg):
        y = np.zeros(2 * self.nbody)
        lib().orc_force(self._h, _p(np.ascontiguousarray(xg, np.float64)), _p(y))
        return y

    def circ(self):
        y = np.zeros(self.L)
        lib().orc_circ(self._h, _p(y))
        return y

    def set_dst_order(self, order):
        """0: transform along i first (default); 1: along j first."""
        lib().orc_set_dst_order(self._h, int(order))

    def step1(self, first):
        """STEP 1 of advance only (intermediate vorticity on every level)."""
        lib().orc_step1(self._h, 1 if first else 0)

    def set_cholesky(self, cholmat, cholvec):
        self.cholmat[:] = np.asarray(cholmat, dtype=np.float64).ravel()
        self.cholvec[:] = np.asarray(cholvec, dtype=np.float64).ravel()
        self._have_chol = True

    def advance(self, motion=None, preprocess=None):
        """One reference time step.  Returns (force, circ, slip)."""
        if preprocess is None:
            preprocess = not getattr(self, "_have_chol", False)
        it = C.c_int(self.itime)
        force = np.zeros(2 * self.nbody)
        circ = np.zeros(self.L)
        slip = np.zeros(1)
        mo = None if motion is None else np.ascontiguousarray(motion, np.float64)
        rc = lib().orc_advance(self._h, C.byref(it), _p(mo), 1 if preprocess else 0, _p(force), _p(circ),
                               _p(slip))
        if rc:
            raise RuntimeError(f"choldc failed at row {rc}")
        if preprocess:
            self._have_chol = True
        self.itime = it.value
        return force, circ, float(slip[0])
