"""ctypes binding of the gfx950 IBFS library (include/ibfs.h).

``Solver`` mirrors the reference's module procedures: ``advance`` is
src/operators.f90:62-188, ``vort2flux``/``curl``/``rot``/``coarsify``/
``get_bc``/``reg``/``regT``/``a_times``/``cholsl`` are the operators of
src/operators.f90 and ``dst``/``ctci``/``ainv`` those of src/myfft.f90.
Arrays are numpy float64 in the reference layouts (Fortran column-major,
flattened).  Errors raise ``IBFSError`` with the library's message, the way
the reference STOPs.

There is no CPU fallback: without ``libibfs_hip.so`` the import fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libibfs_hip.so")

PLAIN_FP64 = 0
MIXED = 1

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class IBFSError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ibfs_amd.build` "
                          "(the IBFS step has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.ibfs_create.restype = C.c_int
    L.ibfs_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_double, _D, _D, _I, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.ibfs_dims.argtypes = [C.c_void_p, _I]
    L.ibfs_set_state.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
    L.ibfs_get_state.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
    L.ibfs_set_cholesky.argtypes = [C.c_void_p, _D, _D]
    L.ibfs_build_cholesky.argtypes = [C.c_void_p, _D, _D]
    L.ibfs_step.argtypes = [C.c_void_p, _I, _D, _D, _D, _D, _D]
    L.ibfs_advance.argtypes = [C.c_void_p, _I, C.c_int, _D, _D]
    for nm in ("ibfs_op_dst", "ibfs_op_ctci", "ibfs_op_ainv", "ibfs_op_rot", "ibfs_op_reg", "ibfs_op_regT",
               "ibfs_op_a_times", "ibfs_op_cholsl"):
        getattr(L, nm).argtypes = [C.c_void_p, _D, _D]
    L.ibfs_op_curl.argtypes = [C.c_void_p, _D, _D, _D]
    L.ibfs_op_coarsify.argtypes = [C.c_void_p, _D, _D, _D]
    L.ibfs_op_get_bc.argtypes = [C.c_void_p, _D, C.c_double, _D]
    L.ibfs_op_rot_nonlinear.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
    L.ibfs_op_vort2flux.argtypes = [C.c_void_p, _D, _D, C.c_int]
    L.ibfs_last_advance_ms.argtypes = [C.c_void_p, _D]
    L.ibfs_time_dst_pass.argtypes = [C.c_void_p, C.c_int, C.c_int, _D]
    L.ibfs_time_dst_gemm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _D]
    L.ibfs_set_solver.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
    L.ibfs_cg_iterations.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.ibfs_move_body.argtypes = [C.c_void_p, _D, _D]
    L.ibfs_set_actuators.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), _D]
    L.ibfs_case_set_actuators.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), _D]
    L.ibfs_op_dst_gemm.argtypes = [C.c_void_p, C.c_int, _D, _D]
    L.ibfs_time_copy.argtypes = [C.c_void_p, C.c_long, C.c_int, _D]
    L.ibfs_time_step_kernel.argtypes = [C.c_void_p, C.c_int, _D, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.ibfs_tile_shapes.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.ibfs_level_tile_shapes.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.ibfs_last_error.restype = C.c_char_p
    L.ibfs_last_error.argtypes = [C.c_void_p]
    L.ibfs_destroy.argtypes = [C.c_void_p]
    _IP = C.POINTER(C.c_int)
    L.ibfs_ensemble_create.argtypes = [C.c_int, _IP, _IP, _IP, _IP, _D, _D, _D, _D, _D, _D, _D, _IP, C.c_int,
                                       C.c_int, C.POINTER(C.c_void_p)]
    L.ibfs_ensemble_info.argtypes = [C.c_void_p, _IP, _IP, _IP]
    L.ibfs_case_dims.argtypes = [C.c_void_p, C.c_int, _IP]
    L.ibfs_case_set_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D, _D]
    L.ibfs_case_get_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D, _D]
    L.ibfs_case_set_cholesky.argtypes = [C.c_void_p, C.c_int, _D, _D]
    L.ibfs_case_build_cholesky.argtypes = [C.c_void_p, C.c_int, _D, _D]
    L.ibfs_ensemble_build_cholesky.argtypes = [C.c_void_p]
    L.ibfs_objective.argtypes = [C.c_void_p, _IP, C.c_int, C.c_int, _D, C.c_double, C.c_double, _D]
    L.ibfs_tangent_linear.argtypes = [C.c_void_p, _IP, C.c_int, C.c_int, _D, _D, C.c_double, C.c_double, _D, _D]
    L.ibfs_adjoint_gradient.argtypes = [C.c_void_p, _IP, C.c_int, C.c_int, C.c_int, _D, C.c_double, C.c_double,
                                        _D, _D]
    L.ibfs_adjoint_forward.argtypes = [C.c_void_p, _IP, C.c_int, C.c_int, _D, _D]
    L.ibfs_adjoint_reverse.argtypes = [C.c_void_p, _D, _D]
    L.ibfs_step_vjp.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
    return L


_lib = _load()
EXPORTED = [n for n in dir(_lib) if n.startswith("ibfs_")]


def lib():
    return _lib


def _p(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_D)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Solver:
    """One case on the GPU (one handle, one HIP stream)."""

    def __init__(self, m, n, mgridlev, dt, re, len_, offx, offy, xb, yb, bdyno=None, unif_flow=True,
                 precision=PLAIN_FP64):
        xb = _f64(xb)
        yb = _f64(yb)
        nb = xb.size
        bd = None if bdyno is None else np.ascontiguousarray(bdyno, dtype=np.int32)
        h = C.c_void_p()
        rc = _lib.ibfs_create(m, n, mgridlev, nb, dt, re, len_, offx, offy, _p(xb), _p(yb),
                              None if bd is None else bd.ctypes.data_as(_I), 1 if unif_flow else 0, precision,
                              C.byref(h))
        self._h = h
        if rc:
            msg = _lib.ibfs_last_error(h).decode() if h.value else "invalid arguments"
            self.close()
            raise IBFSError(f"ibfs_create failed ({rc}): {msg}")
        d = (C.c_int * 12)()
        _lib.ibfs_dims(h, d)
        (self.m, self.n, self.L, self.nb, self.nf, self.nq, self.nw, self.nbc, self.nbody, self.bmin, self.bmax,
         self.nsmear) = list(d)
        self.dt = dt
        self.delta = len_ / float(np.float32(m))  # src/grid.f90:64 (REAL(m))
        self.itime = 0
        self.ncols = 2 * self.nbody + self.L + 1

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.ibfs_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise IBFSError(f"{what} failed ({rc}): {_lib.ibfs_last_error(self._h).decode()}")

    # ---- state ----
    def set_state(self, omega=None, rhs_old=None, q=None, q0=None, fb=None):
        a = [_f64(v) for v in (omega, rhs_old, q, q0, fb)]
        self._chk(_lib.ibfs_set_state(self._h, *[_p(v) for v in a]), "ibfs_set_state")

    def get_state(self):
        om = np.zeros(self.nw * self.L)
        ro = np.zeros(self.nw * self.L)
        q = np.zeros(self.nq * self.L)
        q0 = np.zeros(self.nq * self.L)
        fb = np.zeros(self.nf)
        self._chk(_lib.ibfs_get_state(self._h, _p(om), _p(ro), _p(q), _p(q0), _p(fb)), "ibfs_get_state")
        return {"omega": om, "rhs_old": ro, "q": q, "q0": q0, "fb": fb}

    def set_cholesky(self, cholmat, cholvec):
        self._chk(_lib.ibfs_set_cholesky(self._h, _p(_f64(cholmat)), _p(_f64(cholvec))), "ibfs_set_cholesky")

    def build_cholesky(self):
        cm = np.zeros(self.nf * self.nf)
        cv = np.zeros(self.nf)
        self._chk(_lib.ibfs_build_cholesky(self._h, _p(cm), _p(cv)), "ibfs_build_cholesky")
        return cm, cv

    # ---- time stepping ----
    def advance(self, nsteps=1, motion=None):
        """nsteps calls of advance; returns an (nsteps, 2*nbody+L+1) history of
        [fx.., fy.., circ(1..L), slip]."""
        it = C.c_int(self.itime)
        hist = np.zeros(nsteps * self.ncols)
        self._chk(_lib.ibfs_advance(self._h, C.byref(it), nsteps, _p(_f64(motion)), _p(hist)), "ibfs_advance")
        self.itime = it.value
        return hist.reshape(nsteps, self.ncols)

    def advance_timed(self, nsteps, motion=None, hist=False):
        """nsteps steps; returns device milliseconds measured with HIP events."""
        it = C.c_int(self.itime)
        h = np.zeros(nsteps * self.ncols) if hist else None
        self._chk(_lib.ibfs_advance(self._h, C.byref(it), nsteps, _p(_f64(motion)), _p(h)), "ibfs_advance")
        self.itime = it.value
        ms = C.c_double()
        _lib.ibfs_last_advance_ms(self._h, C.byref(ms))
        return ms.value

    def time_dst_pass(self, axis, reps=200):
        """Mean device time (ms) of one pass over a W field per case of the
        first group: axis 0 = DST-I along i, 1 = along j, 2 = the fused
        column pair of ctci (along j, times laminv, along j)."""
        ms = C.c_double()
        self._chk(_lib.ibfs_time_dst_pass(self._h, axis, reps, C.byref(ms)), "ibfs_time_dst_pass")
        return ms.value

    def set_solver(self, stationary, cgtol=1e-7, cg_max_iter=3000):
        """stationary: the Cholesky factor (default); otherwise cjgr
        (src/operators.f90:393-441) on the device, warm-started from fb."""
        self._chk(_lib.ibfs_set_solver(self._h, 1 if stationary else 0, float(cgtol), int(cg_max_iter)),
                  "ibfs_set_solver")

    def move_body(self, xb, yb):
        """New body point positions (the setup_reg after move_body,
        src/operators.f90:142-144); the factor of the old geometry is dropped."""
        self._chk(_lib.ibfs_move_body(self._h, _p(_f64(xb)), _p(_f64(yb))), "ibfs_move_body")

    def set_actuators(self, npt, params):
        """Actuator slots (actuation/u_jet, src/grid.f90:579-633): npt[a]
        points for slot a, the last sum(npt) points of the case in slot
        order; params[a] = (x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j), the
        slot's read_act_parameters namelist.  The jets are evaluated on the
        device every step and replace the motion argument at those points."""
        npt = np.ascontiguousarray(npt, dtype=np.int32).ravel()
        par = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        if par.size != 7 * npt.size:
            raise ValueError("params needs 7 values per actuator")
        self._chk(_lib.ibfs_set_actuators(self._h, int(npt.size), npt.ctypes.data_as(C.POINTER(C.c_int)),
                                          _p(par) if par.size else None), "ibfs_set_actuators")

    def cg_iterations(self, nsteps):
        """CG iterations of each step of the last advance call."""
        out = (C.c_int * max(nsteps, 1))()
        self._chk(_lib.ibfs_cg_iterations(self._h, int(nsteps), out), "ibfs_cg_iterations")
        return np.array(out[:nsteps])

    def time_dst_gemm(self, axis, ks=4, reps=200):
        """Mean device time (ms) of the dense matrix-core form of a DST-I pass
        (axis 0 = along i, 1 = along j), ks waves splitting the inner
        dimension; FP64 or FP32 as the handle's precision."""
        ms = C.c_double()
        self._chk(_lib.ibfs_time_dst_gemm(self._h, axis, ks, reps, C.byref(ms)), "ibfs_time_dst_gemm")
        return ms.value

    def op_dst_gemm(self, x, ks=4):
        """2-D DST-I of a W field through the two dense matrix-core passes (FP64)."""
        out = np.zeros(self.nw)
        self._chk(_lib.ibfs_op_dst_gemm(self._h, ks, _p(_f64(x)), _p(out)), "ibfs_op_dst_gemm")
        return out

    # ---- sensitivities (mean drag/lift over a window w.r.t. the surface velocity u_B) ----
    def _ncases(self):
        return getattr(self, "ncases", 1)

    def _nf_total(self):
        return sum(d["nf"] for d in self.dims) if hasattr(self, "dims") else self.nf

    def objective(self, nspin, nwin, ub=None, cd=1.0, cl=0.0):
        it = C.c_int(self.itime)
        J = np.zeros(self._ncases())
        self._chk(_lib.ibfs_objective(self._h, C.byref(it), nspin, nwin, _p(_f64(ub)), cd, cl, _p(J)),
                  "ibfs_objective")
        self.itime = it.value
        return J

    def tangent(self, nspin, nwin, ub, dub, cd=1.0, cl=0.0):
        it = C.c_int(self.itime)
        J = np.zeros(self._ncases())
        dJ = np.zeros(self._ncases())
        self._chk(_lib.ibfs_tangent_linear(self._h, C.byref(it), nspin, nwin, _p(_f64(ub)), _p(_f64(dub)), cd, cl,
                                           _p(J), _p(dJ)), "ibfs_tangent_linear")
        self.itime = it.value
        return J, dJ

    def adjoint(self, nspin, nwin, ckpt, ub, cd=1.0, cl=0.0):
        it = C.c_int(self.itime)
        J = np.zeros(self._ncases())
        g = np.zeros(self._nf_total())
        self._chk(_lib.ibfs_adjoint_gradient(self._h, C.byref(it), nspin, nwin, ckpt, _p(_f64(ub)), cd, cl, _p(J),
                                             _p(g)), "ibfs_adjoint_gradient")
        self.itime = it.value
        return J, g

    def adjoint_forward(self, nsteps, ckpt, ub=None):
        """Forward sweep of an adjoint run: nsteps steps with motion = delta *
        u_B, checkpoints every ckpt steps kept on the device for
        adjoint_reverse; returns the (nsteps, rowlen) history rows."""
        it = C.c_int(self.itime)
        rl = self.rowlen if hasattr(self, "rowlen") else self.ncols
        hist = np.zeros(nsteps * rl)
        self._chk(_lib.ibfs_adjoint_forward(self._h, C.byref(it), nsteps, ckpt, _p(_f64(ub)), _p(hist)),
                  "ibfs_adjoint_forward")
        self.itime = it.value
        return hist.reshape(nsteps, rl)

    def adjoint_reverse(self, ghist):
        """Reverse sweep of the last adjoint_forward: ghist = the cotangents of
        its history rows (force columns and circulations; slip entries 0).
        Returns dJ/du_B; the state stays at the forward sweep's end."""
        g = np.zeros(self._nf_total())
        self._chk(_lib.ibfs_adjoint_reverse(self._h, _p(_f64(np.ravel(ghist))), _p(g)), "ibfs_adjoint_reverse")
        return g

    def step_vjp(self, itime, omega, q, a_omega=None, a_rhs_old=None, a_q=None, a_row=None):
        """Vector-Jacobian product of the step that makes itime (from the state
        omega, q at its start; the factor of the handle; motion = 0 here, the
        product does not depend on it): the cotangents of the step's outputs
        (omega, rhs_old, q and its history row) give those of its inputs
        (omega, rhs_old, q) and of its motion argument."""
        W, Q = self.nw * self.L, self.nq * self.L
        out = {"omega": np.zeros(W), "rhs_old": np.zeros(W), "q": np.zeros(Q), "motion": np.zeros(self.nf)}
        self._chk(_lib.ibfs_step_vjp(self._h, int(itime), _p(_f64(omega)), _p(_f64(q)), _p(_f64(a_omega)),
                                     _p(_f64(a_rhs_old)), _p(_f64(a_q)), _p(_f64(a_row)), _p(out["omega"]),
                                     _p(out["rhs_old"]), _p(out["q"]), _p(out["motion"])), "ibfs_step_vjp")
        return out

    def reset(self):
        """Back to rest at itime 0 (zero vorticity, fluxes, forces)."""
        z = lambda n: np.zeros(n)
        if hasattr(self, "dims"):
            for c, d in enumerate(self.dims):
                self.set_state(c, omega=z(d["nw"] * d["L"]), rhs_old=z(d["nw"] * d["L"]), q=z(d["nq"] * d["L"]),
                               fb=z(d["nf"]))
        else:
            self.set_state(omega=z(self.nw * self.L), rhs_old=z(self.nw * self.L), q=z(self.nq * self.L),
                           fb=z(self.nf))
        self.itime = 0

    def ngroups(self):
        """Number of groups (cases of one grid size) of the handle."""
        ng = C.c_int()
        self._chk(_lib.ibfs_tile_shapes(self._h, 0, C.byref(ng), None), "ibfs_tile_shapes")
        return ng.value

    def tile_shapes(self, group=0):
        """FFTs per workgroup the tuner chose for one group of the handle:
        {kind: P} for the row pass, column pass, column pair, step-1 row
        pass and step-1 column pair (0: generic pass)."""
        out = (C.c_int * 5)()
        ng = C.c_int()
        self._chk(_lib.ibfs_tile_shapes(self._h, group, C.byref(ng), out), "ibfs_tile_shapes")
        return dict(zip(("row", "col", "col2", "step1_row", "step1_col2"), list(out)))

    def level_tile_shapes(self, group=0):
        """The shapes of the launches that transform every level at once
        (0: no such launch of that kind); 8 is the register-blocked kernel."""
        out = (C.c_int * 5)()
        self._chk(_lib.ibfs_level_tile_shapes(self._h, group, out), "ibfs_level_tile_shapes")
        return dict(zip(("row", "col", "col2", "step1_row", "step1_col2"), list(out)))

    def time_step_kernel(self, nsteps=20):
        """Mean in-step duration (us) of the step's dominant kernel (the fused
        ctci column pair), from events on its own dispatches during nsteps
        eager steps (the state advances), the number of launches and the
        number of W fields those launches transform (one per level and case
        of each launch)."""
        us = C.c_double()
        n = C.c_int()
        f = C.c_int()
        self._chk(_lib.ibfs_time_step_kernel(self._h, nsteps, C.byref(us), C.byref(n), C.byref(f)),
                  "ibfs_time_step_kernel")
        return us.value, n.value, f.value

    def time_copy(self, n, reps=200):
        """Mean device time (ms) of a streaming copy of n doubles."""
        ms = C.c_double()
        self._chk(_lib.ibfs_time_copy(self._h, int(n), reps, C.byref(ms)), "ibfs_time_copy")
        return ms.value

    def split(self, row):
        nb = self.nbody
        return row[:nb], row[nb:2 * nb], row[2 * nb:2 * nb + self.L], row[-1]

    # ---- operators ----
    def _op(self, name, *ins, nout):
        out = np.zeros(nout)
        self._chk(getattr(_lib, name)(self._h, *[_p(_f64(v)) for v in ins], _p(out)), name)
        return out

    def dst(self, x):
        return self._op("ibfs_op_dst", x, nout=self.nw)

    def ctci(self, x):
        return self._op("ibfs_op_ctci", x, nout=self.nw)

    def ainv(self, x):
        return self._op("ibfs_op_ainv", x, nout=self.nw)

    def rot(self, x):
        return self._op("ibfs_op_rot", x, nout=self.nw)

    def reg(self, h0):
        return self._op("ibfs_op_reg", h0, nout=self.nq)

    def regT(self, h):
        return self._op("ibfs_op_regT", h, nout=self.nf)

    def a_times(self, x):
        return self._op("ibfs_op_a_times", x, nout=self.nf)

    def cholsl(self, b):
        return self._op("ibfs_op_cholsl", b, nout=self.nf)

    def coarsify(self, crhs, rhs):
        return self._op("ibfs_op_coarsify", crhs, rhs, nout=self.nw)

    def rot_nonlinear(self, omega, q, q0, bc):
        return self._op("ibfs_op_rot_nonlinear", omega, q, q0, bc, nout=self.nw)

    def curl(self, x, coarse=None):
        out = np.zeros(self.nq)
        self._chk(_lib.ibfs_op_curl(self._h, _p(_f64(x)), _p(_f64(coarse)), _p(out)), "ibfs_op_curl")
        return out

    def get_bc(self, r, fac):
        out = np.zeros(self.nbc)
        self._chk(_lib.ibfs_op_get_bc(self._h, _p(_f64(r)), fac, _p(out)), "ibfs_op_get_bc")
        return out

    def vort2flux(self, omega, nlev):
        om = np.array(omega, dtype=np.float64, copy=True)
        vel = np.zeros(self.nq * max(nlev, 1))
        self._chk(_lib.ibfs_op_vort2flux(self._h, _p(om), _p(vel), nlev), "ibfs_op_vort2flux")
        return om, vel[: self.nq * nlev]


class Ensemble:
    """Independent cases advanced in lockstep by one handle (cases of equal
    grid size share kernel launches; one HIP graph per ensemble step).

    cases: list of dicts with the Solver keyword arguments (m, n, mgridlev,
    dt, re, len_, offx, offy, xb, yb, bdyno)."""

    def __init__(self, cases, unif_flow=True, precision=PLAIN_FP64):
        nc = len(cases)
        ia = lambda key: (C.c_int * nc)(*[int(c[key]) for c in cases])
        da = lambda key: (C.c_double * nc)(*[float(c[key]) for c in cases])
        nb = (C.c_int * nc)(*[len(c["xb"]) for c in cases])
        xb = np.ascontiguousarray(np.concatenate([np.asarray(c["xb"], float) for c in cases]))
        yb = np.ascontiguousarray(np.concatenate([np.asarray(c["yb"], float) for c in cases]))
        bd = np.ascontiguousarray(np.concatenate(
            [np.zeros(len(c["xb"]), np.int32) if c.get("bdyno") is None else np.asarray(c["bdyno"], np.int32)
             for c in cases]).astype(np.int32))
        h = C.c_void_p()
        rc = _lib.ibfs_ensemble_create(nc, ia("m"), ia("n"), ia("mgridlev"), nb, da("dt"), da("re"), da("len_"),
                                       da("offx"), da("offy"), _p(xb), _p(yb), bd.ctypes.data_as(_I),
                                       1 if unif_flow else 0, precision, C.byref(h))
        self._h = h
        if rc:
            msg = _lib.ibfs_last_error(h).decode() if h.value else "invalid arguments"
            self.close()
            raise IBFSError(f"ibfs_ensemble_create failed ({rc}): {msg}")
        self.ncases = nc
        rl = C.c_int()
        col0 = (C.c_int * nc)()
        _lib.ibfs_ensemble_info(h, None, C.byref(rl), col0)
        self.rowlen = rl.value
        self.col0 = list(col0)
        self.dims = []
        for c in range(nc):
            d = (C.c_int * 12)()
            _lib.ibfs_case_dims(h, c, d)
            self.dims.append(dict(zip(("m", "n", "L", "nb", "nf", "nq", "nw", "nbc", "nbody", "bmin", "bmax",
                                       "nsmear"), list(d))))
        self.itime = 0

    close = Solver.close
    __del__ = Solver.__del__
    _chk = Solver._chk
    tile_shapes = Solver.tile_shapes
    level_tile_shapes = Solver.level_tile_shapes
    ngroups = Solver.ngroups
    time_dst_pass = Solver.time_dst_pass
    _ncases = Solver._ncases
    _nf_total = Solver._nf_total
    objective = Solver.objective
    tangent = Solver.tangent
    adjoint = Solver.adjoint
    adjoint_forward = Solver.adjoint_forward
    adjoint_reverse = Solver.adjoint_reverse
    reset = Solver.reset

    def build_cholesky(self):
        self._chk(_lib.ibfs_ensemble_build_cholesky(self._h), "ibfs_ensemble_build_cholesky")

    def build_case_cholesky(self, c):
        d = self.dims[c]
        cm = np.zeros(d["nf"] * d["nf"])
        cv = np.zeros(d["nf"])
        self._chk(_lib.ibfs_case_build_cholesky(self._h, c, _p(cm), _p(cv)), "ibfs_case_build_cholesky")
        return cm, cv

    def set_case_cholesky(self, c, cholmat, cholvec):
        self._chk(_lib.ibfs_case_set_cholesky(self._h, c, _p(_f64(cholmat)), _p(_f64(cholvec))),
                  "ibfs_case_set_cholesky")

    def advance(self, nsteps=1, motion=None):
        it = C.c_int(self.itime)
        hist = np.zeros(nsteps * self.rowlen)
        self._chk(_lib.ibfs_advance(self._h, C.byref(it), nsteps, _p(_f64(motion)), _p(hist)), "ibfs_advance")
        self.itime = it.value
        return hist.reshape(nsteps, self.rowlen)

    def advance_timed(self, nsteps, motion=None):
        it = C.c_int(self.itime)
        self._chk(_lib.ibfs_advance(self._h, C.byref(it), nsteps, _p(_f64(motion)), None), "ibfs_advance")
        self.itime = it.value
        ms = C.c_double()
        _lib.ibfs_last_advance_ms(self._h, C.byref(ms))
        return ms.value

    def case_rows(self, hist, c):
        d = self.dims[c]
        ncol = 2 * d["nbody"] + d["L"] + 1
        return hist[:, self.col0[c]:self.col0[c] + ncol]

    def get_state(self, c):
        d = self.dims[c]
        W, Q = d["nw"] * d["L"], d["nq"] * d["L"]
        out = {"omega": np.zeros(W), "rhs_old": np.zeros(W), "q": np.zeros(Q), "q0": np.zeros(Q),
               "fb": np.zeros(d["nf"])}
        self._chk(_lib.ibfs_case_get_state(self._h, c, *[_p(out[k]) for k in ("omega", "rhs_old", "q", "q0", "fb")]),
                  "ibfs_case_get_state")
        return out

    def set_state(self, c, omega=None, rhs_old=None, q=None, q0=None, fb=None):
        a = [_f64(v) for v in (omega, rhs_old, q, q0, fb)]
        self._chk(_lib.ibfs_case_set_state(self._h, c, *[_p(v) for v in a]), "ibfs_case_set_state")
