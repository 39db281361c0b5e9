"""The Fortran 90 host (ibfs-if2d_amd/host) as a drop-in for the reference
program: same input/ibfs.inp namelist, same advance loop, same output files
(output/force.dat, circ.dat, slip.dat, restart ibfsNNNNNNN.var, input/ibfs.chd),
with the step done by the gfx950 library.  Its outputs are compared with the
oracle run of the same case, and a restart from a saved state continues the
run exactly."""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle as orc
from conftest import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ibfs-if2d_amd", "host", "ibfs_gpu")

NML = """ &READ_PARAMETERS ISTART = {istart}, ISTOP = {istop}, ISAVE = 5, M = 32,
 N = 32, DT = 0.02, RE = 100., CGTOL = 1.E-7, CG_MAX_ITER = 3000,
 UNIF_FLOW = T, LEN = 4.0, OFFSETX = 2.0, OFFSETY = 2.0, MGRIDLEV = 3,
 MAKE_VORTEX = F, STATIONARY = {stationary} /
"""


def run_host(workdir, istart, istop, nblock=1, stationary=True):
    os.makedirs(os.path.join(workdir, "input"), exist_ok=True)
    os.makedirs(os.path.join(workdir, "output"), exist_ok=True)
    with open(os.path.join(workdir, "input", "ibfs.inp"), "w") as f:
        f.write(NML.format(istart=istart, istop=istop, stationary="T" if stationary else "F"))
    with open(os.path.join(workdir, "input", "ibfs_gpu.inp"), "w") as f:
        f.write(f" &IBFS_GPU PRECISION = 0, NBLOCK = {nblock} /\n")
    r = subprocess.run([EXE], cwd=workdir, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def read_table(path, ncol=None):
    """A list-directed output file; with ncol, the values are taken in order
    and reshaped (list-directed records may wrap onto several lines)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(v.replace("D", "E")) for v in line.split()])
    if ncol is not None:
        return np.array([v for r in rows for v in r]).reshape(-1, ncol)
    return np.array(rows)


def read_record(path):
    """One Fortran unformatted sequential record of float64."""
    with open(path, "rb") as f:
        n = struct.unpack("<i", f.read(4))[0]
        data = np.frombuffer(f.read(n), dtype=np.float64)
        assert struct.unpack("<i", f.read(4))[0] == n
    return data


@pytest.fixture(scope="module")
def host_run(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (run __graft_entry__.build())")
    d = str(tmp_path_factory.mktemp("host"))
    r = run_host(d, 0, 10)
    return d, r


def oracle_run(d, nsteps):
    from ibfs_amd import geometry

    xb, yb, bd = geometry.read_geom(os.path.join(d, "input", "geom.inp"))
    o = orc.Oracle(32, 32, 3, 0.02, 100.0, 4.0, 2.0, 2.0, xb, yb, bd)
    chd = read_record(os.path.join(d, "input", "ibfs.chd"))
    nf = o.nf
    o.set_cholesky(chd[: nf * nf], chd[nf * nf:])
    rows = [o.advance() for _ in range(nsteps)]
    return o, rows


def test_host_outputs_match_oracle(host_run):
    d, _ = host_run
    force = read_table(os.path.join(d, "output", "force.dat"))
    circ = read_table(os.path.join(d, "output", "circ.dat"))
    slip = read_table(os.path.join(d, "output", "slip.dat"))
    assert force.shape == (10, 3) and circ.shape == (10, 4) and slip.shape == (10, 2)
    assert np.array_equal(force[:, 0], np.arange(1, 11))
    o, rows = oracle_run(d, 10)
    of = np.array([r[0] for r in rows])
    # list-directed output keeps ~15-17 significant digits
    assert np.max(np.abs(force[:, 1:] - of)) < 1e-11 * np.max(np.abs(of))
    oc = np.array([r[1] for r in rows])
    assert np.max(np.abs(circ[:, 1:] - oc)) < 1e-11 * np.sum(np.abs(o.omega))
    osl = np.array([r[2] for r in rows])
    assert np.max(np.abs(slip[:, 1] - osl)) < 1e-11 * o.delta
    # restart file at istop: omega, fb, rhs_old, q, q0 (src/variables.f90:152)
    var = read_record(os.path.join(d, "output", "ibfs0000010.var"))
    W, Q = o.nw * o.L, o.nq * o.L
    assert var.size == 2 * W + o.nf + 2 * Q
    assert relerr(var[:W], o.omega) < 1e-11
    assert relerr(var[W:W + o.nf], o.fb) < 1e-11
    assert relerr(var[W + o.nf:2 * W + o.nf], o.rhs_old) < 1e-10
    assert relerr(var[2 * W + o.nf:2 * W + o.nf + Q], o.q) < 1e-11


def test_host_restart_continues_the_run(host_run, tmp_path):
    d, _ = host_run
    d2 = str(tmp_path)
    shutil.copytree(os.path.join(d, "input"), os.path.join(d2, "input"))
    os.makedirs(os.path.join(d2, "output"))
    shutil.copy(os.path.join(d, "output", "ibfs0000005.var"), os.path.join(d2, "output"))
    run_host(d2, 5, 10)
    a = read_record(os.path.join(d, "output", "ibfs0000010.var"))
    b = read_record(os.path.join(d2, "output", "ibfs0000010.var"))
    assert np.array_equal(a, b)
    f1 = read_table(os.path.join(d, "output", "force.dat"))
    f2 = read_table(os.path.join(d2, "output", "force.dat"))
    assert np.array_equal(f1[5:], f2)


def test_host_blocked_calls_write_the_same_files(host_run, tmp_path):
    d, _ = host_run
    d3 = str(tmp_path)
    shutil.copytree(os.path.join(d, "input"), os.path.join(d3, "input"))
    run_host(d3, 0, 10, nblock=4)
    for fn in ("force.dat", "circ.dat", "slip.dat"):
        assert np.array_equal(read_table(os.path.join(d, "output", fn)), read_table(os.path.join(d3, "output", fn)))


@pytest.fixture(scope="module")
def host_cg_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("host_cg"))
    run_host(d, 0, 10, stationary=False)
    return d


def test_host_cg_form_matches_oracle(host_cg_run):
    """STATIONARY = F, the reference's default: each step's force system is
    solved by cjgr (src/operators.f90:169-171) on the device, with the
    namelist's cgtol and cg_max_iter; no factor and no ibfs.chd.  The CG
    path depends on rounding (A is slightly nonsymmetric), so the force
    history is held to 10x the spread of two oracle runs that differ only in
    the rounding of their transforms."""
    d = host_cg_run
    assert not os.path.exists(os.path.join(d, "input", "ibfs.chd"))
    with open(os.path.join(d, "output", "ibfs.inf")) as f:
        inf = f.read()
    assert inf.count("CJGR used") == 10
    from ibfs_amd import geometry

    xb, yb, bd = geometry.read_geom(os.path.join(d, "input", "geom.inp"))
    oa = orc.Oracle(32, 32, 3, 0.02, 100.0, 4.0, 2.0, 2.0, xb, yb, bd)
    ob = orc.Oracle(32, 32, 3, 0.02, 100.0, 4.0, 2.0, 2.0, xb, yb, bd)
    ob.set_dst_order(1)
    for o in (oa, ob):
        o.set_solver(False, cgtol=1e-7, cg_max_iter=3000)
    fa = np.array([oa.advance()[0] for _ in range(10)])
    fb_ = np.array([ob.advance()[0] for _ in range(10)])
    force = read_table(os.path.join(d, "output", "force.dat"))[:, 1:]
    floor = np.max(np.abs(fb_ - fa))
    assert np.max(np.abs(force - fa)) <= 10 * floor + 1e-12 * np.max(np.abs(fa))


def test_host_moving_body_without_motion(host_cg_run, tmp_path):
    """bdyno = 1 (a moving body): every step calls the host's move_body and
    rebuilds the smearing tables on the device (src/operators.f90:142-144).
    With no motion prescribed, as shipped, the run is the fixed-body CG run
    up to the CG tolerance: a moving body's CG applies a_times in every
    product, the fixed body's multiplies by A built once, so the iterates
    differ by rounding that CG carries to its stopping test (cgtol 1e-7)."""
    d = host_cg_run
    d2 = str(tmp_path)
    shutil.copytree(os.path.join(d, "input"), os.path.join(d2, "input"))
    gp = os.path.join(d2, "input", "geom.inp")
    with open(gp) as f:
        lines = f.read().split("\n")
    out = [lines[0]]
    for ln in lines[1:]:
        p = ln.split()
        if len(p) == 3:
            out.append(f"{p[0]} {p[1]} 1")
    with open(gp, "w") as f:
        f.write("\n".join(out) + "\n")
    run_host(d2, 0, 10, stationary=False)
    # forces on the force scale; the net circulation of this mirror-symmetric
    # case is rounding noise (~1e-16), compared absolutely; slip on delta
    delta = 4.0 / 32
    for fn, tol in (("force.dat", None), ("circ.dat", 1e-12), ("slip.dat", 1e-6 * delta)):
        a = read_table(os.path.join(d, "output", fn))
        b = read_table(os.path.join(d2, "output", fn))
        assert np.array_equal(a[:, 0], b[:, 0])
        bound = 1e-5 * np.max(np.abs(a[:, 1:])) if tol is None else tol
        assert np.max(np.abs(a[:, 1:] - b[:, 1:])) <= bound, fn


def test_host_rebuilds_a_stale_factor(host_run, tmp_path):
    """A cached input/ibfs.chd built for another case is not reused (the
    reference reuses whatever file is there, src/operators.f90:41-43): a run
    with another dt and the first run's factor and key rebuilds the factor
    and writes the same files as a run that starts without any."""
    d, _ = host_run
    assert os.path.exists(os.path.join(d, "input", "ibfs.chd.key"))
    stale, fresh = str(tmp_path / "stale"), str(tmp_path / "fresh")
    shutil.copytree(os.path.join(d, "input"), os.path.join(stale, "input"))
    shutil.copytree(os.path.join(d, "input"), os.path.join(fresh, "input"))
    for fn in ("ibfs.chd", "ibfs.chd.key"):
        os.remove(os.path.join(fresh, "input", fn))
    global NML
    nml0 = NML
    NML = NML.replace("DT = 0.02", "DT = 0.015")
    try:
        run_host(stale, 0, 4)
        run_host(fresh, 0, 4)
    finally:
        NML = nml0
    with open(os.path.join(stale, "output", "ibfs.inf")) as f:
        assert "built for another case" in f.read()
    assert np.array_equal(read_table(os.path.join(stale, "output", "force.dat")),
                          read_table(os.path.join(fresh, "output", "force.dat")))
    a = read_record(os.path.join(stale, "input", "ibfs.chd"))
    b = read_record(os.path.join(fresh, "input", "ibfs.chd"))
    assert np.array_equal(a, b)


def test_post_writes_the_tecplot_files(host_run, tmp_path):
    """ibfs_post (the flow of src/post.f90) reads the saved restart files and
    writes write_tecplot's zone per level (src/operators.f90:702-763):
    coordinates, vertex velocity of q + q0 over the level spacing, the
    finest level's reg(fb) force density and vorticity over the spacing
    squared.  Checked against the same quantities from the restart record
    and the oracle's reg, to the 5 significant digits of the E14.5 format."""
    d, _ = host_run
    d2 = str(tmp_path)
    shutil.copytree(os.path.join(d, "input"), os.path.join(d2, "input"))
    os.makedirs(os.path.join(d2, "output"))
    os.makedirs(os.path.join(d2, "post"))
    for it in (5, 10):
        shutil.copy(os.path.join(d, "output", f"ibfs{it:07d}.var"), os.path.join(d2, "output"))
    with open(os.path.join(d2, "input", "ibfs.inp"), "w") as f:
        f.write(NML.format(istart=5, istop=10, stationary="T"))
    r = subprocess.run([os.path.join(ROOT, "ibfs-if2d_amd", "host", "ibfs_post")], cwd=d2, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    from ibfs_amd import geometry

    xb, yb, bd = geometry.read_geom(os.path.join(d2, "input", "geom.inp"))
    o = orc.Oracle(32, 32, 3, 0.02, 100.0, 4.0, 2.0, 2.0, xb, yb, bd)
    m = n = 32
    W, Q = o.nw * o.L, o.nq * o.L
    var = read_record(os.path.join(d2, "output", "ibfs0000010.var"))
    om, fb = var[:W], var[W:W + o.nf]
    q, q0 = var[2 * W + o.nf:2 * W + o.nf + Q], var[2 * W + o.nf + Q:]
    etf = o.reg(fb)
    iu = lambda i, j: (i - 1) + (j - 1) * (m + 1)
    iv = lambda i, j: (m + 1) * n + (i - 1) + (j - 1) * m
    for lev in (1, 2, 3):
        fn = os.path.join(d2, "post", f"ibfs_g{lev}t0000010_vel.dat")
        with open(fn) as f:
            lines = f.read().splitlines()
        assert lines[0].strip().startswith("TITLE") and "DATAPACKING=POINT" in lines[10]  # 12 header lines
        # Ew.d drops the E of three-digit exponents (0.12345-100)
        num = lambda v: float(re.sub(r"(\d)([+-]\d{3})$", r"\1E\2", v))
        data = np.array([[num(v) for v in ln.split()] for ln in lines[12:] if ln.strip()])
        assert data.shape == ((m - 1) * (n - 1), 7)
        fac = 2.0 ** (lev - 1)
        dl = fac * o.delta
        ql, q0l = q[(lev - 1) * o.nq:lev * o.nq], q0[(lev - 1) * o.nq:lev * o.nq]
        exp = []
        for j in range(2, n + 1):
            for i in range(2, m + 1):
                exp.append([fac * o.delta * (i - 1) - (fac - 1) / 2 * 4.0 - 2.0,
                            fac * o.delta * (j - 1) - (fac - 1) / 2 * o.delta * n - 2.0,
                            0.5 * (ql[iu(i, j)] + ql[iu(i, j - 1)] + q0l[iu(i, j)] + q0l[iu(i, j - 1)]) / dl,
                            0.5 * (ql[iv(i, j)] + ql[iv(i - 1, j)] + q0l[iv(i, j)] + q0l[iv(i - 1, j)]) / dl,
                            o.delta * (etf[iu(i, j)] + etf[iu(i, j - 1)]) / 0.02,
                            o.delta * (etf[iv(i, j)] + etf[iv(i - 1, j)]) / 0.02,
                            om[(lev - 1) * o.nw + (i - 2) + (j - 2) * (m - 1)] / dl / dl])
        exp = np.array(exp)
        for c in range(7):
            # E14.5 keeps 5 significant digits: half a unit of the 5th is 5e-5 relative
            scale = np.max(np.abs(exp[:, c]))
            assert np.all(np.abs(data[:, c] - exp[:, c]) <= 5.01e-5 * np.abs(exp[:, c]) + 1e-9 * scale), (lev, c)


ACT = [" &READ_ACT_PARAMETERS X_J = 0.9, Y_J = 0.0, H_J = 0.3, U_J = 0.5, U_J_OSC = 0.25, THETA_J = 0.0,"
       " OMG_J = 3.0 /\n",
       " &READ_ACT_PARAMETERS X_J = 0.0, Y_J = 0.9, H_J = 0.1, U_J = 0.0, U_J_OSC = 0.8, THETA_J = 90.0,"
       " OMG_J = 5.0 /\n"]


def test_host_actuators_match_oracle(tmp_path):
    """Two input/actuator.NNN.inp files and no geom.inp: the host appends
    the slots to the default circle with body numbers -1, -2
    (src/grid.f90:383-392, 636-718), writes geom.inp and slot.NNN.inp, and
    the library drives the jets every step.  force.dat gets the columns of
    bodies -2..0; every output matches the oracle with actuation's motion.
    A second run reads the slots back from geom.inp and slot.NNN.inp
    (src/grid.f90:286-300) and gives the same files."""
    d = str(tmp_path / "a")
    os.makedirs(os.path.join(d, "input"))
    for k, txt in enumerate(ACT):
        with open(os.path.join(d, "input", f"actuator.{k + 1:03d}.inp"), "w") as f:
            f.write(txt)
    run_host(d, 0, 10)
    from ibfs_amd import geometry

    xb, yb, bd = geometry.read_geom(os.path.join(d, "input", "geom.inp"))
    assert list(np.unique(bd)) == [-2, -1, 0]
    npt = [int(open(os.path.join(d, "input", f"slot.{k:03d}.inp")).read().split()[0]) for k in (1, 2)]
    assert npt == [3, 1] and np.count_nonzero(bd == -1) == 3 and np.count_nonzero(bd == -2) == 1
    cx, cy, _ = geometry.circle(4.0 / 32, 1.0, 0.0, 0.0)
    sx, sy, _ = geometry.actuator_slot(0.9, 0.0, 0.3, 4.0 / 32, 0.0, -1)
    assert np.allclose(xb, np.concatenate([cx, sx, [0.0]]), rtol=0, atol=1e-15)
    assert np.allclose(yb, np.concatenate([cy, sy, [0.9]]), rtol=0, atol=1e-15)
    force = read_table(os.path.join(d, "output", "force.dat"), 7)
    assert force.shape == (10, 7) and np.array_equal(force[:, 0], np.arange(1, 11))
    params = [(0.9, 0.0, 0.3, 0.5, 0.25, 0.0, 3.0), (0.0, 0.9, 0.1, 0.0, 0.8, 90.0, 5.0)]
    o = orc.Oracle(32, 32, 3, 0.02, 100.0, 4.0, 2.0, 2.0, xb, yb, bd)
    chd = read_record(os.path.join(d, "input", "ibfs.chd"))
    o.set_cholesky(chd[: o.nf * o.nf], chd[o.nf * o.nf:])
    of = []
    for it in range(1, 11):
        f, c, s = o.advance(motion=orc.actuation(it, o.nb, o.delta, 0.02, npt, params))
        of.append(f)
    of = np.array(of)
    for b in range(3):
        cols = [b, 3 + b]
        assert np.max(np.abs(force[:, 1:][:, cols] - of[:, cols])) < 1e-11 * np.max(np.abs(of[:, cols])), b
    # a second run from the written geometry: slots counted from slot.NNN.inp
    d2 = str(tmp_path / "b")
    shutil.copytree(os.path.join(d, "input"), os.path.join(d2, "input"))
    run_host(d2, 0, 10)
    assert np.array_equal(read_table(os.path.join(d2, "output", "force.dat"), 7), force)


def test_host_stops_on_a_missing_restart_file(tmp_path):
    """istart = 5 with no output/ibfs0000005.var: the host stops with a
    message naming the file (the reference would hit end-of-file)."""
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "input"))
    os.makedirs(os.path.join(d, "output"))
    with open(os.path.join(d, "input", "ibfs.inp"), "w") as f:
        f.write(NML.format(istart=5, istop=10, stationary="T"))
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 or "restart file not found" in (r.stdout + r.stderr)
    assert "ibfs0000005.var not found" in r.stdout


def test_host_stops_on_a_bad_cached_factor(host_run, tmp_path):
    """A cached input/ibfs.chd whose factor has a non-positive pivot (a
    non-SPD matrix) is refused by the library (IBFS_ECHOL) and the host stops
    with the library's message, as the reference stops in choldc."""
    d, _ = host_run
    d2 = str(tmp_path)
    shutil.copytree(os.path.join(d, "input"), os.path.join(d2, "input"))
    os.makedirs(os.path.join(d2, "output"))
    path = os.path.join(d2, "input", "ibfs.chd")
    raw = bytearray(open(path, "rb").read())
    n = struct.unpack("<i", raw[:4])[0] // 8
    nf = int((-1 + (1 + 4 * n) ** 0.5) / 2)
    struct.pack_into("<d", raw, 4 + 8 * (nf * nf + 3), -1.0)  # cholvec(4) < 0
    open(path, "wb").write(bytes(raw))
    with open(os.path.join(d2, "input", "ibfs.inp"), "w") as f:
        f.write(NML.format(istart=0, istop=2, stationary="T"))
    r = subprocess.run([EXE], cwd=d2, capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert "ibfs_set_cholesky" in out and "code" in out and "non-positive" in out
