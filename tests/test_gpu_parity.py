"""Parity of the gfx950 library (through its C-ABI) with the CPU oracle.

FP64 plain form: every operator and the whole step must agree with the
oracle to 1e-11 relative (the bound the north-star contract states for
fields and force histories in FP64); the step-level comparisons use
max|x - x_ref| / max|x_ref| per level / per force component."""
import numpy as np
import pytest

import oracle as orc
from conftest import CASES, case_args, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-11


def pair(name):
    from ibfs_amd import hip

    a = case_args(name)
    o = orc.Oracle(**a)
    g = hip.Solver(**a)
    assert (g.m, g.n, g.L, g.nb, g.nf, g.nq, g.nw) == (o.m, o.n, o.L, o.nb, o.nf, o.nq, o.nw)
    return o, g


@pytest.fixture(scope="module", params=list(CASES))
def pg(request):
    return pair(request.param)


def test_dst_ctci_ainv(pg, rng):
    o, g = pg
    x = rng.standard_normal(o.nw)
    assert relerr(g.dst(x), o.dst(x)) < 1e-13
    assert relerr(g.ctci(x), o.ctci(x)) < 1e-12
    assert relerr(g.ainv(x), o.ainv(x)) < 1e-12


def test_dense_dst_matches_oracle(pg, rng):
    """The matrix-core form of the DST-I passes (measured against the FFT
    passes, not used by the step) gives the same 2-D transform; the cases
    include m != n, so a transposed operand or swapped pass would fail."""
    o, g = pg
    x = rng.standard_normal(o.nw)
    ref = o.dst(x)
    for ks in (1, 2, 4, 8):
        assert relerr(g.op_dst_gemm(x, ks), ref) < 1e-13, f"ks={ks}"


def test_stencil_operators(pg, rng):
    o, g = pg
    s = rng.standard_normal(o.nw)
    c = rng.standard_normal(o.nw)
    assert relerr(g.curl(s), o.curl(s, np.zeros(o.nbc))) == 0.0
    assert relerr(g.curl(s, c), o.curl(s, o.get_bc(c, 1.0))) < 1e-15
    x = rng.standard_normal(o.nq)
    assert relerr(g.rot(x), o.rot(x)) == 0.0
    assert relerr(g.coarsify(c, s), o.coarsify(c, s)) == 0.0
    for fac in (0.25, 1.0):
        assert relerr(g.get_bc(s, fac), o.get_bc(s, fac)) == 0.0
    q = rng.standard_normal(o.nq)
    q0 = rng.standard_normal(o.nq)
    bc = o.get_bc(c, 0.25)
    ref = o.rot(o.nonlinear(s, q, q0, bc))
    assert relerr(g.rot_nonlinear(s, q, q0, bc), ref) < 1e-15


def test_smearing(pg, rng):
    o, g = pg
    assert g.nsmear == o.nsmear
    f = rng.standard_normal(o.nf)
    h = rng.standard_normal(o.nq)
    assert relerr(g.reg(f), o.reg(f)) < 1e-15
    assert relerr(g.regT(h), o.regT(h)) < 1e-15


def test_vort2flux(pg, rng):
    o, g = pg
    om = rng.standard_normal(o.nw * o.L)
    for nlev in (1, o.L):
        o_om, o_vel = o.vort2flux(om, nlev)
        g_om, g_vel = g.vort2flux(om, nlev)
        assert relerr(g_om, o_om) == 0.0  # coarsify is exact arithmetic reordering-free
        for k in range(nlev):
            sl = slice(k * o.nq, (k + 1) * o.nq)
            assert relerr(g_vel[sl], o_vel[sl]) < TOL


def test_a_times_and_factor(pg, rng):
    o, g = pg
    v = rng.standard_normal(o.nf)
    assert relerr(g.a_times(v), o.a_times(v)) < TOL
    cm, cv = g.build_cholesky()
    A = np.stack([o.a_column(i) for i in range(o.nf)], axis=1)
    cmr = cm.reshape(o.nf, o.nf).T
    assert relerr(np.triu(cmr), np.triu(A)) < TOL
    o.cholmat[:] = A.T.ravel()
    assert o.choldc() == 0
    assert relerr(cv, o.cholvec) < TOL
    assert relerr(np.tril(cmr, -1), np.tril(o.cholmat.reshape(o.nf, o.nf).T, -1)) < TOL
    b = rng.standard_normal(o.nf)
    assert relerr(g.cholsl(b), o.cholsl(b)) < 1e-10


@pytest.mark.parametrize("name", list(CASES))
def test_advance_matches_oracle(name):
    """20 steps from rest (the first one the explicit Euler start): vorticity,
    rhs_old and fluxes per level, fb, and the force/circulation/slip
    histories."""
    o, g = pair(name)
    cm, cv = g.build_cholesky()
    o.set_cholesky(cm, cv)  # the same factor on both sides; the factor itself is checked above
    nsteps = 20
    oh = []
    for _ in range(nsteps):
        f, c, s = o.advance()
        oh.append(np.concatenate([f, c, [s]]))
    oh = np.array(oh)
    gh = g.advance(nsteps)
    st = g.get_state()
    for k in range(o.L):
        w = slice(k * o.nw, (k + 1) * o.nw)
        qq = slice(k * o.nq, (k + 1) * o.nq)
        assert relerr(st["omega"][w], o.omega[w]) < TOL, f"omega level {k + 1}"
        # rhs_old is the nonlinear term: differences of products of q and
        # omega, so it carries ~10x the relative rounding of the fields
        assert relerr(st["rhs_old"][w], o.rhs_old[w]) < 10 * TOL, f"rhs_old level {k + 1}"
        assert relerr(st["q"][qq], o.q[qq]) < TOL, f"q level {k + 1}"
    assert relerr(st["fb"], o.fb) < TOL
    # Force histories are normalised by the force scale of the body (max over
    # both components and all steps) and circulations by the vorticity scale
    # sum|omega_k|: in mirror-symmetric cases the lift and the net
    # circulation are rounding noise, so a per-column relative error is
    # meaningless there.
    nbd = o.nbody
    for b in range(nbd):
        cols = [b, nbd + b]
        scale = np.max(np.abs(oh[:, cols]))
        assert np.max(np.abs(gh[:, cols] - oh[:, cols])) < TOL * scale, f"force of body {b}"
    for k in range(o.L):
        scale = np.sum(np.abs(o.omega[k * o.nw:(k + 1) * o.nw]))
        col = 2 * nbd + k
        assert np.max(np.abs(gh[:, col] - oh[:, col])) < TOL * scale, f"circulation level {k + 1}"
    # slip is a small difference of O(delta) terms: compare on the delta scale
    assert np.max(np.abs(gh[:, -1] - oh[:, -1])) < TOL * o.delta


@pytest.mark.parametrize("name", ["cyl32", "naca48x40"])
def test_advance_split_calls_equal_one_call(name):
    """Graph replay across calls: 3 + 5 + 1 steps == 9 steps in one call, bit for bit."""
    from ibfs_amd import hip

    a = case_args(name)
    g1 = hip.Solver(**a)
    g2 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    g2.set_cholesky(cm, cv)
    h1 = np.concatenate([g1.advance(3), g1.advance(5), g1.advance(1)])
    h2 = g2.advance(9)
    assert np.array_equal(h1, h2)
    s1, s2 = g1.get_state(), g2.get_state()
    for key in ("omega", "q", "rhs_old", "fb"):
        assert np.array_equal(s1[key], s2[key]), key


def test_general_q0_matches_oracle():
    """A q0 that is not in the freestream form (the step-1 loader then reads
    the q0 array rather than the per-level freestream value): 10 steps
    against the oracle, then the freestream q0 restored on both sides and 5
    more steps."""
    o, g = pair("naca48x40")
    cm, cv = g.build_cholesky()
    o.set_cholesky(cm, cv)
    q0_free = o.q0.copy()
    q0 = q0_free + 0.1 * np.abs(q0_free).max() * np.random.default_rng(3).standard_normal(q0_free.size)
    o.q0[:] = q0
    g.set_state(q0=q0)

    def compare(nsteps):
        oh = np.array([np.concatenate([f, c, [s]]) for f, c, s in (o.advance() for _ in range(nsteps))])
        gh = g.advance(nsteps)
        st = g.get_state()
        for k in range(o.L):
            w = slice(k * o.nw, (k + 1) * o.nw)
            qq = slice(k * o.nq, (k + 1) * o.nq)
            assert relerr(st["omega"][w], o.omega[w]) < TOL, f"omega level {k + 1}"
            assert relerr(st["q"][qq], o.q[qq]) < TOL, f"q level {k + 1}"
        assert relerr(st["fb"], o.fb) < TOL
        nb2 = 2 * o.nbody
        assert np.max(np.abs(gh[:, :nb2] - oh[:, :nb2])) < TOL * np.max(np.abs(oh[:, :nb2]))

    compare(10)
    o.q0[:] = q0_free
    g.set_state(q0=q0_free)
    compare(5)


def test_set_state_round_trip(rng):
    from ibfs_amd import hip

    a = case_args("cyl32")
    g = hip.Solver(**a)
    st = {k: rng.standard_normal(v.size) for k, v in g.get_state().items()}
    g.set_state(**st)
    back = g.get_state()
    for k in st:
        assert np.array_equal(back[k], st[k]), k


def test_errors_are_reported():
    from ibfs_amd import hip

    a = case_args("cyl32")
    a["m"] = 30  # not a multiple of 4 (src/grid.f90:67-69)
    with pytest.raises(hip.IBFSError):
        hip.Solver(**a)
    a = case_args("cyl32")
    g = hip.Solver(**a)
    with pytest.raises(hip.IBFSError, match="Cholesky"):
        g.advance(1)  # the factor is required before the first step
    # a factor with a non-positive pivot (what choldc reports for a matrix
    # that is not SPD, src/operators.f90:782-784) is refused: IBFS_ECHOL
    cm, cv = g.build_cholesky()
    bad = cv.copy()
    bad[3] = 0.0
    with pytest.raises(hip.IBFSError, match=r"\(3\).*non-positive"):
        g.set_cholesky(cm, bad)
    # a non-finite motion entry is rejected before any step runs: IBFS_EINVAL
    g.set_cholesky(cm, cv)
    mo = np.zeros(g.nf)
    mo[5] = np.nan
    with pytest.raises(hip.IBFSError, match=r"\(1\).*non-finite"):
        g.advance(2, motion=mo)
    assert g.itime == 0
    h = g.advance(1)  # the handle still steps
    assert np.all(np.isfinite(h))


def levels_batch(a):
    """Whether the library batches step 1 over the levels of this case: every
    level k >= 1 damps an edge source below 1e-17 within min(m,n)/4 cells,
    kappa_k min(m,n)/4 >= ln 1e17 with kappa = acosh(1 + 1/(2 beta)),
    beta = dt / (2 Re Delta_k^2), and the band of w_k = ceil(ln 1e17 / kappa_k)
    cells each level with a coarser level carries stays under a quarter of the
    grid."""
    m, n, L = a["m"], a["n"], a["mgridlev"]
    need, q = np.log(1e17), min(m, n) // 4
    if L < 2:
        return False
    for k in range(L):
        dk = a["len_"] / m * 2.0**k
        kappa = np.arccosh(1.0 + 1.0 / (2.0 * (0.5 * a["dt"] / (a["re"] * dk * dk))))
        if k >= 1 and kappa * q < need:
            return False
        if k + 1 < L and 4 * int(np.ceil(need / kappa)) >= min(m, n):
            return False
    return True


@pytest.mark.parametrize("name", ["cyl32", "naca48x40", "cyl40r5"])
def test_level_batched_step1_equals_level_by_level(name, monkeypatch):
    """Step 1 with the levels batched (the bulk solve of every level, then
    each level's correction for the new coarser edge values on the bands next
    to its edges) against the reference's level-by-level order
    (IBFS_STEP1_LEVELS=0): the same step to rounding, over 20 steps.  cyl32's
    finest band would reach a quarter of its 32-cell grid, so it keeps the
    level-by-level order and the two runs are bit-identical."""
    from ibfs_amd import hip

    a = case_args(name)
    batched = levels_batch(a)
    assert batched == (name != "cyl32")
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_STEP1_LEVELS", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(20), g0.advance(20)
    s1, s0 = g1.get_state(), g0.get_state()
    nb2 = 2 * g1.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-12 * np.max(np.abs(h0[:, :nb2]))
    for k in range(g1.L):
        w = slice(k * g1.nw, (k + 1) * g1.nw)
        assert relerr(s1["omega"][w], s0["omega"][w]) < 1e-12, f"omega level {k + 1}"
    # the batched form did run where it applies (the two forms round differently)
    assert np.array_equal(s1["omega"], s0["omega"]) != batched


# one-case grids with three to five levels (the batched vort2flux applies
# two or three powers of its ring map there), square and oblong
V2F_GRIDS = {
    "cyl64L4": dict(m=64, n=64, mgridlev=4, dt=0.01, re=100.0, len_=4.0, offx=2.0, offy=2.0),
    "rect64x48L5": dict(m=64, n=48, mgridlev=5, dt=0.01, re=150.0, len_=4.0, offx=2.0, offy=1.5),
}


def v2f_args(name):
    if name in CASES:
        return case_args(name)
    from ibfs_amd import geometry

    a = dict(V2F_GRIDS[name])
    xb, yb, bd = geometry.circle(a["len_"] / a["m"], 1.0, 0.0, 0.0)
    return dict(a, xb=xb, yb=yb, bdyno=bd)


@pytest.mark.parametrize("name", ["cyl32", "naca48x40", "cyl64L4", "rect64x48L5"])
def test_level_batched_vort2flux_equals_level_by_level(name, monkeypatch):
    """vort2flux with every level's ctci in one batch, the edge values of the
    levels from the ring vectors R_k = sum_j T^j ring(ctci(omega_{k+j})) and
    the edge terms added in closed form, against the reference's coarse-to-
    fine order (IBFS_V2F_LEVELS=0): the same steps to rounding over 20 steps,
    forces, circulation, vorticity and fluxes on every level."""
    from ibfs_amd import hip

    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_V2F_LEVELS", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(20), g0.advance(20)
    s1, s0 = g1.get_state(), g0.get_state()
    nb2 = 2 * g1.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-12 * np.max(np.abs(h0[:, :nb2]))
    # circulation: against the sum of |omega| dx^2 (a symmetric wake's is ~0)
    circ = slice(nb2, nb2 + g1.L)
    cscale = max(np.sum(np.abs(s0["omega"][k * g1.nw:(k + 1) * g1.nw])) * (a["len_"] / a["m"] * 2**k) ** 2
                 for k in range(g1.L))
    assert np.max(np.abs(h1[:, circ] - h0[:, circ])) < 1e-12 * cscale
    for k in range(g1.L):
        w = slice(k * g1.nw, (k + 1) * g1.nw)
        assert relerr(s1["omega"][w], s0["omega"][w]) < 1e-12, f"omega level {k + 1}"
        v = slice(k * g1.nq, (k + 1) * g1.nq)
        assert relerr(s1["q"][v], s0["q"][v]) < 1e-12, f"q level {k + 1}"
    assert not np.array_equal(s1["q"], s0["q"])  # the two forms did run (they round differently)


@pytest.mark.parametrize("name", ["naca48x40", "cyl40r5"])
def test_band_correction_real_space_equals_spectral(name, monkeypatch):
    """The band correction of the batched step 1 as short sums with the
    tabulated real-space kernels H_d (one launch, the default) against its
    spectral form, a fused DST-I column pair per band depth and pair of
    opposite edges (IBFS_BAND_FFT=1): the same steps to rounding over 20
    steps (the dropped tail of H_d is below 1e-17 of the edge values)."""
    from ibfs_amd import hip

    a = case_args(name)
    assert levels_batch(a)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_BAND_FFT", "1")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(20), g0.advance(20)
    s1, s0 = g1.get_state(), g0.get_state()
    nb2 = 2 * g1.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-13 * np.max(np.abs(h0[:, :nb2]))
    for k in range(g1.L):
        w = slice(k * g1.nw, (k + 1) * g1.nw)
        assert relerr(s1["omega"][w], s0["omega"][w]) < 1e-13, f"omega level {k + 1}"
    assert not np.array_equal(s1["omega"], s0["omega"])  # the two forms did run


@pytest.mark.parametrize("name", ["cyl40r5", "cyl64L4", "rect64x48L5"])
def test_step3_real_space_equals_spectral(name, monkeypatch):
    """Step 3 (omega_1 -= ainv(rot(reg(fb)))) as a short sum with the
    tabulated real-space kernel of ainv around the active vertices (the
    default where they keep the kernel's reach from the edges) against the
    spectral ctci (IBFS_STEP3_LOCAL=0): the same 20 steps to rounding (the
    kernel's dropped tail is below 1e-17 of a unit source)."""
    from ibfs_amd import hip

    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_STEP3_LOCAL", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(20), g0.advance(20)
    s1, s0 = g1.get_state(), g0.get_state()
    nb2 = 2 * g1.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-13 * np.max(np.abs(h0[:, :nb2]))
    for k in range(g1.L):
        w = slice(k * g1.nw, (k + 1) * g1.nw)
        assert relerr(s1["omega"][w], s0["omega"][w]) < 1e-13, f"omega level {k + 1}"
        v = slice(k * g1.nq, (k + 1) * g1.nq)
        assert relerr(s1["q"][v], s0["q"][v]) < 1e-13, f"q level {k + 1}"
    assert relerr(s1["fb"], s0["fb"]) < 1e-12
    assert not np.array_equal(s1["omega"], s0["omega"])  # the real-space form did run


@pytest.mark.parametrize("name", ["rect64x48L5", "cyl64L4"])
def test_split_streamfunction_is_bit_identical(name, monkeypatch):
    """Step 4's batched vort2flux keeps the left/right edge terms apart from
    the streamfunction (the last row and column passes independent, in one
    launch where both are the same size-specialised pass) and the tail adds
    them, before get_bc's averages at the edges too: the same sums as the
    column pass adding into the streamfunction (IBFS_SPLIT_S=0), so the
    states and histories after 10 steps are equal bit for bit (square and
    oblong grids)."""
    from ibfs_amd import hip

    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_SPLIT_S", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(10), g0.advance(10)
    s1, s0 = g1.get_state(), g0.get_state()
    assert np.array_equal(h1, h0)
    for k in s0:
        assert np.array_equal(s1[k], s0[k]), k


@pytest.mark.parametrize("name", ["rect64x48L5", "cyl64L4"])
def test_ring_chain_on_symmetry_representatives(name, monkeypatch):
    """The ring map T commutes with the flips of the grid (and, on square
    grids, its transpose), so the ring chain reads T's powers on one row per
    symmetry orbit and permutes the vectors instead (the default) or reads
    every row (IBFS_RING_SYM=0).  The built T is symmetric only to the
    rounding of its GEMMs, so the two agree to rounding over 10 steps and
    are not identical (the symmetric path was taken: four symmetries on the
    oblong grid, eight on the square one)."""
    from ibfs_amd import hip

    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_RING_SYM", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(10), g0.advance(10)
    s1, s0 = g1.get_state(), g0.get_state()
    nb2 = 2 * g1.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-12 * np.max(np.abs(h0[:, :nb2]))
    for k in ("omega", "q"):
        assert relerr(s1[k], s0[k]) < 1e-13, k
    assert relerr(s1["fb"], s0["fb"]) < 1e-12
    assert not np.array_equal(s1["omega"], s0["omega"])


@pytest.mark.parametrize("name", ["rect64x48L5", "cyl64L4"])
def test_fp64_step2_on_the_rings(name, monkeypatch):
    """Step 2 of the level-by-level FP64 vort2flux (IBFS_V2F_LEVELS=0, the
    form of batches) computes the levels below the finest only on their
    rings, by the ring lines of the column pair's output (the default), or
    whole (IBFS_STEP2_RING=0): the same values to rounding, so 10 steps agree
    to rounding and the force solution differs (the ring form was taken)."""
    from ibfs_amd import hip

    monkeypatch.setenv("IBFS_V2F_LEVELS", "0")
    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_STEP2_RING", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(10), g0.advance(10)
    s1, s0 = g1.get_state(), g0.get_state()
    nb2 = 2 * g1.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-12 * np.max(np.abs(h0[:, :nb2]))
    for k in ("omega", "q"):
        assert relerr(s1[k], s0[k]) < 1e-13, k
    assert relerr(s1["fb"], s0["fb"]) < 1e-12
    assert not np.array_equal(s1["fb"], s0["fb"])


@pytest.mark.parametrize("name", ["rect64x48L5", "cyl64L4", "cyl40r5"])
def test_coarsify_tiles_are_bit_identical(name, monkeypatch):
    """The coarsify chain two levels per launch in LDS tiles (the default)
    against one place per thread (IBFS_COARSIFY_TILES=0): both evaluate
    every coarsified value with the same expression from the same inputs, so
    the states and histories after 10 steps are equal bit for bit (five
    levels: two tiled launches; four: one tiled launch and the last level on
    its own; three: one tiled launch)."""
    from ibfs_amd import hip

    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_COARSIFY_TILES", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(10), g0.advance(10)
    s1, s0 = g1.get_state(), g0.get_state()
    assert np.array_equal(h1, h0)
    for k in s0:
        assert np.array_equal(s1[k], s0[k]), k


@pytest.mark.parametrize("name", ["naca48x40", "cyl64L4"])
def test_step2_window_is_bit_identical(name, monkeypatch):
    """Step 2 transforms level 0's streamfunction only on the rows and
    columns of the rectangle that holds every vertex the force stencils read
    (the row and column transforms there are the full pass's): the state and
    the history after 10 steps equal those of the whole-field passes
    (IBFS_S_WINDOW=0) bit for bit."""
    from ibfs_amd import hip

    a = v2f_args(name)
    g1 = hip.Solver(**a)
    cm, cv = g1.build_cholesky()
    monkeypatch.setenv("IBFS_S_WINDOW", "0")
    g0 = hip.Solver(**a)
    g0.set_cholesky(cm, cv)
    h1, h0 = g1.advance(10), g0.advance(10)
    s1, s0 = g1.get_state(), g0.get_state()
    assert np.array_equal(h1, h0)
    for k in s0:
        assert np.array_equal(s1[k], s0[k]), k
