"""Parity at the production grid sizes (the size-specialised DST passes):
transforms at every configured FFT length and at sizes outside that set
(the generic passes), and a few full steps of the
200^2 x 3-level cylinder (the CPU reference case) and of the 512^2 x 5-level
airfoil (the headline case) against the oracle.  The oracle cannot afford
its own preprocess at these sizes, so both sides use the device factor,
after a spot check of columns of A against the oracle's a_times."""
import numpy as np
import pytest

import oracle as orc
from conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-11


def solver_pair(m, L, len_, offx, offy, body, dt, re):
    from ibfs_amd import geometry, hip

    delta = len_ / m
    if body == "circle":
        xb, yb, bd = geometry.circle(delta, 1.0, 0.0, 0.0)
    else:
        xb, yb, bd = geometry.naca4(delta, aoa_deg=20.0)
    a = dict(m=m, n=m, mgridlev=L, dt=dt, re=re, len_=len_, offx=offx, offy=offy, xb=xb, yb=yb, bdyno=bd)
    return orc.Oracle(**a), hip.Solver(**a)


@pytest.mark.parametrize("m", [128, 200, 256, 300, 320, 384, 512, 600, 1024])
def test_transforms_at_production_sizes(m, rng):
    o, g = solver_pair(m, 3, 4.0, 2.0, 2.0, "circle", 0.004, 100.0)
    assert min(g.tile_shapes().values()) >= 1  # every kind on a specialised pass
    x = rng.standard_normal(o.nw)
    assert relerr(g.dst(x), o.dst(x)) < 1e-13
    assert relerr(g.ctci(x), o.ctci(x)) < 1e-12
    assert relerr(g.ainv(x), o.ainv(x)) < 1e-12
    if m <= 512:
        assert relerr(g.op_dst_gemm(x, 4), o.dst(x)) < 1e-13


@pytest.mark.parametrize("m,n", [(1200, 16), (16, 1200), (480, 480)])
def test_transforms_on_the_generic_passes(m, n, rng):
    """Sizes with no size-specialised pass run the generic one, with one
    FFT per workgroup where a pass has few sequences and several where it
    has many: a transform longer than 2048 (2m = 2400), the 1199 short
    columns of the 1200 x 16 grid (two FFTs per workgroup), and 479^2."""
    from ibfs_amd import geometry, hip

    len_ = 4.0
    delta = len_ / m
    xb, yb, bd = geometry.circle(delta, 4 * delta, 0.0, 0.0)
    a = dict(m=m, n=n, mgridlev=2, dt=0.25 * delta, re=100.0, len_=len_, offx=len_ / 2, offy=len_ * n / m / 2,
             xb=xb, yb=yb, bdyno=bd)
    o, g = orc.Oracle(**a), hip.Solver(**a)
    assert set(g.tile_shapes().values()) == {0}  # every kind on the generic pass
    x = rng.standard_normal(o.nw)
    assert relerr(g.dst(x), o.dst(x)) < 1e-13
    assert relerr(g.ctci(x), o.ctci(x)) < 1e-12
    assert relerr(g.ainv(x), o.ainv(x)) < 1e-12


def run_pair(o, g, nsteps, spot_cols):
    cm, cv = g.build_cholesky()
    A_cols = cm.reshape(o.nf, o.nf).T  # [i, j]: upper triangle holds A
    for j in spot_cols:
        col = o.a_column(j)
        # rows i <= j of column j are stored in the upper triangle
        assert relerr(A_cols[: j + 1, j], col[: j + 1]) < TOL
    o.set_cholesky(cm, cv)
    oh = np.array([np.concatenate([f, c, [s]]) for f, c, s in (o.advance() for _ in range(nsteps))])
    gh = g.advance(nsteps)
    st = g.get_state()
    for k in range(o.L):
        w = slice(k * o.nw, (k + 1) * o.nw)
        qq = slice(k * o.nq, (k + 1) * o.nq)
        assert relerr(st["omega"][w], o.omega[w]) < TOL, f"omega level {k + 1}"
        assert relerr(st["q"][qq], o.q[qq]) < TOL, f"q level {k + 1}"
    assert relerr(st["fb"], o.fb) < TOL
    fscale = np.max(np.abs(oh[:, :2]))
    assert np.max(np.abs(gh[:, :2] - oh[:, :2])) < TOL * fscale


def test_cylinder_200_three_levels():
    """The CPU reference case (200^2, 3 levels, Re 100, dt 0.004): 4 steps."""
    o, g = solver_pair(200, 3, 4.0, 2.0, 2.0, "circle", 0.004, 100.0)
    assert o.nb == 157 and g.nb == 157
    run_pair(o, g, 4, [0, 78, 313])


@pytest.mark.parametrize("m", [128, 300])
def test_cylinder_three_levels_other_sizes(m):
    """Cylinder steps at the other specialised lengths (2m = 256: one wave
    per FFT; 2m = 600: radices 4, 2, 3, 5, 5): 3 steps."""
    o, g = solver_pair(m, 3, 4.0, 2.0, 2.0, "circle", 1.0 / m, 100.0)
    run_pair(o, g, 3, [0, o.nf // 2, o.nf - 1])


def test_airfoil_512_five_levels():
    """The headline case (512^2, 5 levels, NACA0012 at 20 deg, Re 300): 2 steps."""
    o, g = solver_pair(512, 5, 4.0, 1.0, 2.0, "naca", 0.002, 300.0)
    assert g.nb == 261
    run_pair(o, g, 2, [0, 130, 521])


def test_step_switches_at_512_five_levels(monkeypatch):
    """The headline case under the switches that select equivalent forms:
    the coarsify chain in LDS tiles or one place per thread, step 4's
    coarsify on the tiles step 3 can change or on every tile, step 2's
    level-0 transforms on the force-stencil rectangle or the whole field,
    and step 4's left/right terms kept apart for the tail or added into the
    streamfunction, give bit-identical states after 3 steps; the band correction in real
    space or as spectral column pairs, step 3 in real space or as the
    spectral ctci, the ring vectors' bottom and top lines by one FFT or
    by direct sums, and the ring chain on the representative rows of T's
    symmetry orbits or on every row, agree to rounding."""
    from ibfs_amd import hip

    def run(**env):
        for k in ("IBFS_COARSIFY_TILES", "IBFS_COARSIFY_INC", "IBFS_S_WINDOW", "IBFS_BAND_FFT", "IBFS_STEP3_LOCAL",
                  "IBFS_RING_FFT", "IBFS_SPLIT_S", "IBFS_RING_SYM"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        _, g = solver_pair(512, 5, 4.0, 1.0, 2.0, "naca", 0.002, 300.0)
        g.build_cholesky()
        h = g.advance(3)
        return h, g.get_state()

    h, st = run()
    for env in ({"IBFS_COARSIFY_TILES": "0"}, {"IBFS_COARSIFY_INC": "0"}, {"IBFS_S_WINDOW": "0"},
                {"IBFS_SPLIT_S": "0"}):
        h1, st1 = run(**env)
        assert np.array_equal(h1, h), env
        for k in st:
            assert np.array_equal(st1[k], st[k]), (env, k)
    for env in ({"IBFS_BAND_FFT": "1"}, {"IBFS_STEP3_LOCAL": "0"}, {"IBFS_RING_FFT": "0"}, {"IBFS_RING_SYM": "0"}):
        h1, st1 = run(**env)
        for k in ("omega", "q"):
            assert relerr(st1[k], st[k]) < 1e-13, (env, k)
        # the force system amplifies a rounding-level change of omega by its
        # condition number
        assert relerr(st1["fb"], st["fb"]) < 1e-12, env
        assert not np.array_equal(st1["omega"], st["omega"]), env


def test_tile_shapes_do_not_change_results(monkeypatch):
    """The tuner's choice of FFTs per workgroup (1, 2 or 4) must not change a
    single bit: sequences 2p and 2p+1 share FFT p whatever the shape.  Forced
    shapes and the tuned one on the 200^2 x 3 cylinder (FFT length 400, all
    three shapes fit) give identical states after 3 steps."""
    from ibfs_amd import hip

    runs = {}
    for tile in ("1", "2", "4", None):
        if tile is None:
            monkeypatch.delenv("IBFS_TILE", raising=False)
        else:
            monkeypatch.setenv("IBFS_TILE", tile)
        _, g = solver_pair(200, 3, 4.0, 2.0, 2.0, "circle", 0.004, 100.0)
        shapes = g.tile_shapes()
        if tile is not None:
            assert set(shapes.values()) == {int(tile)}
        g.build_cholesky()
        g.advance(3)
        runs[tile] = (shapes, g.get_state())
    base = runs["1"][1]
    for tile, (shapes, st) in runs.items():
        print("tile", tile, shapes)
        for k in base:
            assert np.array_equal(st[k], base[k]), k


@pytest.mark.parametrize("m", [128, 256, 512])
def test_register_blocked_passes_do_not_change_results(monkeypatch, m):
    """The register-blocked kernels (tile code 8: FFT lengths 2m = 256, 512,
    1024, 16 values per lane, two radix-4 stages per LDS exchange; tile code
    16: the column passes with staged loads and stores) run the butterflies
    of the LDS passes with the same twiddles, so forcing them (IBFS_TILE=8 or
    16; the step-1 loader stays on P = 4), forcing P = 1, and the tuned
    choice give identical states after 3 steps of a 4-level cylinder, whose
    level-batched launches are tuned on their own batch."""
    from ibfs_amd import hip

    runs = {}
    for tile in ("1", "8", "16", None):
        if tile is None:
            monkeypatch.delenv("IBFS_TILE", raising=False)
        else:
            monkeypatch.setenv("IBFS_TILE", tile)
        _, g = solver_pair(m, 4, 4.0, 2.0, 2.0, "circle", 0.5 * 4.0 / m, 100.0)
        shapes, lshapes = g.tile_shapes(), g.level_tile_shapes()
        if tile == "8":
            assert shapes["row"] == 8 and shapes["col2"] == 8 and shapes["step1_row"] == 4
        if tile == "16":
            assert shapes["row"] == 8 and shapes["col"] == 16 and shapes["col2"] == 16 and lshapes["col2"] == 16
        g.build_cholesky()
        g.advance(3)
        runs[tile] = (shapes, lshapes, g.get_state())
    base = runs["1"][2]
    for tile, (shapes, lshapes, st) in runs.items():
        print("m", m, "tile", tile, shapes, "levels", lshapes,
              {k: f"{relerr(st[k], base[k]):.1e}" for k in base if k != "q0"})
    for tile, (shapes, lshapes, st) in runs.items():
        for k in base:
            assert np.array_equal(st[k], base[k]), k


@pytest.mark.parametrize("tile", ["8", "16"])
@pytest.mark.parametrize("m", [128, 256, 512])
def test_register_blocked_transforms(monkeypatch, m, tile, rng):
    """The register-blocked kernels forced on (IBFS_TILE=8, or 16 for the
    staged column passes) for the operator entry points: dst, ctci and ainv
    against the oracle."""
    monkeypatch.setenv("IBFS_TILE", tile)
    o, g = solver_pair(m, 3, 4.0, 2.0, 2.0, "circle", 0.004, 100.0)
    assert g.tile_shapes()["row"] == 8 and g.tile_shapes()["col2"] == int(tile)
    x = rng.standard_normal(o.nw)
    errs = (relerr(g.dst(x), o.dst(x)), relerr(g.ctci(x), o.ctci(x)), relerr(g.ainv(x), o.ainv(x)))
    print("m", m, "dst %.1e ctci %.1e ainv %.1e" % errs)
    assert errs[0] < 1e-13 and errs[1] < 1e-12 and errs[2] < 1e-12


def test_tandem_1024_six_levels():
    """The large case (1024^2 x 6 levels, dx 0.01, two d = 1 cylinders 3 d
    apart, 628 points, Re 200, dt 0.0015): 2 FP64 steps with the device
    factor on both sides.  At Nf = 1256 this pins the step's explicit-inverse
    force solve (rows of A^-1 = L^-T L^-1) against the oracle's two triangular
    solves of cholsl (src/operators.f90:794-819)."""
    from ibfs_amd import cases, hip

    a = cases.tandem_1024()
    o, g = orc.Oracle(**a), hip.Solver(**a)
    assert g.nb == 628 and g.L == 6
    run_pair(o, g, 2, [0, 627, 1255])
