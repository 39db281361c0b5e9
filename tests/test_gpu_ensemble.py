"""Ensembles and the mixed-precision form.

* An ensemble runs each case with the same kernels and arithmetic as a
  single-case handle, so every case must match its own single-case run bit
  for bit (cases of different sizes, level counts, Re, dt and bodies mixed).
* The mixed form (FP32 fields and transforms, FP64 force solve, reductions
  and one refinement sweep) is measured against the FP64 form; the bounds
  are the ones the chosen form is held to (force histories <= 1e-5 relative
  to the force scale) plus per-field bounds recorded with their values."""
import numpy as np
import pytest

from conftest import case_args, relerr

pytestmark = pytest.mark.gpu


def ens_cases():
    a = case_args("cyl32")
    b = case_args("cyl32")
    b.update(mgridlev=2, re=250.0, dt=0.015)
    c = case_args("naca48x40")
    d = case_args("cyl40r5")
    e = case_args("cyl32")
    e.update(offx=1.5, offy=2.25, re=60.0)
    return [a, b, c, d, e]


def test_ensemble_matches_single_case_runs(monkeypatch):
    """Both sides run step 1 and vort2flux level by level
    (IBFS_STEP1_LEVELS=0, IBFS_V2F_LEVELS=0): a one- or two-case group
    otherwise batches its levels, which sums the edge terms separately and so
    rounds differently (those forms are held to the oracle by the parity
    tests)."""
    from ibfs_amd import hip

    monkeypatch.setenv("IBFS_STEP1_LEVELS", "0")
    monkeypatch.setenv("IBFS_V2F_LEVELS", "0")

    cases = ens_cases()
    ens = hip.Ensemble(cases)
    assert ens.ncases == 5
    ens.build_cholesky()
    singles = []
    for c in cases:
        s = hip.Solver(**c)
        s.build_cholesky()
        singles.append(s)
    hist = np.concatenate([ens.advance(4), ens.advance(7)])
    for i, s in enumerate(singles):
        sh = s.advance(11)
        assert np.array_equal(ens.case_rows(hist, i), sh), f"case {i} history"
        es, ss = ens.get_state(i), s.get_state()
        for key in ("omega", "q", "rhs_old", "fb"):
            assert np.array_equal(es[key], ss[key]), f"case {i} {key}"


@pytest.mark.parametrize("switch", ["IBFS_STEP1_SPLIT", "IBFS_COARSIFY_INC"])
def test_equivalent_forms_in_an_uneven_batch(monkeypatch, switch):
    """An FP64 batch of uneven cases (2 to 5 levels, step 1 and vort2flux
    level by level), each switch on (the default) and off, for 6 steps:
    IBFS_STEP1_SPLIT, the freestream source evaluated by its own kernel into
    a scratch field and read by a plain row pass, or inside the row pass's
    loader (the same arithmetic; a case without the level being solved gets
    zeros from either form); IBFS_COARSIFY_INC, step 4's coarsify on the
    tiles whose inputs step 3 changed (each case its own rectangle), or on
    every tile.  The states are identical."""
    from ibfs_amd import hip

    monkeypatch.setenv("IBFS_STEP1_LEVELS", "0")
    monkeypatch.setenv("IBFS_V2F_LEVELS", "0")
    runs = []
    for on in ("1", "0"):
        monkeypatch.setenv(switch, on)
        ens = hip.Ensemble(ens_cases())
        ens.build_cholesky()
        hist = ens.advance(6)
        runs.append((hist, [ens.get_state(i) for i in range(ens.ncases)]))
    (h1, s1), (h0, s0) = runs
    assert np.array_equal(h1, h0)
    for i in range(len(s1)):
        for key in ("omega", "q", "rhs_old", "fb"):
            assert np.array_equal(s1[i][key], s0[i][key]), f"case {i} {key}"


@pytest.mark.parametrize("name", ["naca48x40", "cyl40r5"])
def test_mixed_step2_on_the_rings(name, monkeypatch):
    """The mixed form's step 2 needs the FP64 streamfunctions of the levels
    below the finest only on their rings (get_bc of the next finer level):
    the ring lines of the column pair's output (the default) or the whole
    last row pass (IBFS_STEP2_RING=0) give those values to FP64 rounding, so
    after 20 steps every field is within the mixed form's 1e-5 of the other,
    and the force solution differs (the ring form was taken)."""
    from ibfs_amd import hip

    a = case_args(name)
    g64 = hip.Solver(**a)
    cm, cv = g64.build_cholesky()
    runs = []
    for on in ("1", "0"):
        monkeypatch.setenv("IBFS_STEP2_RING", on)
        g = hip.Solver(**a, precision=hip.MIXED)
        g.set_cholesky(cm, cv)
        runs.append((g.advance(20), g.get_state()))
    (h1, s1), (h0, s0) = runs
    nb2 = 2 * g64.nbody
    assert np.max(np.abs(h1[:, :nb2] - h0[:, :nb2])) < 1e-5 * np.max(np.abs(h0[:, :nb2]))
    for k in ("omega", "q", "fb"):
        assert relerr(s1[k], s0[k]) < 1e-5, k
    assert not np.array_equal(s1["fb"], s0["fb"])


def test_ensemble_factor_equals_single_factor():
    from ibfs_amd import hip

    cases = ens_cases()[:3]
    ens = hip.Ensemble(cases)
    for i, c in enumerate(cases):
        cm, cv = ens.build_case_cholesky(i)
        s = hip.Solver(**c)
        cm2, cv2 = s.build_cholesky()
        assert np.array_equal(cm, cm2) and np.array_equal(cv, cv2)


def mixed_vs_fp64(args, nsteps):
    from ibfs_amd import hip

    g64 = hip.Solver(**args)
    g32 = hip.Solver(**args, precision=hip.MIXED)
    cm, cv = g64.build_cholesky()
    g32.set_cholesky(cm, cv)
    h64 = g64.advance(nsteps)
    h32 = g32.advance(nsteps)
    s64, s32 = g64.get_state(), g32.get_state()
    nb = g64.nbody
    fscale = np.max(np.abs(h64[:, : 2 * nb]))
    ferr = np.max(np.abs(h32[:, : 2 * nb] - h64[:, : 2 * nb])) / fscale
    errs = {"force": ferr}
    for k in range(g64.L):
        w = slice(k * g64.nw, (k + 1) * g64.nw)
        qq = slice(k * g64.nq, (k + 1) * g64.nq)
        errs[f"omega{k + 1}"] = relerr(s32["omega"][w], s64["omega"][w])
        errs[f"q{k + 1}"] = relerr(s32["q"][qq], s64["q"][qq])
    errs["fb"] = relerr(s32["fb"], s64["fb"])
    return errs


@pytest.mark.parametrize("name", ["cyl32", "naca48x40", "cyl40r5"])
def test_mixed_form_small(name):
    errs = mixed_vs_fp64(case_args(name), 50)
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():  # every field of every level
        assert v < 1e-5, f"{k} {v:.2e}"


def mixed_vs_oracle(args, nsteps):
    """Per-field error of the mixed form against the oracle (both with the
    device FP64 factor): omega and q per level, fb, and the force history
    relative to the force scale."""
    import oracle as orc
    from ibfs_amd import hip

    g64 = hip.Solver(**args)
    cm, cv = g64.build_cholesky()
    g64.close()
    g32 = hip.Solver(**args, precision=hip.MIXED)
    g32.set_cholesky(cm, cv)
    o = orc.Oracle(**args)
    o.set_cholesky(cm, cv)
    oh = np.array([np.concatenate([f, c, [s]]) for f, c, s in (o.advance() for _ in range(nsteps))])
    gh = g32.advance(nsteps)
    st = g32.get_state()
    nb = g32.nbody
    errs = {"force": np.max(np.abs(gh[:, : 2 * nb] - oh[:, : 2 * nb])) / np.max(np.abs(oh[:, : 2 * nb]))}
    for k in range(o.L):
        w = slice(k * o.nw, (k + 1) * o.nw)
        qq = slice(k * o.nq, (k + 1) * o.nq)
        errs[f"omega{k + 1}"] = relerr(st["omega"][w], o.omega[w])
        errs[f"q{k + 1}"] = relerr(st["q"][qq], o.q[qq])
    errs["fb"] = relerr(st["fb"], o.fb)
    return errs


def cylinder(m, levels=3):
    from ibfs_amd import cases

    return cases.cylinder_grid(m, levels)


@pytest.mark.parametrize("which,nsteps", [("cyl32", 50), ("naca48x40", 50), ("cyl40r5", 50), (128, 10),
                                          (300, 5), (600, 3)])
def test_mixed_form_against_oracle(which, nsteps):
    """The mixed form held to the FP32 bar (1e-5 per field) against the
    oracle itself, on the small cases and on cylinders whose FP32 transforms
    run the size-specialised passes of 2m = 256, 600 and 1200."""
    args = case_args(which) if isinstance(which, str) else cylinder(which)
    errs = mixed_vs_oracle(args, nsteps)
    print(which, {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-5, f"{k} {v:.2e}"


def test_mixed_form_headline_case():
    """512^2 x 5 levels NACA0012 at 20 deg, Re 300: 300 steps against FP64
    and 6 against the oracle, every field within 1e-5."""
    from ibfs_amd import geometry

    xb, yb, bd = geometry.naca4(4.0 / 512, aoa_deg=20.0)
    args = dict(m=512, n=512, mgridlev=5, dt=0.002, re=300.0, len_=4.0, offx=1.0, offy=2.0, xb=xb, yb=yb, bdyno=bd)
    errs = mixed_vs_fp64(args, 300)
    print("naca512 vs fp64, 300 steps", {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():  # every field of every level
        assert v < 1e-5, f"{k} {v:.2e}"
    errs = mixed_vs_oracle(args, 6)
    print("naca512 vs oracle, 6 steps", {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-5, f"{k} {v:.2e}"
