"""Tangent-linear and adjoint runs on mixed handles (FP32 fields, FP64 force
system), checked against the FP64 runs, and the large batched adjoint
configuration: 8 copies of the 1024^2 x 6 tandem-cylinder case (Re 200,
dt 0.0015, 628 IB points), each with its own seeded surface-velocity
perturbation, objective mean lift + mean drag, in one handle.  The full
configuration runs 2000 + 400 steps with checkpoints every 100 (bench.py
--config5); here the windows are shortened to keep the test short."""
import numpy as np
import pytest

from conftest import case_args

pytestmark = pytest.mark.gpu


def pair(name):
    from ibfs_amd import hip

    s64 = hip.Solver(**case_args(name))
    cm, cv = s64.build_cholesky()
    sm = hip.Solver(**case_args(name), precision=hip.MIXED)
    sm.set_cholesky(cm, cv)
    return s64, sm


def test_mixed_adjoint_against_fp64_and_mixed_tlm():
    s64, sm = pair("cyl32")
    ub = np.random.default_rng(0).normal(0.0, 0.05, s64.nf)
    J64, g64 = s64.adjoint(12, 8, 5, ub, 1.0, 0.5)
    Jm, gm = sm.adjoint(12, 8, 5, ub, 1.0, 0.5)
    gerr = np.linalg.norm(gm - g64) / np.linalg.norm(g64)
    print("J fp64", J64[0], "mixed", Jm[0], "gradient error", gerr)
    assert abs(Jm[0] - J64[0]) <= 1e-4 * abs(J64[0])
    assert gerr < 1e-4
    rng = np.random.default_rng(3)
    for _ in range(2):
        du = rng.standard_normal(s64.nf)
        sm.reset()
        Jt, dJ = sm.tangent(12, 8, ub, du, 1.0, 0.5)
        print("mixed adjoint", gm @ du, "mixed tlm", dJ[0])
        assert abs(gm @ du - dJ[0]) <= 1e-4 * (abs(dJ[0]) + 1e-3 * np.linalg.norm(gm) * np.linalg.norm(du))


def test_batched_mixed_adjoint_1024_tandem(monkeypatch):
    from ibfs_amd import cases, hip

    c = cases.tandem_1024()
    s64 = hip.Solver(**c)
    assert s64.nb == 628 and s64.L == 6
    cm, cv = s64.build_cholesky()
    nf = s64.nf
    rng = np.random.default_rng(5)
    ubs = [rng.normal(0.0, 0.05, nf) for _ in range(8)]
    ns, nw, ck = 60, 20, 40
    ens = hip.Ensemble([c] * 8, precision=hip.MIXED)
    for k in range(8):
        ens.set_case_cholesky(k, cm, cv)
    J, g = ens.adjoint(ns, nw, ck, np.concatenate(ubs), 1.0, 1.0)
    assert g.shape == (8 * nf,) and np.all(np.isfinite(g))
    # every case of the batch is the single-case run of that case, bit for bit
    # (step 1 level by level on both sides: the batch of 8 runs it that way)
    monkeypatch.setenv("IBFS_STEP1_LEVELS", "0")
    sm = hip.Solver(**c, precision=hip.MIXED)
    sm.set_cholesky(cm, cv)
    for k in (0, 5):
        sm.reset()
        Jk, gk = sm.adjoint(ns, nw, ck, ubs[k], 1.0, 1.0)
        assert Jk[0] == J[k]
        assert np.array_equal(gk, g[k * nf:(k + 1) * nf])
    # gradient error of the mixed form against the FP64 adjoint
    for k in (0, 5):
        s64.reset()
        J64, g64 = s64.adjoint(ns, nw, ck, ubs[k], 1.0, 1.0)
        gk = g[k * nf:(k + 1) * nf]
        err = np.linalg.norm(gk - g64) / np.linalg.norm(g64)
        print(f"case {k}: J fp64 {J64[0]:.10f} mixed {J[k]:.10f}; gradient error {err:.2e}")
        assert abs(J[k] - J64[0]) <= 1e-4 * abs(J64[0])
        assert err < 1e-3
