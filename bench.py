#!/usr/bin/env python3
"""Throughput of the multi-domain IBFS time step on one MI355X.

Workload (the metric's configuration): 512 x 512 cells per level, 5 levels,
finest domain [-1,3] x [-2,2], stationary NACA0012 (chord 1, 20 deg nose-up
about the quarter chord, IB spacing = grid spacing -> 261 points), Re = 300,
dt = 0.002, impulsive start from rest (synthetic: no data files).  One step
is one call of the reference's advance (src/operators.f90:62-188).

Prints one JSON line.  ``value`` is steps/s of the plain FP64 form (the
drop-in).  ``chosen`` reports the mixed FP32/FP64 form for one case and for
a batch of independent copies of the case in one graph per step (aggregate
case-steps/s), with its per-field error against the FP64 run.  ``roofline`` is
the step's dominant kernel (the fused column pair of ctci, timed live inside
steps with events on its own dispatches); ``cpu_baseline`` is the oracle's FP64
step, single thread, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ibfs-if2d_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# HBM bytes per launch of the fused column pair at 512^2 (FP64): FETCH_SIZE
# and WRITE_SIZE per dispatch from two rocprofv3 --pmc runs of this build,
# each scaled by what the counter reads for a copy kernel of the same size;
# the summary of those runs records the figure on a line
# "traffic_bytes_per_launch: N"
PAIR_COUNTERS = os.path.join(ROOT, "profiles", "dst_col2_512x5_fp64_hbm_counters.md")


def pair_traffic():
    """(bytes per launch, source) from the committed counter summary, or
    (None, reason) when it is absent."""
    try:
        with open(PAIR_COUNTERS) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1].split()[0]), os.path.relpath(PAIR_COUNTERS, ROOT)
    except OSError:
        pass
    return None, "no counter summary found"


def step_bytes(m, n, L, nf, s=8):
    """Algorithmic HBM bytes of one step: every 2-D DST-I reads and writes one
    W field (6L+2 per step: two per level in step 1, which transforms only
    the increment, two per level in each vort2flux, two in the projection),
    step 1 reads q, q0, reads/writes rhs_old and reads omega twice (stencil,
    increment) per level, coarsify/projection reads, curl writes, and the
    Cholesky factor read (fixed formula, independent of how the kernels are
    fused)."""
    W = (m - 1) * (n - 1)
    Nq = (m + 1) * n + (n + 1) * m
    return s * (2 * (6 * L + 2) * W + L * (2 * Nq + 4 * W) + (2 * L - 1) * W + (L + 1) * Nq) + 8 * nf * (nf - 1)


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def cpu_baseline(case, cm, cv, budget_s=15.0):
    import oracle as orc

    o = orc.Oracle(**case)
    o.set_cholesky(cm, cv)
    o.advance()  # itime 1 (explicit Euler start), untimed
    n = 0
    t0 = time.perf_counter()
    while True:
        o.advance()
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 200:
            break
    return n / el, n


MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense matrix peaks


def dense_vs_fft(hip, cases, sizes=(128, 200, 256, 320, 384, 512)):
    """One DST-I pass over an (m-1)^2 field, single case: the FFT pass (the
    tuned shape the step uses) against the dense product with the DST-I
    matrix on the matrix cores (best of ks = 1, 2, 4, 8 waves per output
    tile), along i and along j, in FP64 and FP32 (the mixed form's
    element type).  flops of the dense pass: 2 (m-1)^3."""
    out = {}
    for prec, name in ((hip.PLAIN_FP64, "fp64"), (hip.MIXED, "fp32")):
        rows = {}
        for m in sizes:
            s = hip.Solver(**cases.cylinder_grid(m), precision=prec)
            r = {}
            for axis, ax in ((0, "row"), (1, "col")):
                fft = s.time_dst_pass(axis, 200) * 1e3
                best = min((s.time_dst_gemm(axis, ks, 200) * 1e3, ks) for ks in (1, 2, 4, 8))
                flops = 2.0 * (m - 1) ** 3
                r[ax] = {"fft_us": fft, "mfma_us": best[0], "mfma_ks": best[1],
                         "mfma_tflops": flops / (best[0] * 1e-6) / 1e12,
                         "mfma_util": flops / (best[0] * 1e-6) / 1e12 / MFMA_PEAK_TFLOPS[name],
                         "faster": "mfma" if best[0] < fft else "fft"}
            # the step's column work is the fused pair (two transforms, one
            # launch); its dense form is two column products
            pair = s.time_dst_pass(2, 200) * 1e3
            r["col_pair"] = {"fft_us": pair, "mfma_us": 2 * r["col"]["mfma_us"],
                             "faster": "mfma" if 2 * r["col"]["mfma_us"] < pair else "fft"}
            rows[str(m)] = r
            log(f"dense vs FFT {name} m={m}: row FFT {r['row']['fft_us']:.2f} us / MFMA {r['row']['mfma_us']:.2f} us "
                f"(ks {r['row']['mfma_ks']}, {100 * r['row']['mfma_util']:.1f}% of peak); col FFT "
                f"{r['col']['fft_us']:.2f} / MFMA {r['col']['mfma_us']:.2f} us")
            del s
        out[name] = rows
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16, help="copies of the case in the batched chosen form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-chosen", action="store_true", help="skip the chosen-form legs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--adjoint-spin", type=int, default=2000, help="spin-up steps of the adjoint leg (0 with "
                    "--adjoint-window 0 skips it)")
    ap.add_argument("--adjoint-window", type=int, default=400)
    ap.add_argument("--ensemble", type=int, default=1, help="run the 64-case uneven ensemble leg (0: skip)")
    ap.add_argument("--config5", type=int, default=1, help="run the batched 1024^2 x 6 tandem adjoint leg (0: skip)")
    ap.add_argument("--config5-spin", type=int, default=2000)
    ap.add_argument("--config5-window", type=int, default=400)
    ap.add_argument("--cg-steps", type=int, default=20, help="steps of the stationary = F leg (CG force solve; "
                    "0: skip)")
    ap.add_argument("--dense-sweep", type=int, default=1, help="time the dense matrix-core DST-I passes against "
                    "the FFT passes at every production size (0: skip)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (independent cases would shard "
              "one per GPU with no communication; one case does not shard)", file=sys.stderr)
        return 2

    from ibfs_amd import cases, hip

    case = cases.naca_512()
    K, W0 = args.steps, max(args.warmup, 1)

    # ---- plain FP64, one case
    g = hip.Solver(**case)
    t0 = time.perf_counter()
    cm, cv = g.build_cholesky()
    t_pre = time.perf_counter() - t0
    log(f"preprocess (device a_times x {g.nf} + choldc): {t_pre:.3f} s")
    h64 = g.advance(W0)
    t0 = time.perf_counter()
    dev_ms = g.advance_timed(K)
    wall = time.perf_counter() - t0
    steps_s = K / wall
    log(f"plain FP64: {K} steps wall {wall * 1e3:.1f} ms (device {dev_ms:.1f} ms) -> {steps_s:.1f} steps/s")

    # ---- dominant kernel of the step: the fused column pair of ctci (DST-I
    # along j, times laminv, DST-I along j) of vort2flux; with the levels
    # batched one launch transforms every level's field.  Per launch it reads
    # each field once, the table laminv (W doubles, shared by the levels)
    # once, and writes each field once.  In-step duration: events on the
    # kernel's own dispatches during 20 eager steps of a second solver (its
    # state advances; the rocprofv3 trace of the graph-replayed step is the
    # check); isolated back-to-back launches are reported beside it
    gk = hip.Solver(**case)
    gk.set_cholesky(cm, cv)
    gk.advance(2)
    pair_us, pair_n, pair_fields = gk.time_step_kernel(20)
    del gk
    fields_per_launch = pair_fields / max(pair_n, 1)
    pair_bytes = int(round((2 * fields_per_launch + 1) * 8 * g.nw))
    t_pair = g.time_dst_pass(2, 200)
    t_row = g.time_dst_pass(0, 200)
    t_col = g.time_dst_pass(1, 200)
    achieved = pair_bytes / (pair_us * 1e-6) / 1e9
    traffic, traffic_src = pair_traffic()
    B = step_bytes(g.m, g.n, g.L, g.nf)
    log(f"column pair in the step {pair_us:.2f} us over {pair_n} launches of {fields_per_launch:g} fields "
        f"({pair_bytes / 1e6:.2f} MB, {achieved:.0f} GB/s); one field isolated "
        f"{t_pair * 1e3:.2f} us; DST pass: row {t_row * 1e3:.2f} us, column {t_col * 1e3:.2f} us; step bytes "
        f"{B / 1e6:.1f} MB -> whole-step {B * steps_s / 1e9:.1f} GB/s")
    # streaming copies: one of the pass's size (the counter calibration point)
    # and a large one (the measured copy bandwidth of this GPU)
    t_copy_w = g.time_copy(g.nw, 200)
    n_big = 1 << 25  # 32 Mi doubles = 256 MiB each way
    t_copy_big = g.time_copy(n_big, 20)
    copy_gbs = 2 * 8 * n_big / (t_copy_big * 1e-3) / 1e9
    log(f"copy of W doubles {t_copy_w * 1e3:.2f} us; 256 MiB copy {copy_gbs:.0f} GB/s")

    cg = None
    if args.cg_steps:
        # ---- stationary = F (the reference's default): cjgr on the device
        # a_times every step, warm-started from the previous fb, cgtol 1e-7
        gc = hip.Solver(**case)
        gc.set_solver(False, cgtol=1e-7, cg_max_iter=3000)
        gc.advance(5)
        its0 = gc.cg_iterations(5)
        t0 = time.perf_counter()
        gc.advance(args.cg_steps)
        wcg = time.perf_counter() - t0
        its = gc.cg_iterations(args.cg_steps)
        cg = {"steps_per_s": args.cg_steps / wcg, "ms_per_step": 1e3 * wcg / args.cg_steps,
              "cg_iterations_first5": [int(v) for v in its0], "cg_iterations_mean": float(np.mean(its)),
              "ms_per_iteration": 1e3 * wcg / max(1, int(np.sum(its))), "cgtol": 1e-7,
              "note": "eager launches, one host read-back per CG iteration"}
        log(f"stationary = F (CG force solve): {cg['steps_per_s']:.1f} steps/s, {cg['cg_iterations_mean']:.1f} "
            f"iterations per step ({cg['ms_per_iteration']:.3f} ms per iteration); first steps {cg['cg_iterations_first5']}")
        del gc

    dense = None
    if args.dense_sweep:
        dense = dense_vs_fft(hip, cases)

    chosen = None
    if not args.no_chosen:
        # ---- mixed form, one case: force error against the FP64 run over the warm-up window
        gm = hip.Solver(**case, precision=hip.MIXED)
        gm.set_cholesky(cm, cv)
        hm = gm.advance(W0)
        nb = g.nbody
        ferr = float(np.max(np.abs(hm[:, :2 * nb] - h64[:, :2 * nb])) / np.max(np.abs(h64[:, :2 * nb])))
        t0 = time.perf_counter()
        gm.advance_timed(K)
        wm = time.perf_counter() - t0
        mixed_s = K / wm
        # per-field error against the FP64 run: both forms have taken W0 + K
        # steps from the same start; Wf more give the window of the mean forces
        Wf = 500
        hf, hmf = g.advance(Wf), gm.advance(Wf)
        a, b = g.get_state(), gm.get_state()

        def rel_levels(key, per):
            return [float(np.max(np.abs(b[key][k * per:(k + 1) * per] - a[key][k * per:(k + 1) * per]))
                          / np.max(np.abs(a[key][k * per:(k + 1) * per]))) for k in range(g.L)]

        # circulation per level against the level's sum of |omega| dx^2 (the
        # net circulation of an impulsive start stays near zero)
        dx = case["len_"] / case["m"]
        circ_err = [float(np.max(np.abs(hmf[:, 2 * nb + k] - hf[:, 2 * nb + k]))
                          / (np.sum(np.abs(a["omega"][k * g.nw:(k + 1) * g.nw])) * (dx * 2 ** k) ** 2))
                    for k in range(g.L)]
        F64, Fm = hf[:, :2 * nb], hmf[:, :2 * nb]
        mean_rel = lambda cols: float(abs(Fm[:, cols].sum(1).mean() - F64[:, cols].sum(1).mean())
                                      / abs(F64[:, cols].sum(1).mean()))
        field_err = {"after_steps": W0 + K + Wf, "omega_per_level": rel_levels("omega", g.nw),
                     "q_per_level": rel_levels("q", g.nq),
                     "fb": float(np.max(np.abs(b["fb"] - a["fb"])) / np.max(np.abs(a["fb"]))),
                     "force_history_window": float(np.max(np.abs(Fm - F64)) / np.max(np.abs(F64))),
                     "mean_drag_window": mean_rel(slice(0, nb)), "mean_lift_window": mean_rel(slice(nb, 2 * nb)),
                     "circulation_per_level": circ_err,
                     "window_steps": Wf,
                     "note": "max |mixed - fp64| / max |fp64| per field (per level for omega and q); mean forces "
                             "over the last window_steps steps, relative; circulation per level against the "
                             "level's sum of |omega| dx^2"}
        ferr = max(ferr, field_err["force_history_window"], field_err["mean_drag_window"],
                   field_err["mean_lift_window"])
        log(f"mixed FP32/FP64, 1 case: {mixed_s:.1f} steps/s; force error vs FP64 {ferr:.2e} (worst of the first "
            f"{W0} steps, the last {Wf} steps' history and their mean drag/lift); omega per level "
            + " ".join(f"{x:.1e}" for x in field_err["omega_per_level"]) + "; q per level "
            + " ".join(f"{x:.1e}" for x in field_err["q_per_level"]))
        del gm
        # ---- mixed form, batched: args.batch independent copies in one graph per step
        res = {}
        for prec, name in ((hip.MIXED, "mixed"), (hip.PLAIN_FP64, "fp64")):
            ens = hip.Ensemble([case] * args.batch, precision=prec)
            for c in range(args.batch):
                ens.set_case_cholesky(c, cm, cv)
            ens.advance(W0)
            kb = max(20, K // 4)
            t0 = time.perf_counter()
            ens.advance_timed(kb)
            wb = time.perf_counter() - t0
            res[name] = args.batch * kb / wb
            res[name + "_tiles"] = ens.tile_shapes(0)
            # the dominant kernel over the whole batch (the chip filled)
            res[name + "_pair_us"] = ens.time_dst_pass(2, 100) * 1e3
            log(f"{name}, batch of {args.batch}: {res[name]:.1f} case-steps/s ({kb} steps, {wb * 1e3:.1f} ms); "
                f"tile shapes {res[name + '_tiles']}")
            del ens
        esz = {"mixed": 4, "fp64": 8}
        batched_roofline = {name: {"launch_us": res[name + "_pair_us"],
                                   "bytes_per_launch": 3 * esz[name] * g.nw * args.batch,
                                   "achieved_GBs": 3 * esz[name] * g.nw * args.batch / (res[name + "_pair_us"] * 1e-6) / 1e9}
                            for name in ("mixed", "fp64")}
        for name, r in batched_roofline.items():
            r["frac"] = r["achieved_GBs"] / HBM_PEAK_GBS
            log(f"column pair over the batch of {args.batch} ({name}): {r['launch_us']:.1f} us -> "
                f"{r['achieved_GBs']:.0f} GB/s ({100 * r['frac']:.1f}% of 8 TB/s)")
        chosen = {"mixed_one_case_steps_s": mixed_s, "batched_mixed_case_steps_s": res["mixed"],
                  "column_pair_over_batch": batched_roofline,
                  "batched_fp64_case_steps_s": res["fp64"], "batch": args.batch,
                  "tile_shapes": {"one_case_fp64": g.tile_shapes(0), "batched_mixed": res["mixed_tiles"],
                                  "batched_fp64": res["fp64_tiles"]},
                  "speedup_vs_plain": res["mixed"] / steps_s, "force_err_vs_fp64": ferr,
                  "field_err_vs_fp64": field_err}

    adjoint = None
    if args.adjoint_spin + args.adjoint_window > 0:
        # ---- discrete adjoint: mean drag over the window after the spin-up,
        # gradient w.r.t. u_B ~ N(0, 0.05^2) at the 188 points of the 384^2 x 4 cylinder
        ac = cases.cylinder_384_adjoint()
        sa = hip.Solver(**ac)
        sa.build_cholesky()
        ub = np.random.default_rng(0).normal(0.0, 0.05, sa.nf)
        t0 = time.perf_counter()
        Ja, ga = sa.adjoint(args.adjoint_spin, args.adjoint_window, 50, ub, 1.0, 0.0)
        ta = time.perf_counter() - t0
        sa.reset()
        t0 = time.perf_counter()
        sa.objective(args.adjoint_spin, args.adjoint_window, ub, 1.0, 0.0)
        tf = time.perf_counter() - t0
        adjoint = {"grads_per_s": 1.0 / ta, "seconds_per_gradient": ta, "forward_only_s": tf,
                   "case": "384^2 x 4 levels cylinder (188 IB points), Re 150, dt 0.003",
                   "spin_up_steps": args.adjoint_spin, "window_steps": args.adjoint_window, "checkpoint_every": 50,
                   "parameters": int(sa.nf), "J_mean_drag": float(Ja[0]), "grad_norm": float(np.linalg.norm(ga))}
        log(f"adjoint: {ta:.2f} s per gradient ({args.adjoint_spin}+{args.adjoint_window} steps, {sa.nf} params; "
            f"forward alone {tf:.2f} s), J = {Ja[0]:.6f}")
        del sa

    ensemble = None
    if args.ensemble:
        # ---- 64 uneven cases (sizes 256..512, 3..5 levels, cylinders and airfoils), mixed form
        cs = cases.ensemble_64()
        ens = hip.Ensemble(cs, precision=hip.MIXED)
        t0 = time.perf_counter()
        ens.build_cholesky()
        tb = time.perf_counter() - t0
        ens.advance(W0)
        kb = 100
        t0 = time.perf_counter()
        ens.advance_timed(kb)
        we = time.perf_counter() - t0
        ensemble = {"cases": len(cs), "case_steps_per_s": len(cs) * kb / we, "ms_per_ensemble_step": 1e3 * we / kb,
                    "factor_build_s": tb, "precision": "mixed FP32/FP64",
                    "tile_shapes_per_group": [ens.tile_shapes(k) for k in range(ens.ngroups())]}
        log(f"ensemble of {len(cs)} uneven cases: {len(cs) * kb / we:.0f} case-steps/s ({1e3 * we / kb:.2f} ms per "
            f"ensemble step; factors {tb:.1f} s)")
        del ens

    config5 = None
    if args.config5:
        # ---- batched adjoint: 8 copies of the 1024^2 x 6 tandem-cylinder case
        # (628 IB points, Re 200, dt 0.0015), each with its own seeded u_B,
        # objective mean lift + mean drag, mixed form, checkpoints every 100
        # steps; gradient error against the FP64 adjoint of case 0
        tc = cases.tandem_1024()
        s64 = hip.Solver(**tc)
        cm5, cv5 = s64.build_cholesky()
        nf5 = s64.nf
        rng = np.random.default_rng(5)
        ubs = [rng.normal(0.0, 0.05, nf5) for _ in range(8)]
        sp5, wn5 = args.config5_spin, args.config5_window
        ens = hip.Ensemble([tc] * 8, precision=hip.MIXED)
        for k in range(8):
            ens.set_case_cholesky(k, cm5, cv5)
        ens.advance(1)  # graph capture and first-step set-up outside the timings
        ens.reset()
        t0 = time.perf_counter()
        ens.objective(sp5, wn5, np.concatenate(ubs), 1.0, 1.0)
        tf5 = time.perf_counter() - t0
        ens.reset()
        t0 = time.perf_counter()
        J5, g5 = ens.adjoint(sp5, wn5, 100, np.concatenate(ubs), 1.0, 1.0)
        ta5 = time.perf_counter() - t0
        del ens
        t0 = time.perf_counter()
        J64, g64 = s64.adjoint(sp5, wn5, 100, ubs[0], 1.0, 1.0)
        t64 = time.perf_counter() - t0
        gerr = float(np.linalg.norm(g5[:nf5] - g64) / np.linalg.norm(g64))
        config5 = {"case": "8 x (1024^2 x 6 levels, tandem cylinders D=1 gap 3D, 628 IB points, Re 200, dt 0.0015)",
                   "spin_up_steps": sp5, "window_steps": wn5, "checkpoint_every": 100, "parameters_per_case": nf5,
                   "objective": "mean lift + mean drag", "precision": "mixed FP32/FP64",
                   "forward_case_steps_per_s": 8 * (sp5 + wn5) / tf5,
                   "adjoint_s": ta5, "gradients_per_s": 8 / ta5,
                   "adjoint_case_steps_per_s": 8 * (sp5 + wn5) / ta5,
                   "fp64_single_adjoint_s": t64,
                   "grad_err_vs_fp64_case0": gerr, "J_err_vs_fp64_case0": float(abs(J5[0] - J64[0]) / abs(J64[0]))}
        log(f"config 5: 8 mixed adjoints of 1024^2 x 6 tandem ({sp5}+{wn5} steps): {ta5:.1f} s -> "
            f"{8 / ta5:.3f} gradients/s; forward {8 * (sp5 + wn5) / tf5:.0f} case-steps/s; FP64 single adjoint "
            f"{t64:.1f} s; gradient error vs FP64 {gerr:.2e}, J error {config5['J_err_vs_fp64_case0']:.2e}")
        del s64

    cpu = None
    if not args.no_cpu:
        v, n = cpu_baseline(case, cm, cv, args.cpu_seconds)
        cpu = {"value": v, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{n} steps (itime 2..{n + 1}) of the same 512^2 x 5 NACA0012 case, oracle FP64 step "
                         f"(own radix-4/2/3/5 FFT DST-I, gcc -O3), 1 thread of {os.cpu_count()} "
                         f"({platform.processor() or platform.machine()})"}
        log(f"CPU baseline: {v:.3f} steps/s over {n} steps -> GPU/CPU {steps_s / v:.1f}x")

    out = {
        "metric": "time steps/s at 512^2 x 5 levels (plain FP64)",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W0,
        "ms_per_step": 1e3 / steps_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "512x512 cells x 5 levels, NACA0012 at 20 deg (261 IB points), Re 300, dt 0.002, "
                               "impulsive start", "m": g.m, "n": g.n, "levels": g.L, "nb": g.nb,
                   "precision": "plain FP64", "parallelism": "single GPU, one case"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": f"dst_col2_n (fused column pair of ctci over {fields_per_launch:g} 511x511 FP64 "
                               "fields, one per level)",
                     "launch_us": pair_us, "launch_us_note": "mean in-step duration over the kernel's "
                     f"{pair_n} dispatches in 20 steps (events on the dispatches)", "bytes_per_launch": pair_bytes,
                     "fields_per_launch": fields_per_launch,
                     "isolated_launch_us": t_pair * 1e3, "row_pass_us": t_row * 1e3, "col_pass_us": t_col * 1e3,
                     "copy_same_size_us": t_copy_w * 1e3, "measured_copy_GBs": copy_gbs},
        "step_roofline": {"bytes_per_step": B, "achieved_GBs": B * steps_s / 1e9,
                          "frac": B * steps_s / 1e9 / HBM_PEAK_GBS},
        "tile_shapes": {"single_field": g.tile_shapes(0), "level_batched": g.level_tile_shapes(0)},
        "preprocess_s": t_pre,
        "chosen": chosen,
        "adjoint": adjoint,
        "ensemble64": ensemble,
        "tandem1024_adjoint_batch8": config5,
        "cpu_baseline": cpu,
        "dst_dense_mfma_vs_fft": dense,
        "cg_form": cg,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.exit(main())
