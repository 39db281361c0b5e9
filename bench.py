#!/usr/bin/env python3
"""Throughput of the multi-domain IBFS time step on one MI355X.

Workload (the metric's configuration): 512 x 512 cells per level, 5 levels,
finest domain [-1,3] x [-2,2], stationary NACA0012 (chord 1, 20 deg nose-up
about the quarter chord, IB spacing = grid spacing -> 261 points), Re = 300,
dt = 0.002, impulsive start from rest (synthetic: no data files).  One step
is one call of the reference's advance (src/operators.f90:62-188).

Prints one JSON line: steps/s of the plain FP64 form, the roofline of the
dominant kernel (a 1-D DST-I pass, timed live with HIP events on the
library's stream) and a CPU baseline (the oracle's FP64 step, single
thread, on a bounded sample of the same workload).
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
WORKLOAD = dict(m=512, n=512, mgridlev=5, dt=0.002, re=300.0, len_=4.0, offx=1.0, offy=2.0)


def workload_geometry(m=512, len_=4.0):
    from ibfs_amd import geometry

    return geometry.naca4(len_ / m, aoa_deg=20.0, x_le=0.0)


def step_bytes(m, n, L, nf, s=8):
    """Algorithmic HBM bytes of one step: every 2-D DST-I reads and writes one
    W field (7L+2 per step), step 1 reads q, q0 and reads/writes rhs_old per
    level, coarsify/projection reads, curl writes, and the Cholesky factor
    read (fixed formula, independent of how the kernels are fused)."""
    W = (m - 1) * (n - 1)
    Nq = (m + 1) * n + (n + 1) * m
    return s * (2 * (7 * L + 2) * W + L * (2 * Nq + 2 * W) + (2 * L - 1) * W + (L + 1) * Nq) + 8 * nf * (nf - 1)


def cpu_baseline(xb, yb, bd, cm, cv, budget_s=15.0):
    import oracle as orc

    w = WORKLOAD
    o = orc.Oracle(w["m"], w["n"], w["mgridlev"], w["dt"], w["re"], w["len_"], w["offx"], w["offy"], xb, yb, bd)
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (independent cases shard "
              "one per GPU with no communication; see DESIGN.md)", file=sys.stderr)
        return 2

    from ibfs_amd import hip

    w = WORKLOAD
    xb, yb, bd = workload_geometry(w["m"], w["len_"])
    g = hip.Solver(w["m"], w["n"], w["mgridlev"], w["dt"], w["re"], w["len_"], w["offx"], w["offy"], xb, yb, bd)
    t0 = time.perf_counter()
    cm, cv = g.build_cholesky()
    t_pre = time.perf_counter() - t0
    print(f"# preprocess (device a_times x {g.nf} + choldc): {t_pre:.3f} s", file=sys.stderr, flush=True)

    g.advance(max(args.warmup, 1))
    t0 = time.perf_counter()
    dev_ms = g.advance_timed(args.steps, hist=True)
    wall = time.perf_counter() - t0
    steps_s = args.steps / wall
    print(f"# {args.steps} steps: wall {wall * 1e3:.1f} ms, device {dev_ms:.1f} ms", file=sys.stderr, flush=True)

    # dominant kernel: 1-D DST-I pass over one W field (read W + write W doubles)
    W = g.nw
    pass_bytes = 2 * 8 * W
    t_row = g.time_dst_pass(0, 200)
    t_col = g.time_dst_pass(1, 200)
    t_pass = max(t_row, t_col)
    achieved = pass_bytes / (t_pass * 1e-3) / 1e9
    B = step_bytes(g.m, g.n, g.L, g.nf)
    print(f"# DST pass: row {t_row * 1e3:.2f} us, column {t_col * 1e3:.2f} us; step bytes {B / 1e6:.1f} MB "
          f"-> whole-step {B * steps_s / 1e9:.1f} GB/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        v, n = cpu_baseline(xb, yb, bd, cm, cv, args.cpu_seconds)
        cpu = {"value": v, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{n} steps (itime 2..{n + 1}) of the same 512^2 x 5 NACA0012 case, oracle FP64 step "
                         f"(own radix-4/2/3/5 FFT DST-I, gcc -O3), 1 thread of {os.cpu_count()} "
                         f"({platform.processor() or platform.machine()})"}
        print(f"# CPU baseline: {v:.3f} steps/s over {n} steps -> GPU/CPU {steps_s / v:.1f}x", file=sys.stderr)

    out = {
        "metric": "time steps/s at 512^2 x 5 levels (plain FP64)",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "dst_pass (1-D DST-I over a 511x511 FP64 field)",
                     "launch_us": t_pass * 1e3, "bytes_per_launch": pass_bytes},
        "step_roofline": {"bytes_per_step": B, "achieved_GBs": B * steps_s / 1e9,
                          "frac": B * steps_s / 1e9 / HBM_PEAK_GBS},
        "preprocess_s": t_pre,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.exit(main())
