#!/usr/bin/env python3
"""Headline benchmark: DOF-updates/s of the ADER-DG time step, 3-D Euler,
N=5, 81^3 cells (the precision workload: entropy wave + 8 seeded Fourier
modes, periodic, CFL 0.9), on one MI355X.

A step is one non-fused global time step of the reference
(Riemann on every face -> surface integral + update -> dt = min over
cells -> space-time predictor + volume integral), all on the device with
inputs resident in HBM.  DOF-updates/s = cells * (N+1)^3 * steps / wall time.

Prints ONE JSON line (rank 0).  The CPU baseline times the oracle (the
CPU restatement of the reference's generic kernels) on a bounded sample of
the same workload's cells on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-updates/s, 3D Euler N=5 81^3: FP32-MFMA vs FP64 GPU at L2 err ≤1e-6"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
EULER_FLUX_FLOPS = 40  # flops of one Euler flux evaluation (3 directions)


def predictor_flops(B, V, D, picard_total, cells):
    """Algorithmic FLOPs of the predictor + volume integral kernel:
    per Picard iteration and cell: 6 V B^5 (3 directional Kxi contractions) +
    2 V B^5 (iK1) + (c_F + 4V) B^4 (fluxes, rhs0, residual); per cell once:
    2 (3V + D) B^4 (time averages) + 12 (D + V) B^3 (face extrapolation) +
    6 V B^4 (volume integral)."""
    per_iter = 8 * V * B ** 5 + (EULER_FLUX_FLOPS + 4 * V) * B ** 4
    fixed = 2 * (3 * V + D) * B ** 4 + 12 * (D + V) * B ** 3 + 6 * V * B ** 4
    return per_iter * picard_total + fixed * cells


def cpu_baseline(exa_grid, order, n, seconds, threads):
    """Oracle on a bounded sample of the workload's cells (same data, same dx)."""
    from oracle import oracle as O
    O.lib()
    B = order + 1
    s = O.make_solver(order, 5, 0)
    m = 64
    elapsed = 0.0
    while True:
        luh = exa_grid.cells(0, m)
        g = O.Grid(s, (m, 1, 1), lo=(0.0, 0.0, 0.0), hi=(m / n, 1.0 / n, 1.0 / n))
        g.set_solution(luh)
        t0 = time.perf_counter()
        g.sample_step(m, nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed > 0.25 * seconds or m >= n ** 3:
            break
        m = min(n ** 3, int(m * max(2.0, 0.5 * seconds / max(elapsed, 1e-3))))
    return dict(value=m * B ** 3 / elapsed, unit="DOF-updates/s", cores=threads, kind="port",
                sample=f"one time step (predictor, volume, Riemann on 3 faces/cell, surface, update, dt) "
                       f"of the first {m} cells of the {n}^3 N={order} workload state, "
                       f"oracle (C restatement of the generic kernels), OpenMP {threads} threads, "
                       f"{elapsed:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=81, help="cells per direction")
    ap.add_argument("--order", type=int, default=5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from exahype_amd import capi
    capi.lib()
    order, n = args.order, args.cells
    B, V, D = order + 1, 5, 5
    s = capi.euler_solver(order)
    g = capi.Grid(s, n)
    g.init_euler(rho0=1.0, amp=0.2, k=(1, 1, 1), vel=(1.0, 1.0, 1.0), p=1.0, nmodes=8, mode_amp=0.01,
                 seed=args.seed)
    g.step(args.warmup)
    g.synchronize()

    g.set_profiling(True)
    g.stage_times()  # reset
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    stages = g.stage_times()
    picard_total, pred_cells = g.picard_totals()
    g.set_profiling(False)
    t_now, dt_now, nsteps = g.time()
    if not np.isfinite(dt_now) or dt_now <= 0:
        raise RuntimeError(f"non-finite time step after {nsteps} steps")

    ncells = n ** 3
    dof = ncells * B ** 3
    value = dof * args.steps / wall
    pred_ms, pred_n = stages["predictor"]
    avg_ms = pred_ms / max(pred_n, 1)
    flops_per_launch = predictor_flops(B, V, D, picard_total, pred_cells) / max(pred_n, 1)
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    roofline = dict(bound="mfma", achieved=round(achieved, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / FP64_PEAK_TFLOPS, 4), traffic=None,
                    kernel="predictor_kernel<6,Euler,flux> (predictor + volume integral)",
                    avg_launch_ms=round(avg_ms, 4), flops_per_launch=flops_per_launch,
                    picard_mean=round(picard_total / max(pred_cells, 1), 4))
    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        cpu = cpu_baseline(g, order, n, args.cpu_seconds, threads)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (entropy wave + 8 seeded Fourier modes, generated on the device)",
        "config": {"workload": f"euler3d-precision {n}^3 cells N={order} periodic CFL0.9",
                   "form": "plain FP64", "cells": ncells, "order": order, "dof_per_step": dof,
                   "scalar_unknowns_per_step": dof * V, "seed": args.seed, "t_end": t_now},
        "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in stages.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
