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
FP32_PEAK_TFLOPS = 157.3
# FP32 FMA rate of the vector ALUs without packed FP32 instructions (which this
# build does not use): 256 CUs x 4 SIMDs x 16 lanes x 2 FLOP x 2.4 GHz
FP32_VALU_PEAK_TFLOPS = 78.6
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


ELASTIC_NCP_FLOPS = 42  # flops of one elastic non-conservative product B(Q) grad Q (3 directions)


def predictor_flops_ncp(B, V, D, picard_total, cells, c_s=ELASTIC_NCP_FLOPS):
    """Algorithmic FLOPs of the predictor + volume integral for a PDE with
    only the non-conservative product (the elastic workload): per Picard
    iteration and cell 6 V B^5 (gradients along 3 directions) + 2 V B^5
    (iK1) + (c_S + 3 V) B^4 (product, rhs, residual); per cell once
    2 (V + D) B^4 (time averages) + 12 (D + V) B^3 (face extrapolation) +
    2 V B^3 (volume integral of the averaged product)."""
    per_iter = 8 * V * B ** 5 + (c_s + 3 * V) * B ** 4
    fixed = 2 * (V + D) * B ** 4 + 12 * (D + V) * B ** 3 + 2 * V * B ** 3
    return per_iter * picard_total + fixed * cells


def predictor_executed_flops_fp32(B, V, picard_total, cells):
    """FLOPs the default FP32-bulk predictor executes for a flux-only PDE (the
    algorithmic count above counts every Picard iteration in full): the first
    iteration evaluates one flux and one slice's Kxi contractions per node
    (c_F + 6 V B) and forms lQi^1 in closed form and its residual (8 V B + 5 V);
    later iterations are full; the epilogue adds the once-per-cell terms and,
    except at B = 5, 6 and 8 where the flux averages are accumulated during
    the iterations, re-evaluates the B slice fluxes of the last iterate
    (c_F B^4)."""
    c_f = EULER_FLUX_FLOPS
    first = B ** 3 * (c_f + 6 * V * B + 8 * V * B + 5 * V)
    per_iter = 8 * V * B ** 5 + (c_f + 4 * V) * B ** 4
    fixed = 2 * (3 * V + V) * B ** 4 + 12 * (V + V) * B ** 3 + 6 * V * B ** 4
    recompute = 0 if B in (5, 6, 8) else c_f * B ** 4
    return cells * first + (picard_total - cells) * per_iter + cells * (fixed + recompute)


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


def run_form(capi, order, n, steps, warmup, seed, precision, variant=0, fused=True):
    """One timed run of the stepping loop; returns timings and the final solution."""
    s = capi.euler_solver(order, precision=precision)
    g = capi.Grid(s, n)
    if variant:
        g.set_predictor_variant(variant)
    g.set_fused_corrector(fused)
    g.init_euler(rho0=1.0, amp=0.2, k=(1, 1, 1), vel=(1.0, 1.0, 1.0), p=1.0, nmodes=8, mode_amp=0.01,
                 seed=seed)
    g.step(warmup)
    g.synchronize()
    g.set_profiling(True)
    g.stage_times()  # reset
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    stages = g.stage_times()
    picard_total, pred_cells = g.picard_totals()
    g.set_profiling(False)
    t_now, dt_now, nsteps = g.time()
    if not np.isfinite(dt_now) or dt_now <= 0:
        raise RuntimeError(f"non-finite time step after {nsteps} steps")
    return dict(grid=g, wall=wall, stages=stages, picard_total=picard_total, pred_cells=pred_cells, t=t_now)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "predictor_euler_n5_81cube_pmc_traffic.txt")


def measured_traffic(form, order, n):
    """HBM bytes per predictor launch measured with rocprofv3 --pmc FETCH_SIZE
    and WRITE_SIZE (each scaled by the same counter's reading for a copy of
    known size), as recorded in the profile summary for this workload; None
    when no measurement of this form and size is recorded."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    key = f"traffic_bytes_per_launch {form} N={order} {n}^3"
    for line in open(TRAFFIC_FILE):
        if line.startswith(key + " "):
            return int(line.split()[-1])
    return None


CORR_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "corrector_euler_n5_81cube_pmc_traffic.txt")


def corrector_roofline(r, form, order, n, B, V, D):
    """HBM roofline of the one-kernel corrector (the step's second kernel):
    per cell it reads luh (double, B^3 D), lduh (B^3 V) and the cell's face
    blocks (6 B^2 (D + V)) in the bulk type and writes luh (B^3 V doubles);
    the neighbours' face blocks it also reads are other cells' own blocks, so
    the minimal traffic counts each block once (`traffic`, when recorded,
    shows the re-reads).  Launch time from the grid's HIP events."""
    ms, nl = r["stages"]["update"]
    avg = ms / max(nl, 1)
    bulk = 4 if form == "fp32" else 8
    alg = n ** 3 * (8 * B ** 3 * D + bulk * (B ** 3 * V + 6 * B * B * (D + V)) + 8 * B ** 3 * V)
    a = alg / (avg * 1e-3) / 1e9
    traffic = None
    if os.path.exists(CORR_TRAFFIC_FILE):
        key = f"traffic_bytes_per_launch corrector {form} N={order} {n}^3"
        for line in open(CORR_TRAFFIC_FILE):
            if line.startswith(key + " "):
                traffic = int(line.split()[-1])
    return dict(bound="hbm", achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(a / HBM_PEAK_GBS, 4),
                traffic=traffic, kernel=f"corrector_kernel<6,Euler,{'float' if form == 'fp32' else 'double'}>",
                avg_launch_ms=round(avg, 4), algorithmic_bytes_per_launch=alg)


def roofline_for(r, B, V, D, peak, kernel, compute_unit, valu_peak=None, executed=None):
    """Compute roofline of the predictor.  `bound` is the contract's category
    (compute, i.e. "mfma", not "hbm"); `compute_unit` names the unit that
    executes the contractions, `valu_frac` the fraction of the vector ALUs'
    peak, `executed_frac` the fraction for the FLOPs the kernel executes."""
    pred_ms, pred_n = r["stages"]["predictor"]
    avg_ms = pred_ms / max(pred_n, 1)
    flops_per_launch = predictor_flops(B, V, D, r["picard_total"], r["pred_cells"]) / max(pred_n, 1)
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    out = dict(bound="mfma", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
               frac=round(achieved / peak, 4), traffic=None, kernel=kernel, compute_unit=compute_unit,
               avg_launch_ms=round(avg_ms, 4), flops_per_launch=flops_per_launch,
               picard_mean=round(r["picard_total"] / max(r["pred_cells"], 1), 4))
    if valu_peak:
        out["valu_peak"] = valu_peak
        out["valu_frac"] = round(achieved / valu_peak, 4)
    if executed is not None:
        ex = executed(r["picard_total"], r["pred_cells"]) / max(pred_n, 1)
        out["executed_flops_per_launch"] = ex
        out["executed_frac"] = round(ex / (avg_ms * 1e-3) / 1e12 / peak, 4)
    return out


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}), flush=True)


def cpu_baseline_limiter(n, seconds, threads, gpu_fv_fraction, warmup, steps):
    """The oracle's limiting driver (OpenMP over cells) on the octant of the
    n^3 blast that holds the blast centre: the (n+1)/2 cells per direction
    from the centre cell to the +x,+y,+z walls, at the workload's resolution,
    so that the FV-active fraction of the sample follows the whole grid's;
    the GPU run's warmup steps untimed, then its timed steps (fewer when the
    time budget runs out), so that the sample covers the same stretch of the
    blast's evolution (rate per DOF-update, an intensive figure)."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refutil import blast_cells
    order = 3
    i0 = n // 2
    m = n - i0
    s = O.make_solver(order)
    nodes = np.array([O.table(O.T_NODES, order, i) for i in range(order + 1)])
    g = O.LimitingGrid(s, m, lo=(i0 / n,) * 3, hi=(1.0,) * 3, periodic=False)
    g.set_solution(blast_cells(order, n, nodes=nodes, radius=0.2, box=(i0, m)),
                   blast_cells(order, n, subcells=True, radius=0.2, box=(i0, m)))
    if warmup:
        g.step(warmup)
    st0 = g.status()
    done, t0 = 0, time.perf_counter()
    while done < 1 or (done < steps and time.perf_counter() - t0 < seconds):
        g.step(1)
        done += 1
    el = time.perf_counter() - t0
    st = g.status()
    frac = 0.5 * (float((st0 >= 2).mean()) + float((st >= 2).mean()))
    return dict(value=m ** 3 * (order + 1) ** 3 * done / el, unit="DOF-updates/s", cores=threads, kind="port",
                fv_active_fraction=frac,
                sample=f"steps {warmup + 1}..{warmup + done} of the limiting driver (oracle, OpenMP {threads} threads) "
                       f"on the {m}^3-cell octant of the {n}^3 N=3 blast holding its centre (transmissive walls), "
                       f"{el:.1f} s; FV-active fraction {frac:.4f} (mean of before and after) against "
                       f"{gpu_fv_fraction:.4f} of the GPU grid over the same steps")


LIM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "limiter_status_blast_n3_81cube_pmc_traffic.txt")


def limiter_measured_traffic(n):
    """HBM bytes per status-kernel launch on the limiter workload (rocprofv3
    --pmc, copy-calibrated), as recorded for this grid; None otherwise."""
    if not os.path.exists(LIM_TRAFFIC_FILE):
        return None
    key = f"traffic_bytes_per_launch lim_update_status N=3 {n}^3"
    for line in open(LIM_TRAFFIC_FILE):
        if line.startswith(key + " "):
            return int(line.split()[-1])
    return None


def limiter_rooflines(stages, n, B, V, fv_frac, troubled_frac, picard):
    """Rooflines of the limiter step's three largest kernels.
    Status kernel (one launch per step over every cell; DG surface integral
    and update, stable dt, DMP, status): per DG cell it reads luh, lduh (B^3 V
    each), lFhbnd (6 B^2 V) and the merged extrema (2 x 6 V), writes luh and
    luh_prev (B^3 V each), the local extrema (2 V), the cell dt and two face
    blocks of extrema (4 V); per FV-active cell it reads luh, the extrema and
    the patch extrema (2 V) and writes the local extrema, the dt and all six
    face blocks (12 V); about 76 bytes of statuses, counters and neighbour
    indices per cell either way.
    FV kernel (one launch per step over the FV-active cells): per cell it
    reads the patch with its ghost layers (S^3 V, S = 2N+1+4) and writes the
    new patch interior (L^3 V, L = 2N+1), its projection to DG (B^3 V) and
    its extrema (2 V).
    Predictor (one launch per step; FP64 Picard + volume integral on the
    cells below troubled): the algorithmic FLOPs as for the precision
    workload's FP64 form."""
    cells = n ** 3
    B3, B2 = B ** 3, B * B
    L = 2 * (B - 1) + 1
    S = L + 4
    dg = 8 * (4 * B3 * V + 6 * B2 * V + 12 * V + 2 * V + 1 + 4 * V) + 76
    fv = 8 * (B3 * V + 12 * V + 2 * V + 2 * V + 1 + 12 * V) + 76
    status_bytes = cells * ((1 - fv_frac) * dg + fv_frac * fv)
    fv_bytes = cells * fv_frac * 8 * (S ** 3 * V + L ** 3 * V + B3 * V + 2 * V)
    out = []
    ms, nl = stages["update"]
    avg = ms / max(nl, 1)
    a = status_bytes / (avg * 1e-3) / 1e9
    out.append(dict(bound="hbm", achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(a / HBM_PEAK_GBS, 4), traffic=limiter_measured_traffic(n),
                    kernel="lim_update_status_kernel<3> (DG update, stable dt, DMP, status)",
                    avg_launch_ms=round(avg, 4), algorithmic_bytes_per_launch=int(status_bytes)))
    ms, nl = stages["predictor"]
    avg = ms / max(nl, 1)
    it, pc = picard
    run_cells = pc * (1 - troubled_frac)
    flops = predictor_flops(B, V, V, it, run_cells) / max(nl, 1)
    a = flops / (avg * 1e-3) / 1e12
    out.append(dict(bound="mfma", achieved=round(a, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(a / FP64_PEAK_TFLOPS, 4), traffic=None, compute_unit="VALU",
                    kernel="predictor_kernel<4,Euler,flux,double> (Picard + volume integral)",
                    avg_launch_ms=round(avg, 4), flops_per_launch=flops,
                    picard_mean=round(it / max(run_cells, 1), 4)))
    ms, nl = stages["fv"]
    avg = ms / max(nl, 1)
    a = fv_bytes / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
    out.append(dict(bound="hbm", achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(a / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="lim_fv_update_kernel<3> (MUSCL-Hancock patch update + projection)",
                    avg_launch_ms=round(avg, 4), algorithmic_bytes_per_launch=int(fv_bytes)))
    return out


def bench_limiter(capi, args):
    """Limiter workload: Euler blast, N=3, 81^3, transmissive, FP64."""
    order, n = 3, args.cells
    B, V = order + 1, 5
    g = capi.LimitingGrid(capi.euler_solver(order), n)
    g.init_blast(seed=args.seed, radius=0.2)
    g.step(args.warmup)
    g.synchronize()
    st0, _ = g.status()
    g.set_profiling(True)
    g.stage_times()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    stages = g.stage_times()
    picard = g.picard_totals()
    st, irr = g.status()
    t_now, dt_now, nsteps = g.time()
    if not np.isfinite(dt_now) or dt_now <= 0:
        raise RuntimeError("non-finite time step")
    dof = n ** 3 * B ** 3
    # statuses over the timed steps: the mean of the fractions before and after
    fv_frac = 0.5 * (float((st0 >= 2).mean()) + float((st >= 2).mean()))
    tr_frac = 0.5 * (float((st0 == 3).mean()) + float((st == 3).mean()))
    roofs = limiter_rooflines(stages, n, B, V, fv_frac, tr_frac, picard)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    out = {
        "metric": "DOF-updates/s, 3D Euler blast N=3 81^3 with a-posteriori subcell limiter (FP64)",
        "value": dof * args.steps / wall, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (spherical blast + 1% seeded density noise, generated on the device)",
        "config": {"workload": f"euler3d-limiter-blast {n}^3 cells N={order} transmissive, MUSCL-Hancock 7^3 "
                               f"subcells, DMP 1e-4/1e-3, helper layers 1", "cells": n ** 3,
                   "dof_per_step": dof, "seed": args.seed},
        "stage_ms_per_step": {k: round(v[0] / max(args.steps, 1), 4) for k, v in stages.items()},
        "limited": {"fv_active_fraction": float((st >= 2).mean()), "troubled_fraction": float((st == 3).mean()),
                    "fv_active_fraction_mean_timed": fv_frac, "irregular_steps_total": int(irr)},
        "roofline": roofs[0],
        "rooflines_other_kernels": roofs[1:],
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_limiter(n, args.cpu_seconds, threads,
                                                                               fv_frac, args.warmup, args.steps),
    }
    print(json.dumps(out), flush=True)
    g.close()


FV_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fv_godunov_54cube_L10_pmc_traffic.txt")


def fv_measured_traffic(n, L):
    """HBM bytes per FV update launch from rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (copy-calibrated), as recorded for this grid; None otherwise."""
    if not os.path.exists(FV_TRAFFIC_FILE):
        return None
    key = f"traffic_bytes_per_launch fv_godunov {n}^3 L={L}"
    for line in open(FV_TRAFFIC_FILE):
        if line.startswith(key + " "):
            return int(line.split()[-1])
    return None


def cpu_baseline_fv(L, seconds):
    """The oracle's finite-volumes driver (patches with ghost layers, ghost
    filling, Godunov update, stable dt), single thread, on a periodic grid of
    6^3 cells of L^3 subcells (rate per subcell update, an intensive figure)."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refutil import fv_grid_state, fv_smooth
    m = 6
    g = O.FVGrid(O.make_solver(3, cfl=0.9), O.FV_GODUNOV, L, m, periodic=True)
    g.set_solution(fv_grid_state((m, m, m), L, fv_smooth)[1])
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        g.step(1)
        steps += 1
    el = time.perf_counter() - t0
    return dict(value=m ** 3 * L ** 3 * steps / el, unit="DOF-updates/s (subcells)", cores=1, kind="port",
                sample=f"{steps} steps of the FV driver (oracle) on {m}^3 cells of {L}^3 subcells, periodic smooth "
                       f"Euler state, 1 thread, {el:.1f} s")


AMR_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "amr_volume_n5_27to81_pmc_traffic.txt")


def amr_measured_traffic(kind):
    """HBM bytes per AMR transfer launch (rocprofv3 --pmc, copy-calibrated) as
    recorded for the 27^3 -> 81^3 workload; None otherwise."""
    if not os.path.exists(AMR_TRAFFIC_FILE):
        return None
    key = f"traffic_bytes_per_launch amr_{kind} 27^3->81^3 N=5"
    for line in open(AMR_TRAFFIC_FILE):
        if line.startswith(key + " "):
            return int(line.split()[-1])
    return None


def amr_items(nc, r):
    """Transfers of an nc^3 coarse grid refined r times per direction: rows
    (fine block, coarse block, sx, sy, sz), coarse-major, and the CSR pointer
    of each coarse block's children."""
    nf = nc * r
    c = np.arange(nc ** 3)
    ci, cj, ck = c % nc, (c // nc) % nc, c // (nc * nc)
    s = np.arange(r ** 3)
    sx, sy, sz = s % r, (s // r) % r, s // (r * r)
    fi = (ci[:, None] * r + sx[None, :])
    fj = (cj[:, None] * r + sy[None, :])
    fk = (ck[:, None] * r + sz[None, :])
    items = np.stack([fi + nf * (fj + nf * fk), np.repeat(c[:, None], r ** 3, 1),
                      np.broadcast_to(sx, fi.shape), np.broadcast_to(sy, fi.shape), np.broadcast_to(sz, fi.shape)],
                     -1).reshape(-1, 5).astype(np.int32)
    return items, np.arange(0, nc ** 3 * r ** 3 + 1, r ** 3, dtype=np.int32)


def cpu_baseline_amr(seconds):
    """The oracle's volume prolongation + restriction (single calls, as the
    reference's grid traversal makes them), one thread, N=5, 5 variables."""
    from oracle import oracle as O
    s = O.make_solver(5)
    rng = np.random.default_rng(1)
    uc, uf = rng.standard_normal(216 * 5), rng.standard_normal(216 * 5)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        sub = (n % 3, (n // 3) % 3, (n // 9) % 3)
        O.volume_unknowns_prolongation(s, uc, 0, 1, sub)
        uc = O.volume_unknowns_restriction(s, uc, uf, 0, 1, sub)
        n += 1
    el = time.perf_counter() - t0
    return dict(value=2 * n * 216 * 5 / el, unit="DOF-transfers/s", cores=1, kind="port",
                sample=f"{n} volume prolongations + {n} restrictions (oracle single calls, N=5, 5 variables, one "
                       f"level), 1 thread, {el:.1f} s")


def bench_amr(capi, args):
    """AMR workload: every cell of a 27^3 Euler grid at N=5 refined once (81^3
    fine cells = 531,441 volume prolongations), then all fine cells restricted
    back (27 children added per coarse cell), FP64, device-resident blocks."""
    nc, r, B, D = 27, 3, 6, 5
    s = capi.euler_solver(5)
    rng = np.random.default_rng(args.seed)
    items, ptr = amr_items(nc, r)
    coarse = capi.DeviceArray(rng.standard_normal((nc ** 3, B ** 3 * D)))
    fine = capi.DeviceArray(np.zeros((len(items), B ** 3 * D)))
    d_items, d_ptr = capi.DeviceArray(items), capi.DeviceArray(ptr)
    d_ritems = capi.DeviceArray(items[:, [1, 0, 2, 3, 4]])
    acc = capi.DeviceArray(np.zeros((nc ** 3, B ** 3 * D)))
    for _ in range(args.warmup):
        capi.volume_prolongation_batch(s, fine, coarse, d_items, 1)
        capi.volume_restriction_batch(s, acc, fine, d_ptr, d_ritems, 1)
    tp, tr = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tp.append(capi.volume_prolongation_batch(s, fine, coarse, d_items, 1))
        tr.append(capi.volume_restriction_batch(s, acc, fine, d_ptr, d_ritems, 1))
    wall = time.perf_counter() - t0
    # parity on a sample: restriction of a prolongated field returns the coarse data
    chk = capi.DeviceArray(np.zeros((nc ** 3, B ** 3 * D)))
    capi.volume_restriction_batch(s, chk, fine, d_ptr, d_ritems, 1)
    err = float(np.abs(chk.get() - coarse.get()).max())
    nf = len(items)
    dof = nf * B ** 3 * D
    p_ms, r_ms = float(np.mean(tp)), float(np.mean(tr))
    blk = B ** 3 * D * 8
    alg_p = nf * blk + nc ** 3 * blk          # write every fine block, read every coarse block
    alg_r = nf * blk + 2 * nc ** 3 * blk      # read every fine block, read and write every coarse block
    achieved = alg_p / (p_ms * 1e-3) / 1e9
    out = {
        "metric": "DOF-transfers/s, 3D ADER-DG AMR volume prolongation + restriction, 27^3 -> 81^3 cells, N=5, FP64",
        "value": 2 * dof * args.steps / wall, "unit": "DOF-transfers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded normal coarse data, generated on the host, resident on the device)",
        "config": {"workload": "amr volume transfers: 27^3 coarse cells, one level (81^3 = 531441 fine cells), N=5, "
                               "5 variables; a step = all prolongations + all restrictions",
                   "transfers_per_step": 2 * nf, "seed": args.seed},
        "kernel_ms": {"prolongation": round(p_ms, 4), "restriction": round(r_ms, 4)},
        "restrict_after_prolong_max_err": err,
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 4), traffic=amr_measured_traffic("prolongation"),
                         kernel="amr_kernel<6,3,false> (volume prolongation, one level)",
                         avg_launch_ms=round(p_ms, 4), algorithmic_bytes_per_launch=alg_p,
                         restriction={"avg_launch_ms": round(r_ms, 4), "algorithmic_bytes_per_launch": alg_r,
                                      "achieved_GBs": round(alg_r / (r_ms * 1e-3) / 1e9, 1),
                                      "traffic": amr_measured_traffic("restriction")}),
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_amr(args.cpu_seconds),
    }
    print(json.dumps(out), flush=True)
    for a in (coarse, fine, d_items, d_ptr, d_ritems, acc, chk):
        a.close()


OUTPUT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "output_plot_vertices_n5_81cube_pmc_traffic.txt")


def output_measured_traffic():
    """HBM bytes per vertex-interpolation launch (rocprofv3 --pmc, copy-calibrated); None otherwise."""
    if not os.path.exists(OUTPUT_TRAFFIC_FILE):
        return None
    for line in open(OUTPUT_TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch plot_vertices 81^3 N=5 "):
            return int(line.split()[-1])
    return None


def bench_output(capi, args):
    """Output workload: the Cartesian VTK plotter's interpolation of a 81^3
    Euler grid at N=5 (every cell's 6^3 equidistant vertex values) on the
    device, FP64; a step = one projection of the whole grid."""
    n, order = 81, 5
    B, D = order + 1, 5
    g = capi.Grid(capi.euler_solver(order), n)
    g.init_euler()
    g.step(1)
    g.synchronize()
    for _ in range(args.warmup):
        g.plot_vertices()
    ts = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts.append(g.plot_vertices(timed=True)[1])
    wall = time.perf_counter() - t0  # includes the copy of the 4.6 GB result to the host
    k_ms = float(np.mean(ts))
    vals = n ** 3 * B ** 3 * D
    alg = 2 * vals * 8  # read luh, write the vertex values
    achieved = alg / (k_ms * 1e-3) / 1e9
    out = {
        "metric": "vertex values/s, Cartesian VTK plotter interpolation on the device, 3D Euler N=5 81^3, FP64",
        "value": vals / (k_ms * 1e-3), "unit": "vertex values/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": k_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (entropy wave initial condition generated on the device, one time step)",
        "config": {"workload": f"plot vertices {n}^3 cells N={order}", "values": vals,
                   "wall_ms_per_step_incl_host_copy": round(wall / args.steps * 1e3, 3)},
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 4), traffic=output_measured_traffic(),
                         kernel="plot_vertices_kernel<6> (equidistant vertex interpolation)",
                         avg_launch_ms=round(k_ms, 4), algorithmic_bytes_per_launch=alg),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    g.close()


def bench_fv(capi, args):
    """Finite-volumes workload: the EulerFV demonstrator's Godunov solver in
    3-D, 54^3 cells of 10^3 subcells (540^3), periodic entropy wave + 8
    seeded modes, FP64."""
    n, L = args.fv_cells, args.patch_size
    g = capi.FVGrid(capi.euler_solver(3), L, n)
    g.init_euler(rho0=1.0, amp=0.2, k=(1, 1, 1), vel=(1.0, 1.0, 1.0), p=1.0, nmodes=8, mode_amp=0.01,
                 seed=args.seed)
    g.step(args.warmup)
    g.synchronize()
    g.set_profiling(True)
    g.stage_times()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    stages = g.stage_times()
    t_now, dt_now, _ = g.time()
    if not np.isfinite(dt_now) or dt_now <= 0:
        raise RuntimeError("non-finite time step")
    sub = (n * L) ** 3
    upd_ms, upd_n = stages["update"]
    avg_ms = upd_ms / max(upd_n, 1)
    alg_bytes = sub * 5 * 8 * 2  # read u, write u_new (variables in double)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "DOF-updates/s (subcell updates), 3D Euler finite volumes (Godunov/Rusanov), 54^3 cells x 10^3 "
                  "subcells, FP64",
        "value": sub * args.steps / wall, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (entropy wave + 8 seeded Fourier modes at subcell centres, generated on the device)",
        "config": {"workload": f"euler3d-fv-godunov {n}^3 cells, patch size {L} ({n * L}^3 subcells), periodic, "
                               "CFL 0.9", "subcells": sub, "seed": args.seed},
        "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in stages.items()},
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 4), traffic=fv_measured_traffic(n, L),
                         kernel="fv_godunov_kernel (Godunov update + stable dt of the new solution)",
                         avg_launch_ms=round(avg_ms, 4), algorithmic_bytes_per_launch=alg_bytes),
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_fv(L, args.cpu_seconds),
    }
    print(json.dumps(out), flush=True)
    g.close()


def cpu_baseline_batch(specs, threads, seconds):
    """The oracle's grid time step per order on a 9^3 patch of that order,
    weighted by the workload's DOF mix: the estimated CPU rate on the same
    patch set (rate per DOF-update, an intensive figure)."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refutil import euler_state, node_coords
    O.lib()
    per_dof_s = {}
    budget = seconds / 6.0
    for order in sorted({o for _, o, _, _ in specs}):
        m = 9
        g = O.Grid(O.make_solver(order, 5, 0), m)
        X = node_coords(order, m)
        g.set_solution(euler_state(1.0 + 0.2 * np.sin(2 * np.pi * X.sum(-1)), (1.0, 1.0, 1.0), 1.0).reshape(-1))
        steps, t0 = 0, time.perf_counter()
        while steps < 1 or (time.perf_counter() - t0 < budget and steps < 50):
            g.step(1, nthreads=threads)
            steps += 1
        per_dof_s[order] = (time.perf_counter() - t0) / (steps * m ** 3 * (order + 1) ** 3)
    dof = [n ** 3 * (o + 1) ** 3 for n, o, _, _ in specs]
    est = sum(d * per_dof_s[o] for d, (_, o, _, _) in zip(dof, specs))
    return dict(value=sum(dof) / est, unit="DOF-updates/s", cores=threads, kind="port",
                sample="oracle grid steps of one 9^3 patch per order (entropy wave, periodic), OpenMP "
                       f"{threads} threads, about {budget:.1f} s per order; per-order seconds per DOF-update "
                       "weighted by the patch set's DOF mix")


def batch_profile(capi, specs, precision, steps, hbm_budget):
    """A profiled run of the patch set (timed steps eager, events around every
    corrector and predictor launch): per group (PDE, order) the predictor's
    roofline from the counted Picard iterations, the corrector time, and the
    share of the step the launches leave idle (what one persistent launch
    over all groups could at most recover)."""
    ps = capi.PatchSet(specs, precision=precision, hbm_budget_gb=hbm_budget)
    ps.set_profiling(True)
    ps.run(steps, use_graph=False, autotune=True, warmup=1)
    sec, _, _ = ps.stats()
    peak = FP32_PEAK_TFLOPS if precision == capi.PRECISION_FP32_BULK else FP64_PEAK_TFLOPS
    groups, busy = [], 0.0
    for pde in (capi.PDE_EULER, capi.PDE_ELASTIC):
        for order in range(2, 8):
            pred_ms, corr_ms, launches, it, cells = ps.group_times(pde, order)
            if launches == 0:
                continue
            B = order + 1
            if pde == capi.PDE_EULER:
                flops = predictor_flops(B, 5, 5, it, cells)
            else:
                flops = predictor_flops_ncp(B, 9, 12, it, cells)
            avg = pred_ms / launches
            a = flops / launches / (avg * 1e-3) / 1e12
            busy += pred_ms + corr_ms
            groups.append(dict(pde="euler" if pde == capi.PDE_EULER else "elastic", order=order,
                               predictor_ms_per_step=round(pred_ms / steps, 4),
                               corrector_ms_per_step=round(corr_ms / steps, 4), avg_launch_ms=round(avg, 4),
                               flops_per_launch=flops / launches, achieved_tflops=round(a, 3),
                               frac=round(a / peak, 4), picard_mean=round(it / max(cells, 1), 4),
                               variant=ps.predictor_variant(order, pde)))
    ps.close()
    dom = max(groups, key=lambda g: g["predictor_ms_per_step"])
    roof = dict(bound="mfma", achieved=dom["achieved_tflops"], peak=peak, unit="TFLOP/s", frac=dom["frac"],
                traffic=None, compute_unit="VALU",
                kernel=f"predictor of the {dom['pde']} N={dom['order']} group (the largest predictor time)",
                avg_launch_ms=dom["avg_launch_ms"], flops_per_launch=dom["flops_per_launch"],
                picard_mean=dom["picard_mean"])
    idle = max(0.0, 1.0 - busy * 1e-3 / sec) if sec > 0 else None
    return roof, groups, dict(profiled_ms_per_step=round(sec / steps * 1e3, 4),
                              kernel_ms_per_step=round(busy / steps, 4),
                              idle_fraction=round(idle, 4) if idle is not None else None)


def bench_batch(capi, args):
    """Patch-set workload: 96 seeded periodic Euler patches, sides in
    {9,15,27,45,81}, orders 2..7, entropy waves, per-patch dt; FP32-bulk vs FP64."""
    specs = capi.draw_patches(args.seed)
    keep = 27 ** 3
    res = {}
    for form in ("fp64", "fp32"):
        prec = capi.PRECISION_FP64 if form == "fp64" else capi.PRECISION_FP32_BULK
        t0 = time.perf_counter()
        ps = capi.PatchSet(specs, precision=prec, hbm_budget_gb=args.hbm_budget, keep_max_cells=keep)
        ps.run(args.steps, use_graph=True, autotune=True, warmup=args.warmup)
        wall = time.perf_counter() - t0
        sec, dof, nwaves = ps.stats()
        times = [ps.time(j) for j in range(len(specs))]
        if not all(np.isfinite(dt) and dt > 0 for _, dt in times):
            raise RuntimeError(f"{form}: non-finite time step in the patch set")
        sols = {j: ps.solution(j) for j, sp in enumerate(specs) if sp[0] ** 3 <= keep}
        variants = {o: ps.predictor_variant(o) for o in range(2, 8)}
        ps.close()
        res[form] = dict(sec=sec, dof=dof, nwaves=nwaves, wall=wall, sols=sols, variants=variants, times=times)
        print(f"batch {form}: {nwaves} wave(s), {dof / sec:.3e} DOF/s, {wall:.1f} s with setup", file=sys.stderr,
              flush=True)
    num = sum(float(np.sum((res["fp32"]["sols"][j] - res["fp64"]["sols"][j]) ** 2)) for j in res["fp64"]["sols"])
    den = sum(float(np.sum(res["fp64"]["sols"][j] ** 2)) for j in res["fp64"]["sols"])
    m, m64 = res["fp32"], res["fp64"]
    # the two-pass corrector (face pass + update pass) on the FP32-bulk form
    ps = capi.PatchSet(specs, precision=capi.PRECISION_FP32_BULK, hbm_budget_gb=args.hbm_budget)
    ps.set_fused_corrector(False)
    ps.run(args.steps, use_graph=True, autotune=True, warmup=args.warmup)
    sec2, dof2, _ = ps.stats()
    ps.close()
    roof, groups, launch = batch_profile(capi, specs, capi.PRECISION_FP32_BULK, 2, args.hbm_budget)
    graph_ms = m["sec"] / args.steps * 1e3
    launch["graph_ms_per_step"] = round(graph_ms, 4)
    # the share of the graph step not covered by corrector and predictor
    # launches (dt reductions and gaps): the most one persistent launch over
    # all groups could recover
    launch["idle_fraction_vs_graph"] = round(max(0.0, 1.0 - launch["kernel_ms_per_step"] / graph_ms), 4)
    print(f"batch profile: {launch}", file=sys.stderr, flush=True)
    cells = sum(n ** 3 for n, *_ in specs)
    dof_step = sum(n ** 3 * (o + 1) ** 3 for n, o, _, _ in specs)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    out = {
        "metric": "DOF-updates/s, 96 uneven 3D Euler patches (orders 2..7), per-patch dt, FP32-bulk vs FP64",
        "value": m["dof"] / m["sec"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": m["sec"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic (96 seeded patches, entropy waves generated on the device)",
        "config": {"workload": f"euler3d-patchset 96 patches, sides {{9,15,27,45,81}}, orders 2..7, "
                               f"{cells} cells, {dof_step} DOF per step, periodic CFL0.9",
                   "seed": args.seed, "hbm_budget_gb": args.hbm_budget, "waves": m["nwaves"],
                   "launches": "one per stage and order group, captured once per wave in a hipGraph",
                   "corrector": "one kernel per group (Riemann + surface + update + dt)",
                   "predictor_variant_by_order": m["variants"]},
        "two_pass_corrector_fp32": {"value": dof2 / sec2, "ms_per_step": sec2 / args.steps * 1e3},
        "groups_fp32": groups,
        "launch_overhead_fp32": launch,
        "wall_s_including_setup": round(m["wall"], 3),
        "fp64": {"value": m64["dof"] / m64["sec"], "ms_per_step": m64["sec"] / args.steps * 1e3,
                 "waves": m64["nwaves"], "predictor_variant_by_order": m64["variants"],
                 "wall_s_including_setup": round(m64["wall"], 3)},
        "fp32_vs_fp64": {"speedup": m64["sec"] / m["sec"],
                         "l2_rel_err": (num / den) ** 0.5 if den > 0 else None,
                         "l2_over": f"the {len(m64['sols'])} patches with at most {keep} cells",
                         "max_rel_time_diff": max(abs(a[0] - b[0]) / b[0] for a, b in zip(m["times"], m64["times"]))},
        "roofline": roof,
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_batch(specs, threads, args.cpu_seconds),
    }
    print(json.dumps(out), flush=True)


LINEAR_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "linear_elastic_n4_81cube_pmc_traffic.txt")


def linear_predictor_traffic():
    """HBM bytes per linear-predictor launch on the elastic workload (rocprofv3
    --pmc, copy-calibrated); None otherwise."""
    if not os.path.exists(LINEAR_TRAFFIC_FILE):
        return None
    for line in open(LINEAR_TRAFFIC_FILE):
        if line.startswith("| linear predictor |"):
            return int(line.split("|")[4].strip().replace(",", ""))
    return None


def cpu_baseline_elastic(seconds, threads=16):
    """The oracle's grid driver (nonlinear path with the NCP, and the linear
    kernels) on a periodic 6^3 sample of the elastic workload at N=4, OpenMP
    threads; steps until about seconds/2 per form."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refutil import elastic_cells
    order, m = 4, 6
    B = order + 1
    nodes = np.array([O.table(O.T_NODES, order, i) for i in range(B)])
    luh = elastic_cells(order, m, seed=2024, nodes=nodes, width=0.05)
    out = {}
    for linear in (False, True):
        s = O.elastic_solver(order)
        s.linear = int(linear)
        g = O.Grid(s, m)
        g.set_solution(luh)
        g.step(1, nthreads=threads)  # the initial prediction
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds / 2:
            g.step(1, nthreads=threads)
            steps += 1
        el = time.perf_counter() - t0
        out["linear" if linear else "nonlinear"] = (m ** 3 * B ** 3 * steps / el, steps, el)
    v, st, el = out["nonlinear"]
    return dict(value=v, unit="DOF-updates/s", cores=threads, kind="port",
                linear_value=out["linear"][0],
                sample=f"{st} steps of the oracle's grid driver (nonlinear, NCP) on a periodic {m}^3 sample of the "
                       f"elastic workload at N={order}, OpenMP {threads} threads, {el:.1f} s; the linear kernels "
                       f"{out['linear'][1]} steps in {out['linear'][2]:.1f} s (linear_value)")


def linear_roofline(stages, n, B):
    """Roofline of the linear predictor (the linear form's dominant kernel): it
    reads luh (B^3 x 12 doubles) and writes lduh (B^3 x 9) and both face arrays
    (6 B^2 x (12 + 9)) per cell; launch time from the grid's HIP events."""
    ms, launches = stages["predictor"]
    avg = ms / max(launches, 1)
    alg = n ** 3 * (B ** 3 * 12 + B ** 3 * 9 + 6 * B * B * 21) * 8
    achieved = alg / (avg * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=linear_predictor_traffic(),
                kernel="linear_predictor_kernel<5, Elastic> (Cauchy-Kowalevski + linear volume integral)",
                avg_launch_ms=round(avg, 4), algorithmic_bytes_per_launch=alg,
                note="latency-bound on LDS (neither HBM nor ALU saturated)")


def elastic_roofline(r, B, peak, valu_peak, kernel):
    """Compute roofline of the elastic predictor (the nonlinear form's
    dominant kernel), FLOPs from predictor_flops_ncp and the counted Picard
    iterations; launch time from the grid's HIP events."""
    ms, nl = r["stages"]["predictor"]
    avg = ms / max(nl, 1)
    flops = predictor_flops_ncp(B, 9, 12, r["picard_total"], r["pred_cells"]) / max(nl, 1)
    a = flops / (avg * 1e-3) / 1e12
    out = dict(bound="mfma", achieved=round(a, 3), peak=peak, unit="TFLOP/s", frac=round(a / peak, 4),
               traffic=None, kernel=kernel, compute_unit="VALU", avg_launch_ms=round(avg, 4),
               flops_per_launch=flops, picard_mean=round(r["picard_total"] / max(r["pred_cells"], 1), 4))
    if valu_peak:
        out["valu_peak"] = valu_peak
        out["valu_frac"] = round(a / valu_peak, 4)
    return out


def bench_elastic(capi, args):
    """Elastic workload: velocity-stress NCP form, N=4, 81^3, Gaussian pulse
    of width 0.05 in a seeded two-layer material; FP32-bulk vs FP64, and the
    same problem through the linear kernels (Cauchy-Kowalevski predictor,
    linear Riemann solver) in FP64."""
    order, n = 4, args.cells
    B = order + 1
    res = {}
    forms = ("fp64", "fp32", "linear") if args.form == "both" else (args.form,)
    for form in forms:
        prec = capi.PRECISION_FP32_BULK if form == "fp32" else capi.PRECISION_FP64
        g = capi.Grid(capi.elastic_solver(order, precision=prec, linear=form == "linear"), n)
        g.init_elastic(seed=args.seed, width=0.05)
        g.step(args.warmup)
        g.synchronize()
        g.set_profiling(True)
        g.stage_times()
        t0 = time.perf_counter()
        g.step(args.steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        stages = g.stage_times()
        t_now, dt_now, _ = g.time()
        if not np.isfinite(dt_now) or dt_now <= 0:
            raise RuntimeError(f"{form}: non-finite time step")
        it, pc = g.picard_totals() if form != "linear" else (0, 0)
        res[form] = dict(wall=wall, stages=stages, t=t_now, sol=g.solution().reshape(-1, 12)[:, :9],
                         picard_total=it, pred_cells=pc)
        print(f"elastic {form}: {wall / max(args.steps, 1) * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
        g.close()
    dof = n ** 3 * B ** 3
    if len(forms) == 1:  # one form (profiling runs): its own timing only
        r = res[forms[0]]
        print(json.dumps({"workload": "elastic", "form": forms[0], "steps": args.steps,
                          "ms_per_step": r["wall"] / args.steps * 1e3, "value": dof * args.steps / r["wall"],
                          "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in r["stages"].items()}}),
              flush=True)
        return
    a, b = res["fp32"]["sol"], res["fp64"]["sol"]
    l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    m, m64 = res["fp32"], res["fp64"]
    # the same problem as one patch of the batching scheduler (exa_patchset)
    patch = {}
    for form in ("fp64", "fp32"):
        prec = capi.PRECISION_FP32_BULK if form == "fp32" else capi.PRECISION_FP64
        ps = capi.PatchSet([capi.elastic_patch(n, order, args.seed, width=0.05)], precision=prec,
                           hbm_budget_gb=args.hbm_budget, keep_max_cells=n ** 3)
        ps.run(args.steps, use_graph=True, autotune=True, warmup=args.warmup)
        sec, pdof, _ = ps.stats()
        sol = ps.solution(0).reshape(-1, 12)[:, :9]
        ps.close()
        ref = res[form]["sol"]
        patch[form] = {"value": pdof / sec, "ms_per_step": sec / args.steps * 1e3,
                       "l2_rel_diff_vs_grid": float(np.linalg.norm(sol - ref) / np.linalg.norm(ref))}
        print(f"elastic patch set {form}: {patch[form]}", file=sys.stderr, flush=True)
    out = {
        "metric": "DOF-updates/s, 3D linear elasticity (NCP form) N=4 81^3: FP32-bulk vs FP64",
        "value": dof * args.steps / m["wall"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": m["wall"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic (Gaussian velocity pulse, seeded two-layer material, generated on the device)",
        "config": {"workload": f"elastic3d {n}^3 cells N={order} periodic CFL0.9, 9 variables + 3 parameters",
                   "seed": args.seed, "dof_per_step": dof},
        "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in m["stages"].items()},
        "fp64": {"value": dof * args.steps / m64["wall"], "ms_per_step": m64["wall"] / args.steps * 1e3,
                 "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in m64["stages"].items()}},
        "fp32_vs_fp64": {"speedup": m64["wall"] / m["wall"], "l2_rel_err": l2,
                         "steps_compared": args.warmup + args.steps},
        "patchset": patch,
        # the linear kernels on the same problem: a different scheme in its
        # face states (face-averaged coefficient matrix), so the difference to
        # the nonlinear FP64 run is a discretisation difference, not an error
        "linear_fp64": {"value": dof * args.steps / res["linear"]["wall"],
                        "ms_per_step": res["linear"]["wall"] / args.steps * 1e3,
                        "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4)
                                              for k, v in res["linear"]["stages"].items()},
                        "l2_rel_diff_vs_nonlinear_fp64": float(np.linalg.norm(res["linear"]["sol"] - b)
                                                               / np.linalg.norm(b)),
                        "roofline": linear_roofline(res["linear"]["stages"], n, B)},
        "roofline": elastic_roofline(m, B, FP32_PEAK_TFLOPS, FP32_VALU_PEAK_TFLOPS,
                                     "predictor_kernel<5,Elastic,ncp,float> (Picard + volume integral)"),
        "fp64_roofline": elastic_roofline(m64, B, FP64_PEAK_TFLOPS, None,
                                          "predictor_kernel<5,Elastic,ncp,double> (Picard + volume integral)"),
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_elastic(args.cpu_seconds),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=81, help="cells per direction")
    ap.add_argument("--order", type=int, default=5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--form", choices=("both", "fp32", "fp64", "linear"), default="both",
                    help="both: time FP64 and FP32-bulk and report the error between them (elastic: and the "
                         "linear kernels); one form: time that form only (linear: elastic workload only)")
    ap.add_argument("--predictor", choices=("auto", "mfma"), default="auto",
                    help="precision workload, FP32-bulk form: auto = the default predictor (contractions on the "
                         "vector ALUs); mfma = the predictor with its contractions on the matrix cores")
    ap.add_argument("--corrector", choices=("fused", "two-pass"), default="fused",
                    help="precision workload: fused = Riemann + surface + update + dt in one kernel per step "
                         "(default); two-pass = face pass writing the flux blocks, then the update pass")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fv-cells", type=int, default=54, help="fv: cells per direction")
    ap.add_argument("--patch-size", type=int, default=10, help="fv: subcells per cell and direction")
    ap.add_argument("--workload", choices=("precision", "limiter", "batch", "elastic", "fv", "amr", "output"),
                    default="precision",
                    help="precision: Euler N=5 81^3, FP32-bulk vs FP64 (the headline line); "
                         "limiter: Euler blast N=3 81^3 with the subcell limiter, FP64; "
                         "batch: 96 seeded uneven patches (orders 2..7), FP32-bulk vs FP64; "
                         "elastic: linear elasticity N=4 81^3, FP32-bulk vs FP64; "
                         "fv: Godunov finite volumes, 54^3 cells of 10^3 subcells, FP64; "
                         "amr: volume prolongation + restriction, 27^3 -> 81^3 cells, N=5, FP64; "
                         "output: VTK vertex interpolation of an 81^3 N=5 grid, FP64")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="batch: GB of patch state per wave")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 times): the known-size "
                         "reference for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from exahype_amd import capi
    capi.lib()
    if args.workload == "limiter":
        bench_limiter(capi, args)
        return
    if args.workload == "batch":
        bench_batch(capi, args)
        return
    if args.form == "linear" and args.workload != "elastic":
        sys.exit("--form linear applies to the elastic workload only")
    if args.workload == "amr":
        bench_amr(capi, args)
        return
    if args.workload == "output":
        bench_output(capi, args)
        return
    if args.workload == "elastic":
        bench_elastic(capi, args)
        return
    if args.workload == "fv":
        bench_fv(capi, args)
        return
    order, n = args.order, args.cells
    B, V, D = order + 1, 5, 5
    ncells = n ** 3
    dof = ncells * B ** 3
    forms = {"both": ("fp64", "fp32"), "fp32": ("fp32",), "fp64": ("fp64",)}[args.form]
    res = {}
    sol64 = None
    for form in forms:
        prec = capi.PRECISION_FP64 if form == "fp64" else capi.PRECISION_FP32_BULK
        variant = capi.PREDICTOR_MFMA if (form == "fp32" and args.predictor == "mfma") else 0
        r = run_form(capi, order, n, args.steps, args.warmup, args.seed, prec, variant,
                     fused=args.corrector == "fused")
        res[form] = r
        if form == "fp64" and "fp32" in forms:
            sol64 = r["grid"].solution()
        if form == "fp32" and sol64 is not None:
            sol32 = r["grid"].solution()
            r["l2"] = float(np.linalg.norm(sol32 - sol64) / np.linalg.norm(sol64))
            r["linf"] = float(np.abs(sol32 - sol64).max() / np.abs(sol64).max())
            del sol32
    main_form = "fp32" if "fp32" in res else "fp64"
    m = res[main_form]
    value = dof * args.steps / m["wall"]
    if main_form == "fp32":
        if args.predictor == "mfma":
            roof = roofline_for(m, B, V, D, FP32_PEAK_TFLOPS,
                                "predictor_mfma_kernel<6,Euler,flux> (predictor + volume integral)",
                                "matrix cores (v_mfma_f32_4x4x1f32) for the contractions, vector ALUs for the rest")
        else:
            roof = roofline_for(m, B, V, D, FP32_PEAK_TFLOPS,
                                "predictor_kernel<6,Euler,flux,float> (predictor + volume integral)",
                                "vector ALUs (FP32 FMA, no packed FP32; no MFMA instruction)",
                                valu_peak=FP32_VALU_PEAK_TFLOPS,
                                executed=lambda p, c: predictor_executed_flops_fp32(B, V, p, c))
    else:
        roof = roofline_for(m, B, V, D, FP64_PEAK_TFLOPS, "predictor_kernel<6,Euler,flux,double> (predictor + volume integral)",
                            "vector ALUs (FP64 FMA)")
    roof["traffic"] = measured_traffic(main_form, order, n)
    # algorithmic HBM bytes per launch: read luh (double), write lduh, lQhbnd, lFhbnd (bulk type)
    bulk = 4 if main_form == "fp32" else 8
    roof["algorithmic_bytes_per_launch"] = ncells * (8 * B ** 3 * D + bulk * (B ** 3 * V + 6 * B * B * (D + V)))
    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        cpu = cpu_baseline(m["grid"], order, n, args.cpu_seconds, threads)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if main_form == "fp32" else "f64",
        "data": "synthetic (entropy wave + 8 seeded Fourier modes, generated on the device)",
        "config": {"workload": f"euler3d-precision {n}^3 cells N={order} periodic CFL0.9",
                   "form": ("FP32-bulk: predictor/volume/Riemann f32, surface/update/dt f64"
                            if main_form == "fp32" else "plain FP64"),
                   "cells": ncells, "order": order, "dof_per_step": dof,
                   "scalar_unknowns_per_step": dof * V, "seed": args.seed,
                   "predictor": args.predictor,
                   "corrector": ("one kernel: Riemann on each cell's faces + surface + update + dt"
                                 if args.corrector == "fused" else "two passes: face Riemann, then update + dt")},
        "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in m["stages"].items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if args.corrector == "fused":
        out["corrector_roofline"] = corrector_roofline(m, main_form, order, n, B, V, D)
    if "fp64" in res:
        r64 = res["fp64"]
        out["fp64"] = {"value": dof * args.steps / r64["wall"], "ms_per_step": r64["wall"] / args.steps * 1e3,
                       "stage_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in r64["stages"].items()},
                       "roofline": dict(roofline_for(r64, B, V, D, FP64_PEAK_TFLOPS,
                                                     "predictor_kernel<6,Euler,flux,double>", "vector ALUs (FP64 FMA)"),
                                        traffic=measured_traffic("fp64", order, n))}
    if "fp32" in res and "fp64" in res:
        out["fp32_vs_fp64"] = {"speedup": out["fp64"]["ms_per_step"] / out["ms_per_step"],
                               "l2_rel_err": res["fp32"]["l2"], "linf_rel_err": res["fp32"]["linf"],
                               "steps_compared": args.warmup + args.steps}
    print(json.dumps(out), flush=True)
    for r in res.values():
        r["grid"].close()


if __name__ == "__main__":
    main()
