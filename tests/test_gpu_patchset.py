"""The patch set (exa_patchset): many independent periodic Euler patches of
uneven size and order sharing launches, with per-patch dt, memory waves,
hipGraph replay and the LDS-variant autotuner.

Bars: every patch against the oracle's grid of the same size and order after
several steps (<= 1e-12 relative per DOF, time stamp and next dt to 1e-13);
waves, graph replay and the LDS variants change nothing (bit-identical, or
1e-13 for a different LDS variant); the FP32-bulk form against FP64 (relative
L2 <= 1e-6); at the full workload size (the 96 seeded patches), finite
solutions and per-patch times equal to the sum of positive dts.
"""
import numpy as np
import pytest

from refutil import euler_state, node_coords, rel_per_dof

pytestmark = pytest.mark.gpu

# (n, order, k, amp): every order 2..7, small sides, two patches of one order
SMALL = [(4, 2, (1, 2, 3), 0.2), (3, 3, (2, 1, 1), 0.1), (5, 2, (3, 3, 1), 0.3), (3, 4, (1, 1, 2), 0.05),
         (2, 5, (1, 3, 2), 0.25), (2, 6, (2, 2, 2), 0.15), (2, 7, (1, 1, 1), 0.2)]


def _ic(n, order, k, amp):
    X = node_coords(order, n)
    rho = 1.0 + amp * np.sin(2 * np.pi * (X[..., 0] * k[0] + X[..., 1] * k[1] + X[..., 2] * k[2]))
    return euler_state(rho, (1.0, 1.0, 1.0), 1.0).reshape(-1)


def _run(exa, specs, steps, **kw):
    budget = kw.pop("budget", 100.0)
    prec = kw.pop("precision", exa.PRECISION_FP64)
    ps = exa.PatchSet(specs, precision=prec, hbm_budget_gb=budget, keep_max_cells=1 << 30)
    ps.run(steps, **kw)
    sols = [ps.solution(j) for j in range(len(specs))]
    times = [ps.time(j) for j in range(len(specs))]
    stats = ps.stats()
    ps.close()
    return sols, times, stats


def test_initial_state_is_the_entropy_wave(exa):
    print("case: patch set, 0 steps, initial condition", flush=True)
    sols, times, _ = _run(exa, SMALL, 0, use_graph=False, autotune=False)
    for spec, sol, (t, dt) in zip(SMALL, sols, times):
        assert rel_per_dof(sol, _ic(*spec)) < 1e-14
        assert t == 0.0 and dt > 0.0


def test_patches_vs_oracle(exa, oracle):
    steps = 3
    print(f"case: patch set of {len(SMALL)} patches vs oracle grids, {steps} steps", flush=True)
    sols, times, (sec, dof, nwaves) = _run(exa, SMALL, steps)
    assert nwaves == 1
    assert dof == sum(n ** 3 * (o + 1) ** 3 for n, o, _, _ in SMALL) * steps
    for spec, sol, (t, dt) in zip(SMALL, sols, times):
        n, order = spec[0], spec[1]
        o = oracle.Grid(oracle.make_solver(order, 5, 0), n)
        o.set_solution(_ic(*spec))
        odts, _ = o.step(steps, nthreads=0)
        err = rel_per_dof(sol, o.solution())
        assert err < 1e-12, (spec, err)
        assert abs(t - odts.sum()) <= 1e-13 * odts.sum(), spec
        assert abs(t - o.time()) <= 1e-13 * o.time()
        assert abs(dt - o.stable_dt()) <= 1e-13 * dt, spec


def test_waves_graph_and_variants_change_nothing(exa):
    steps = 2
    print("case: patch set, one wave vs several, graph vs eager, LDS variants", flush=True)
    ref, tref, (_, _, w1) = _run(exa, SMALL, steps, use_graph=False, autotune=False)
    assert w1 == 1
    # 1 MB holds every patch alone (the largest needs about 0.82 MB) but not all
    # of them (about 3.4 MB): several waves
    sols, times, (_, _, w2) = _run(exa, SMALL, steps, budget=1e-3, use_graph=True, autotune=False)
    assert w2 > 1
    for a, b in zip(ref, sols):
        assert np.array_equal(a, b)
    assert tref == times
    # the autotuner may choose either LDS variant: same arithmetic up to rounding order
    sols, _, _ = _run(exa, SMALL, steps, use_graph=True, autotune=True)
    for a, b in zip(ref, sols):
        assert rel_per_dof(a, b) < 1e-13


def test_fp32_bulk_vs_fp64(exa):
    steps = 4
    print("case: patch set FP32-bulk vs FP64", flush=True)
    ref, tref, _ = _run(exa, SMALL, steps)
    sols, times, _ = _run(exa, SMALL, steps, precision=exa.PRECISION_FP32_BULK)
    for spec, a, b in zip(SMALL, ref, sols):
        l2 = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert l2 <= 1e-6, (spec, l2)
    for (t0, _), (t1, _) in zip(tref, times):
        assert abs(t0 - t1) <= 1e-6 * t0


def test_full_workload_96_patches_fp32(exa):
    """The batching workload at its real size (96 seeded patches), FP32-bulk,
    2 steps: finite solutions, positive dts, time = 2 steps of each patch."""
    specs = exa.draw_patches(2024)
    print(f"case: 96-patch workload FP32-bulk 2 steps, {sum(n ** 3 for n, *_ in specs)} cells", flush=True)
    ps = exa.PatchSet(specs, precision=exa.PRECISION_FP32_BULK, hbm_budget_gb=200.0)
    ps.run(2)
    sec, dof, nwaves = ps.stats()
    print(f"  {nwaves} wave(s), {dof / sec:.3e} DOF/s", flush=True)
    for j in range(len(specs)):
        t, dt = ps.time(j)
        assert np.isfinite(t) and t > 0 and np.isfinite(dt) and dt > 0
    ps.close()
