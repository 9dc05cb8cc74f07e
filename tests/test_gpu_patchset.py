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


# elastic patches (velocity-stress NCP form) beside Euler ones: orders 2..5
ELASTIC = [(3, 2, 11, 0.15), (3, 3, 12, 0.2), (2, 4, 13, 0.25), (2, 5, 14, 0.2)]


def _mixed(exa):
    return SMALL[:3] + [exa.elastic_patch(n, o, seed, width=w) for n, o, seed, w in ELASTIC]


def _rel_var(a, b, nv):
    a, b = a.reshape(-1, nv), b.reshape(-1, nv)
    return max(np.abs(a[:, v] - b[:, v]).max() / max(np.abs(b[:, v]).max(), 1e-300) for v in range(nv))


def test_elastic_patches_vs_oracle(exa, oracle):
    """Elastic patches in the Euler patches' scheduler (groups of one PDE and
    order): initial state against the host generator, then every patch
    against the oracle's grid after 3 steps (each variable to 1e-12 of its
    scale, time and next dt to 1e-13)."""
    from refutil import elastic_cells
    steps = 3
    specs = _mixed(exa)
    print(f"case: patch set with {len(ELASTIC)} elastic patches vs oracle grids, {steps} steps", flush=True)
    ic, _, _ = _run(exa, specs, 0, use_graph=False, autotune=False)
    sols, times, (_, dof, _) = _run(exa, specs, steps)
    assert dof == sum(s[0] ** 3 * (s[1] + 1) ** 3 for s in SMALL[:3]) * steps + \
        sum(n ** 3 * (o + 1) ** 3 for n, o, _, _ in ELASTIC) * steps
    for j, spec in enumerate(specs):
        if isinstance(spec, dict):
            n, order = spec["n"], spec["order"]
            nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
            host = elastic_cells(order, n, seed=spec["seed"], nodes=nodes, width=spec["width"])
            assert _rel_var(ic[j], host, 12) < 1e-13, spec
            o = oracle.Grid(oracle.elastic_solver(order), n)
            nv = 12
        else:
            n, order = spec[0], spec[1]
            o = oracle.Grid(oracle.make_solver(order, 5, 0), n)
            nv = 5
        o.set_solution(ic[j])
        odts, _ = o.step(steps, nthreads=0)
        err = _rel_var(sols[j], o.solution(), nv)
        t, dt = times[j]
        print(f"  patch {j} N={order} {n}^3: {err:.2e}", flush=True)
        assert err < 1e-12, (spec, err)
        assert abs(t - o.time()) <= 1e-13 * o.time()
        assert abs(dt - o.stable_dt()) <= 1e-13 * dt, spec


@pytest.mark.parametrize("precision", [0, 1])
def test_patch_corrector_forms_agree(exa, precision):
    """The one-kernel corrector inside patches against the face and update
    passes: FP64 to 1e-13 of each variable's scale, FP32-bulk L2 1e-6."""
    steps = 3
    specs = _mixed(exa)
    print(f"case: patch set corrector forms, precision {precision}", flush=True)
    out = []
    for fused in (True, False):
        ps = exa.PatchSet(specs, precision=precision, hbm_budget_gb=100.0, keep_max_cells=1 << 30)
        ps.set_fused_corrector(fused)
        ps.run(steps, use_graph=True, autotune=False)
        out.append([ps.solution(j) for j in range(len(specs))])
        ps.close()
    for spec, a, b in zip(specs, *out):
        nv = 12 if isinstance(spec, dict) else 5
        if precision == 0:
            assert _rel_var(a, b, nv) < 1e-13, spec
        else:
            assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6, spec


def test_elastic_patches_fp32_bulk_vs_fp64(exa):
    steps = 4
    specs = [exa.elastic_patch(n, o, seed, width=w) for n, o, seed, w in ELASTIC]
    print("case: elastic patches FP32-bulk vs FP64", flush=True)
    ref, tref, _ = _run(exa, specs, steps)
    sols, times, _ = _run(exa, specs, steps, precision=exa.PRECISION_FP32_BULK)
    for spec, a, b in zip(specs, ref, sols):
        a, b = a.reshape(-1, 12)[:, :9], b.reshape(-1, 12)[:, :9]
        l2 = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert l2 <= 1e-6, (spec, l2)


def test_patch_profiling(exa):
    """Profiled run: per-group launches, Picard counts and times; the same
    solutions as the graph run (profiling only adds events)."""
    steps = 2
    specs = _mixed(exa)
    print("case: patch set profiling", flush=True)
    ref, _, _ = _run(exa, specs, steps, autotune=False)
    ps = exa.PatchSet(specs, hbm_budget_gb=100.0, keep_max_cells=1 << 30)
    ps.set_profiling(True)
    ps.run(steps, autotune=False)
    for j, spec in enumerate(specs):
        assert np.array_equal(ps.solution(j), ref[j])
    groups = {(exa.PDE_EULER, s[1]) for s in SMALL[:3]} | {(exa.PDE_ELASTIC, o) for _, o, _, _ in ELASTIC}
    for pde, order in groups:
        pred_ms, corr_ms, launches, picard, cells = ps.group_times(pde, order)
        ncells = sum(n ** 3 for n, o, *_ in (SMALL[:3] if pde == exa.PDE_EULER else ELASTIC) if o == order)
        assert launches == steps and cells == ncells * steps, (pde, order)
        assert picard >= cells and pred_ms > 0 and corr_ms > 0
    assert ps.group_times(exa.PDE_EULER, 7)[2] == 0
    ps.close()
