"""Fused time stepping of the limiting grid on the device
(exa_lgrid_set_fused_time_stepping: LimitingADERDGSolver::fusedTimeStep
LS:806-836 in Runner::runOneTimeStepWithFusedAlgorithmicSteps, Runner.cpp:
1053-1110) against the oracle's restatement (oracle/limiter.c lg_step_fused).

Bars as for the non-fused limiting grid: the limiter status of every cell
identical after every step, the same steps with an irregular change, the same
number of predictor re-runs, dt and time to 1e-13, the solution to 1e-12 of
each variable's scale.  The blast cases have irregular steps (rollback, local
recomputation, re-prediction) and regular steps with re-runs.
"""
import numpy as np
import pytest

from refutil import blast_cells, rel_max

pytestmark = pytest.mark.gpu


def _pair(exa, oracle, order, n, periodic, radius, weight, helper_layers=1, cure=0, godunov=False):
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    luh = blast_cells(order, n, nodes=nodes, radius=radius)
    lim_ic = blast_cells(order, n, subcells=True, radius=radius)
    g = exa.LimitingGrid(exa.euler_solver(order), n,
                         boundary=exa.BC_PERIODIC if periodic else exa.BC_TRANSMISSIVE,
                         helper_layers=helper_layers, steps_till_cured=cure,
                         fv_scheme=exa.FV_GODUNOV if godunov else exa.FV_MUSCLHANCOCK)
    g.upload(luh, lim_ic)
    g.set_fused_time_stepping(True, weight)
    o = oracle.LimitingGrid(oracle.make_solver(order), n, periodic=periodic, helper_layers=helper_layers,
                            steps_till_cured=cure,
                            fv_scheme=oracle.FV_GODUNOV if godunov else oracle.FV_MUSCLHANCOCK)
    o.set_fused(True, weight)
    o.set_solution(luh, lim_ic)
    return g, o


@pytest.mark.parametrize("n,periodic,radius,weight,steps,helper_layers,cure,godunov",
                         [(6, True, 0.3, 1.0, 9, 1, 0, False), (6, True, 0.3, 0.9, 8, 1, 0, False),
                          (6, False, 0.2, 0.95, 7, 1, 0, False), (7, True, 0.3, 1.0, 8, 2, 1, False),
                          (6, True, 0.3, 0.9, 8, 1, 0, True)])
def test_fused_limiting_grid_vs_oracle(exa, oracle, n, periodic, radius, weight, steps, helper_layers, cure,
                                       godunov):
    order = 3
    print(f"case: fused limiting grid vs oracle, N={order} {n}^3 periodic={periodic} r={radius} "
          f"weight={weight} {steps} steps, helper layers {helper_layers}, cure {cure}, "
          f"FV {'Godunov' if godunov else 'MUSCL-Hancock'}", flush=True)
    g, o = _pair(exa, oracle, order, n, periodic, radius, weight, helper_layers, cure, godunov)
    assert np.array_equal(g.status()[0], o.status())
    irr_prev = 0
    for k in range(steps):
        t, dt, _ = g.time()
        assert abs(dt - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        assert abs(t - o.time()) <= 1e-13 * max(o.time(), o.dt()), f"time before step {k}"
        g.step(1)
        ost, _, oirr = o.step(1)
        st, irr = g.status()
        assert np.array_equal(st, ost[0]), f"limiter status after step {k}: {np.nonzero(st != ost[0])[0][:10]}"
        assert irr - irr_prev == int(oirr[0]), f"irregular change in step {k}"
        irr_prev = irr
        assert g.predictor_reruns() == o.predictor_reruns(), f"re-runs after step {k}"
        sol, osol = g.solution().reshape(-1, 5), o.solution().reshape(-1, 5)
        for v in range(5):
            err = np.abs(sol[:, v] - osol[:, v]).max() / max(np.abs(osol[:, v]).max(), 1e-300)
            assert err <= 1e-12 or np.abs(sol[:, v] - osol[:, v]).max() <= 1e-14, f"var {v} step {k}: {err:.3e}"
    print(f"  irregular steps {irr_prev}, re-runs {o.predictor_reruns()}", flush=True)
    assert irr_prev > 0
    if weight == 1.0:
        assert o.predictor_reruns() > 0
    # device patches keep two ghost layers for either scheme, the oracle's
    # Godunov patches one
    L = 2 * order + 1
    go = 1 if godunov else 2
    for c in np.nonzero(st >= helper_layers + 1)[0][:8]:
        a = g.limiter_solution(int(c)).reshape(L + 4, L + 4, L + 4, 5)[2:-2, 2:-2, 2:-2]
        b = o.limiter_solution(int(c)).reshape(L + 2 * go, L + 2 * go, L + 2 * go, 5)[go:-go, go:-go, go:-go]
        assert rel_max(a, b) <= 1e-11
    g.close()


def test_fused_limiting_without_troubled_cells_equals_fused_plain_grid(exa):
    """A smooth flow with a loose DMP never flags a cell: the fused limiting
    grid then equals the fused plain grid (same dt sequence and solution)."""
    order, n, steps = 3, 5, 6
    print(f"case: fused limiting grid without troubled cells vs fused plain grid N={order} {n}^3", flush=True)
    lg = exa.LimitingGrid(exa.euler_solver(order), n, boundary=exa.BC_PERIODIC, relax=1.0, scaling=1.0)
    pg = exa.Grid(exa.euler_solver(order), n)
    pg.init_euler(nmodes=6, mode_amp=0.03, seed=5)
    luh = pg.solution()
    lg.upload(luh)
    lg.set_fused_time_stepping(True, 0.9)
    pg.set_fused_time_stepping(True, 0.9)
    lg.step(steps)
    pg.step(steps)
    st, irr = lg.status()
    assert (st == 0).all() and irr == 0
    assert abs(lg.time()[0] - pg.time()[0]) <= 1e-13 * pg.time()[0]
    assert rel_max(lg.solution(), pg.solution()) <= 1e-12
    assert lg.predictor_reruns() == pg.predictor_reruns()
    lg.close()
    pg.close()


def test_fused_limiting_refusals(exa):
    L = exa.lib()
    g = exa.LimitingGrid(exa.euler_solver(3), 3)
    assert L.exa_lgrid_set_fused_time_stepping(g._h, 1, 0.0) == exa.EXA_E_INVALID
    assert L.exa_lgrid_set_fused_time_stepping(g._h, 1, 1.01) == exa.EXA_E_INVALID
    assert L.exa_lgrid_set_fused_time_stepping(None, 1, 0.9) == exa.EXA_E_INVALID
    g.close()


def test_fused_limiting_full_size_invariants(exa):
    """The limiter workload at its real shape (N=3, 81^3, transmissive blast)
    with fused time stepping: finite and positive, neighbouring statuses
    within one, the time stamp the sum of the corrector dts, and the same
    limited region as the non-fused run after the same steps to within a few
    cells (the dts differ slightly)."""
    print("case: fused limiting 81^3 N=3, 5 steps", flush=True)
    n, steps = 81, 5
    runs = []
    for fused in (False, True):
        g = exa.LimitingGrid(exa.euler_solver(3), n)
        g.init_blast(seed=7, radius=0.2)
        if fused:
            g.set_fused_time_stepping(True, 0.9)
        total = 0.0
        for _ in range(steps):
            total += g.time()[1]
            g.step(1)
        t, _, nsteps = g.time()
        st, irr = g.status()
        sol = g.solution().reshape(-1, 5)
        assert nsteps == steps and abs(t - total) <= 1e-13 * total
        assert np.isfinite(sol).all() and sol[:, 0].min() > 0
        S3 = st.reshape(n, n, n)
        for ax in range(3):
            assert np.abs(np.diff(S3, axis=ax)).max() <= 1
        runs.append(((st >= 2).mean(), irr, g.predictor_reruns()))
        g.close()
    print(f"  FV-active fraction non-fused {runs[0][0]:.4f} fused {runs[1][0]:.4f}, irregular steps "
          f"{runs[0][1]} / {runs[1][1]}, re-runs {runs[1][2]}", flush=True)
    assert 0.0 < runs[1][0] < 0.5
    assert abs(runs[1][0] - runs[0][0]) < 0.01


@pytest.mark.timeout(600)
def test_fused_limiting_grid_vs_oracle_mid_size(exa, oracle):
    """Statuses bit-exact against the oracle's fused driver on a 16^3
    transmissive blast over 20 fused steps (N=3), with the dt of every step,
    the irregular steps and re-runs, and the solution at the end."""
    order, n, steps, weight = 3, 16, 20, 0.9
    print(f"case: fused limiting grid vs oracle, N={order} {n}^3 blast {steps} steps", flush=True)
    g, o = _pair(exa, oracle, order, n, False, 0.1, weight)
    assert np.array_equal(g.status()[0], o.status())
    irr_prev = 0
    for k in range(steps):
        assert abs(g.time()[1] - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        g.step(1)
        ost, _, oirr = o.step(1)
        st, irr = g.status()
        assert np.array_equal(st, ost[0]), f"limiter status after step {k}: {np.nonzero(st != ost[0])[0][:10]}"
        assert irr - irr_prev == int(oirr[0]), f"irregular change in step {k}"
        irr_prev = irr
        assert g.predictor_reruns() == o.predictor_reruns(), f"re-runs after step {k}"
        if k % 5 == 4:
            print(f"  step {k + 1}: statuses identical", flush=True)
    sol, osol = g.solution().reshape(-1, 5), o.solution().reshape(-1, 5)
    for v in range(5):
        assert rel_max(sol[:, v], osol[:, v]) <= 1e-11, v
    print(f"  FV-active {(st >= 2).mean():.3f}, irregular steps {irr_prev}, re-runs {o.predictor_reruns()}", flush=True)
    assert irr_prev > 0 and (st >= 2).any()
    g.close()
