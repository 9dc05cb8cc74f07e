"""The oracle's fused time stepping of the limiting solver
(LimitingADERDGSolver::fusedTimeStep LS:806-836, the SolutionUpdate
endIteration's check SolutionUpdate.cpp:305-318 with TimeStepSizeComputation.cpp:
167-204, updateMeshAndSubdomains Runner.cpp:866-946 with
recomputePredictorLocally LS:1438-1464, LocalRecomputation.cpp:168-222 and the
adapter's updateTimeStepSizesFused ADERDGSolver.cpp:594-612).

Pins:
  * no troubled cell: the fused limiting driver equals the fused plain driver
    (itself pinned in test_oracle_fused.py) bit for bit -- dt sequence,
    solution, time, re-runs;
  * a blast with irregular changes: the dt sequence follows, step by step,
    from the stable dt of each step's solution computed by the plain driver
    (regular step: the fused rule; irregular step: the new stable dt for
    corrector and predictor), and the time stamp is the sum of the dts;
  * after every step every cell below troubled holds the prediction of its
    current solution with the dt the step predicted with (the step's
    predictor dt, or the new dt after a re-run), bit for bit against the
    single-cell predictor: a cell missing from the re-prediction after an
    irregular change, or a wrong dt, keeps a stale prediction;
  * the blast invariants of the non-fused driver (DG of an FV-active cell =
    projection of its patch, a helper's patch = projection of its DG,
    positive density, neighbouring statuses within one).
Parity unpinned (no reference test runs it): which traversal re-predicts
which cell after an irregular change, beyond the consistency above.
"""
import numpy as np
import pytest

from refutil import GAMMA, blast_cells, node_coords


def _smooth_ic(oracle, order, n):
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    X = node_coords(order, n, nodes=nodes)
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * (X[..., 0] + X[..., 1] + X[..., 2]))
    Q = np.stack([rho, rho, rho, rho, 1.0 / (GAMMA - 1) + 1.5 * rho], -1)
    return np.ascontiguousarray(Q.reshape(-1))


@pytest.mark.parametrize("periodic,weight", [(True, 0.9), (False, 1.0), (True, 0.5)])
def test_fused_limiting_without_troubled_cells_is_fused_plain(oracle, periodic, weight):
    order, n, steps = 3, 4, 5
    s = oracle.make_solver(order)
    luh = _smooth_ic(oracle, order, n)
    lg = oracle.LimitingGrid(s, n, periodic=periodic, relax=1.0, scaling=1.0)
    lg.set_fused(True, weight)
    lg.set_solution(luh)
    st, dts, irr = lg.step(steps)
    assert (st == 0).all() and not irr.any()
    pg = oracle.Grid(s, n, transmissive=not periodic)
    pg.set_solution(luh)
    pg.set_fused(True, weight)
    pdts, _ = pg.step(steps)
    assert np.array_equal(dts, pdts)
    assert np.array_equal(lg.solution(), pg.solution())
    assert lg.time() == pg.time()
    assert lg.predictor_reruns() == pg.predictor_reruns()


def _stable_dt(oracle, s, n, luh, periodic):
    g = oracle.Grid(s, n, transmissive=not periodic)
    g.set_solution(luh)
    return g.stable_dt()


@pytest.mark.parametrize("weight", [0.9, 1.0])
def test_fused_limiting_blast(oracle, weight):
    order, n, steps, periodic = 3, 6, 8, True
    s = oracle.make_solver(order)
    B, L, g = order + 1, 2 * order + 1, 2
    S = L + 2 * g
    T = 3
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(B)])
    luh = blast_cells(order, n, nodes=nodes, radius=0.3)
    lim_ic = blast_cells(order, n, subcells=True, radius=0.3)
    grid = oracle.LimitingGrid(s, n, periodic=periodic)
    grid.set_fused(True, weight)
    grid.set_solution(luh, lim_ic)
    dx = 1.0 / n
    dc = dp = grid.dt()
    total = 0.0
    reruns = 0
    seen_irregular = seen_regular = seen_rerun = False
    for step in range(steps):
        st, dts, irr = grid.step(1)
        st, irregular = st[0], bool(irr[0])
        assert dts[0] == dc
        total += dc
        sol = grid.solution()
        s_new = _stable_dt(oracle, s, n, sol, periodic)
        rr = grid.predictor_reruns() - reruns
        reruns += rr
        pred_dt = dp  # the dt this step predicted with (before a re-run)
        if irregular:
            seen_irregular = True
            dc, dp = s_new, s_new
        else:
            seen_regular = True
            violated = dp > s_new
            assert rr == int(violated)
            if violated:
                dc, dp = weight * s_new, weight * s_new
            else:
                dc, dp = dp, min(s_new, 0.5 * (dp + weight * s_new))
        if rr:
            seen_rerun = True
            pred_dt = dc
        assert grid.dt() == dc
        assert grid.time() == pytest.approx(total, rel=1e-15, abs=0)
        # every cell below troubled holds the prediction of its current solution
        cells = sol.reshape(-1, B ** 3 * 5)
        for c in np.nonzero(st < T)[0]:
            qb, fb = grid.prediction(int(c))
            ref = oracle.space_time_predictor(s, cells[c], dx, pred_dt)
            assert np.array_equal(qb, ref["lQhbnd"]) and np.array_equal(fb, ref["lFhbnd"]), (step, int(c))
        assert cells.reshape(-1, 5)[:, 0].min() > 0
        for c in np.nonzero(st > 0)[0][:12]:
            lim = grid.limiter_solution(int(c))
            if st[c] >= 2:
                assert np.abs(oracle.project_on_dg(order, 5, g, lim) - cells[c]).max() < 1e-13
            else:
                proj = oracle.project_on_fv(order, 5, g, cells[c]).reshape(S, S, S, 5)
                inner = lim.reshape(S, S, S, 5)[g:g + L, g:g + L, g:g + L]
                assert np.abs(proj[g:g + L, g:g + L, g:g + L] - inner).max() < 1e-13
        S3 = st.reshape(n, n, n)
        for ax in range(3):
            assert np.abs(S3 - np.roll(S3, 1, axis=ax)).max() <= 1
    assert seen_irregular and seen_regular
    if weight == 1.0:
        assert seen_rerun
