"""The one-kernel corrector (each cell solves the Rusanov problems of its six
faces from its own and its neighbours' face blocks, then surface integral,
solution update and stable dt; riemannSolverNonlinear.cpph:37-117 with
surfaceIntegralNonlinear, solutionUpdate and stableTimeStepSize) against the
two-pass corrector (face pass writing both cells' flux blocks, then the update
pass) on the same grids.

Bar: FP64 to 1e-13 relative per DOF after several steps (the face averages are
summed in the reference's node order in the one-kernel form and by a butterfly
in the face pass: the same sums in another order), the same dt sequence to
1e-13; FP32-bulk to 1e-6 relative L2.  Periodic and transmissive Euler grids,
the elastic (non-conservative product) grid, and conservation of the
one-kernel form on a periodic grid (both cells of a face get the same flux).
The grid-vs-oracle tests (test_gpu_grid.py) run the one-kernel form, the
default, at the reference's 1e-12 bar.
"""
import numpy as np
import pytest

from refutil import cell_weights, rel_per_dof

pytestmark = pytest.mark.gpu


def _pair(exa, make, steps, init):
    out = []
    for fused in (True, False):
        g = make()
        g.set_fused_corrector(fused)
        init(g)
        g.step(steps)
        out.append((g.solution(), g.dt_history(steps + 1)))
        g.close()
    return out


@pytest.mark.parametrize("order,n,boundary", [(2, 6, 0), (3, 5, 1), (5, 4, 0), (5, 4, 1), (7, 3, 0)])
def test_fused_corrector_fp64_euler(exa, order, n, boundary):
    print(f"case: fused corrector fp64 Euler N={order} {n}^3 boundary={boundary}", flush=True)
    (a, da), (b, db) = _pair(exa, lambda: exa.Grid(exa.euler_solver(order), n, boundary=boundary), 5,
                             lambda g: g.init_euler(nmodes=8, mode_amp=0.05, seed=order))
    err = rel_per_dof(a, b)
    print(f"  rel per DOF {err:.2e}", flush=True)
    assert err < 1e-13
    assert np.abs(da - db).max() <= 1e-13 * db.max()


@pytest.mark.parametrize("order,n", [(5, 6), (3, 8)])
def test_fused_corrector_fp32_bulk(exa, order, n):
    print(f"case: fused corrector fp32-bulk Euler N={order} {n}^3", flush=True)
    (a, _), (b, _) = _pair(exa, lambda: exa.Grid(exa.euler_solver(order, precision=exa.PRECISION_FP32_BULK), n),
                           4, lambda g: g.init_euler(nmodes=8, mode_amp=0.05, seed=1))
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  L2 {l2:.2e}", flush=True)
    assert l2 < 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_corrector_elastic(exa, precision):
    print(f"case: fused corrector elastic N=3 5^3 precision={precision}", flush=True)
    (a, _), (b, _) = _pair(exa, lambda: exa.Grid(exa.elastic_solver(3, precision=precision), 5), 4,
                           lambda g: g.init_elastic(seed=3, width=0.15))
    a, b = a.reshape(-1, 12)[:, :9], b.reshape(-1, 12)[:, :9]
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  L2 {l2:.2e}", flush=True)
    assert l2 < (1e-13 if precision == 0 else 1e-6)


def test_fused_corrector_conserves(exa):
    """Periodic FP64: the cell-weighted totals of all five variables stay
    constant to rounding over 10 steps (each face's flux enters both cells)."""
    order, n = 4, 6
    B = order + 1
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(nmodes=8, mode_amp=0.05, seed=5)
    u0 = g.solution()
    g.step(10)
    u = g.solution()
    g.close()
    w3 = cell_weights(order)[None, ..., None]
    tot0 = (w3 * u0.reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    tot = (w3 * u.reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    drift = np.abs(tot - tot0).max() / np.abs(tot0).max()
    print(f"  drift {drift:.2e}", flush=True)
    assert drift < 1e-13


def test_fused_corrector_refused_for_linear_grids(exa):
    g = exa.Grid(exa.elastic_solver(3, linear=True), 2)
    assert exa.lib().exa_grid_set_fused_corrector(g._h, 1) == exa.EXA_E_UNSUPPORTED
    assert exa.lib().exa_grid_set_fused_corrector(g._h, 0) == exa.EXA_OK
    g.close()
