"""A user PDE compiled into the device kernels (include/exahype_amd/user_solver.hpp,
the example in exahype_amd/examples/user_euler.hip) and driven through
exa_grid_create_user: the reference's user-solver route (the generated
AbstractMySolver calling the generic kernels with the user's flux,
eigenvalues, boundaryValues and adjustPointSolution,
UserSolverInterface.h:64-171).

Bars: the user Euler against the built-in Euler on the same grid, <= 1e-12
relative per DOF after 5 steps (the same formulas in another instantiation);
reflecting walls (boundaryValues, boundaryConditions.cpph:24-85) and the
per-step adjustPointSolution (solutionAdjustment.cpph:17-47 after the update
at t + dt) against the oracle's grid with the same hooks, <= 1e-12 of each
variable's scale, the dt sequence to 1e-13; FP32-bulk against FP64 on the
user grid, relative L2 < 1e-6.
"""
import ctypes

import numpy as np
import pytest

from refutil import rel_per_dof

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _table(exa, order):
    return getattr(exa.user_example(), f"exa_user_euler_order{order}")


def _user_grid(exa, order, n, boundary=0, precision=0):
    return exa.Grid.from_user_solver(_table(exa, order), n, precision=precision, boundary=boundary)


def _scaled_err(a, b):
    """max over variables of max|a-b| / max|b| (momenta cross zero at walls)."""
    a, b = a.reshape(-1, 5), b.reshape(-1, 5)
    return float((np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)).max())


def test_user_solver_table(exa):
    g = _user_grid(exa, 5, 2)
    assert (g.user.order, g.user.num_variables, g.user.num_parameters) == (5, 5, 0)
    assert g.user.has_boundary_values == 1 and g.user.has_adjust == 1
    assert g.user.cfl == 0.9
    g.close()


@pytest.mark.parametrize("order,n", [(3, 5), (5, 4)])
def test_user_euler_matches_builtin_periodic(exa, order, n):
    print(f"case: user Euler vs built-in, N={order} {n}^3 periodic", flush=True)
    u = _user_grid(exa, order, n)
    u.adjust_solution()  # adjustPointSolution at t = 0: the entropy wave
    b = exa.Grid(exa.euler_solver(order), n)
    b.init_euler()  # the same entropy wave from the built-in initialiser
    ic_err = rel_per_dof(u.solution(), b.solution())
    print(f"  IC rel per DOF {ic_err:.2e}", flush=True)
    assert ic_err < 1e-14
    b.init_euler(nmodes=8, mode_amp=0.05, seed=order)
    u.upload(b.solution())
    u.step(5)
    b.step(5)
    err = rel_per_dof(u.solution(), b.solution())
    du, db = u.dt_history(6), b.dt_history(6)
    print(f"  rel per DOF {err:.2e}", flush=True)
    assert err < TOL
    assert np.abs(du - db).max() <= 1e-13 * db.max()
    u.close()
    b.close()


def test_user_euler_matches_builtin_transmissive(exa):
    order, n = 5, 4  # the example compiles orders 3 and 5
    print(f"case: user Euler vs built-in, N={order} {n}^3 transmissive", flush=True)
    b = exa.Grid(exa.euler_solver(order), n, boundary=exa.BC_TRANSMISSIVE)
    b.init_euler(nmodes=8, mode_amp=0.05, seed=2)
    u = _user_grid(exa, order, n, boundary=exa.BC_TRANSMISSIVE)
    u.upload(b.solution())
    u.step(5)
    b.step(5)
    err = _scaled_err(u.solution(), b.solution())
    print(f"  scaled err {err:.2e}", flush=True)
    assert err < TOL
    u.close()
    b.close()


def _vs_oracle(exa, oracle, order, n, steps, wall, adjust):
    boundary = exa.BC_USER if wall else exa.BC_PERIODIC
    g = _user_grid(exa, order, n, boundary=boundary)
    g.adjust_solution()
    if adjust:
        g.set_adjust_solution(True)
    luh0 = g.solution()
    g.step(steps)
    gpu = g.solution()
    dts = g.dt_history(steps + 1)
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n, transmissive=wall)
    o.set_user(wall, adjust)
    o.set_solution(luh0)
    odts, _ = o.step(steps)
    g.close()
    return gpu, o.solution(), dts[:steps], odts


@pytest.mark.parametrize("order,n,steps", [(3, 5, 6), (5, 4, 4)])
def test_reflecting_walls_vs_oracle(exa, oracle, order, n, steps):
    print(f"case: user boundaryValues (reflecting walls) N={order} {n}^3 {steps} steps", flush=True)
    gpu, cpu, dts, odts = _vs_oracle(exa, oracle, order, n, steps, wall=True, adjust=False)
    err = _scaled_err(gpu, cpu)
    print(f"  scaled err {err:.2e}", flush=True)
    assert err < TOL
    assert np.abs(dts - odts).max() <= 1e-13 * odts.max()


@pytest.mark.parametrize("order,n,steps", [(3, 5, 6), (5, 3, 4)])
def test_adjust_every_step_vs_oracle(exa, oracle, order, n, steps):
    print(f"case: user adjustPointSolution every step N={order} {n}^3 {steps} steps", flush=True)
    gpu, cpu, dts, odts = _vs_oracle(exa, oracle, order, n, steps, wall=False, adjust=True)
    err = _scaled_err(gpu, cpu)
    print(f"  scaled err {err:.2e}", flush=True)
    assert err < TOL
    assert np.abs(dts - odts).max() <= 1e-13 * odts.max()


def test_walls_and_adjust_vs_oracle(exa, oracle):
    order, n, steps = 3, 4, 5
    print(f"case: user walls + adjust N={order} {n}^3 {steps} steps", flush=True)
    gpu, cpu, dts, odts = _vs_oracle(exa, oracle, order, n, steps, wall=True, adjust=True)
    err = _scaled_err(gpu, cpu)
    print(f"  scaled err {err:.2e}", flush=True)
    assert err < TOL
    assert np.abs(dts - odts).max() <= 1e-13 * odts.max()


def test_fp32_bulk_user_grid(exa):
    order, n, steps = 5, 4, 4
    print(f"case: user grid FP32-bulk vs FP64 N={order} {n}^3 walls", flush=True)
    out = []
    for precision in (exa.PRECISION_FP64, exa.PRECISION_FP32_BULK):
        g = _user_grid(exa, order, n, boundary=exa.BC_USER, precision=precision)
        g.adjust_solution()
        g.step(steps)
        out.append(g.solution())
        g.close()
    l2 = np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])
    print(f"  L2 {l2:.2e}", flush=True)
    assert l2 < 1e-6


def test_user_route_refusals(exa):
    L = exa.lib()
    b = exa.Grid(exa.euler_solver(3), 2)
    # the library's Euler has no boundaryValues / adjustPointSolution
    assert L.exa_grid_set_adjust_solution(b._h, 1) == exa.EXA_E_UNSUPPORTED
    assert L.exa_grid_adjust_solution(b._h) == exa.EXA_E_UNSUPPORTED
    b.close()
    h = ctypes.c_void_p()
    lo = (ctypes.c_double * 3)(0, 0, 0)
    hi = (ctypes.c_double * 3)(1, 1, 1)
    assert L.exa_grid_create(ctypes.byref(exa.euler_solver(3)), 2, 2, 2, lo, hi, exa.BC_USER,
                             ctypes.byref(h)) == exa.EXA_E_UNSUPPORTED
    u = _user_grid(exa, 3, 2, boundary=exa.BC_USER)
    # user grids run the one-kernel corrector only
    assert L.exa_grid_set_fused_corrector(u._h, 0) == exa.EXA_E_UNSUPPORTED
    assert L.exa_grid_set_adjust_solution(u._h, 1) == exa.EXA_OK
    u.close()
