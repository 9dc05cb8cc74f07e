"""Fused time stepping on the device grid (exa_grid_set_fused_time_stepping:
fuse-algorithmic-steps, Runner.cpp:1053-1110, TimeStepSizeComputation.cpp:
167-204) against the oracle's restatement (oracle/oracle.c grid_step_fused).

Bars: FP64 <= 1e-12 relative per DOF (each variable's scale where momenta
cross zero) after the steps, the corrector dt sequence to 1e-13 and the same
number of predictor re-runs; FP32-bulk against FP64 relative L2 < 1e-6.  The
re-run cases start a user grid at rest and let its adjustPointSolution pull
the flow towards the entropy wave, so the stable dt shrinks every step.
"""
import numpy as np
import pytest

from refutil import euler_state, rel_per_dof

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _scaled_err(a, b):
    a, b = a.reshape(-1, 5), b.reshape(-1, 5)
    return float((np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)).max())


@pytest.mark.parametrize("order,n,steps,weight", [(3, 5, 10, 0.9), (5, 4, 8, 0.99), (2, 6, 10, 1.0)])
def test_fused_time_stepping_vs_oracle(exa, oracle, order, n, steps, weight):
    print(f"case: fused time stepping Euler N={order} {n}^3 {steps} steps weight={weight}", flush=True)
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(nmodes=8, mode_amp=0.05, seed=order)
    luh0 = g.solution()
    g.set_fused_time_stepping(True, weight)
    g.step(steps)
    gpu, dts, reruns = g.solution(), g.dt_history(steps + 1), g.predictor_reruns()
    g.close()
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n)
    o.set_solution(luh0)
    o.set_fused(True, weight)
    odts, _ = o.step(steps)
    err = rel_per_dof(gpu, o.solution())
    print(f"  rel per DOF {err:.2e} reruns gpu {reruns} oracle {o.predictor_reruns()}", flush=True)
    assert err < TOL
    assert reruns == o.predictor_reruns()
    assert np.abs(dts[:steps] - odts).max() <= 1e-13 * odts.max()


def test_fused_time_stepping_two_pass_corrector(exa, oracle):
    order, n, steps = 3, 4, 6
    print(f"case: fused time stepping, two-pass corrector N={order} {n}^3", flush=True)
    g = exa.Grid(exa.euler_solver(order), n, boundary=exa.BC_TRANSMISSIVE)
    g.init_euler(nmodes=8, mode_amp=0.05, seed=7)
    luh0 = g.solution()
    g.set_fused_corrector(False)
    g.set_fused_time_stepping(True, 0.9)
    g.step(steps)
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n, transmissive=True)
    o.set_solution(luh0)
    o.set_fused(True, 0.9)
    odts, _ = o.step(steps)
    err = _scaled_err(g.solution(), o.solution())
    print(f"  scaled err {err:.2e}", flush=True)
    assert err < TOL
    assert np.abs(g.dt_history(steps)[:steps] - odts).max() <= 1e-13 * odts.max()
    g.close()


def _user_rest_grid(exa, order, n, boundary, precision=0):
    g = exa.Grid.from_user_solver(getattr(exa.user_example(), f"exa_user_euler_order{order}"), n,
                                  precision=precision, boundary=boundary)
    B = order + 1
    rest = np.tile(euler_state(1.0, (0.0, 0.0, 0.0), 1.0), g.ncells * B ** 3)
    g.upload(rest)
    g.set_adjust_solution(True)
    return g, rest


@pytest.mark.parametrize("order,n,steps,weight,wall", [(3, 4, 12, 1.0, False), (3, 4, 12, 0.95, True),
                                                        (5, 3, 8, 1.0, True)])
def test_fused_time_stepping_reruns_vs_oracle(exa, oracle, order, n, steps, weight, wall):
    print(f"case: fused time stepping re-runs N={order} {n}^3 weight={weight} walls={wall}", flush=True)
    g, rest = _user_rest_grid(exa, order, n, exa.BC_USER if wall else exa.BC_PERIODIC)
    g.set_fused_time_stepping(True, weight)
    g.step(steps)
    gpu, dts, reruns = g.solution(), g.dt_history(steps + 1), g.predictor_reruns()
    g.close()
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n, transmissive=wall)
    o.set_user(wall, True)
    o.set_solution(rest)
    o.set_fused(True, weight)
    odts, _ = o.step(steps)
    err = _scaled_err(gpu, o.solution())
    print(f"  scaled err {err:.2e} reruns gpu {reruns} oracle {o.predictor_reruns()}", flush=True)
    assert o.predictor_reruns() > 0
    assert reruns == o.predictor_reruns()
    assert err < TOL
    assert np.abs(dts[:steps] - odts).max() <= 1e-13 * odts.max()


@pytest.mark.parametrize("variant", ["valu", "mfma"])
def test_fused_time_stepping_fp32_bulk(exa, variant):
    order, n, steps = 5, 4, 8
    print(f"case: fused time stepping FP32-bulk ({variant}) vs FP64 N={order} {n}^3", flush=True)
    out = []
    for precision in (exa.PRECISION_FP64, exa.PRECISION_FP32_BULK):
        g = exa.Grid(exa.euler_solver(order, precision=precision), n)
        if precision == exa.PRECISION_FP32_BULK and variant == "mfma":
            g.set_predictor_variant(exa.PREDICTOR_MFMA)
        g.init_euler(nmodes=8, mode_amp=0.05, seed=3)
        g.set_fused_time_stepping(True, 0.9)
        g.step(steps)
        out.append((g.solution(), g.dt_history(steps), g.predictor_reruns()))
        g.close()
    l2 = np.linalg.norm(out[1][0] - out[0][0]) / np.linalg.norm(out[0][0])
    print(f"  L2 {l2:.2e} reruns {out[0][2]} {out[1][2]}", flush=True)
    assert l2 < 1e-6
    assert out[0][2] == out[1][2]
    assert np.abs(out[1][1] - out[0][1]).max() <= 1e-6 * out[0][1].max()


def test_fused_time_stepping_user_rerun_fp32(exa):
    """The re-run gate on the FP32 user kernels: FP32-bulk against FP64 on
    the accelerating user grid (a re-run every step at weight 1)."""
    order, n, steps = 3, 4, 10
    out = []
    for precision in (exa.PRECISION_FP64, exa.PRECISION_FP32_BULK):
        g, _ = _user_rest_grid(exa, order, n, exa.BC_PERIODIC, precision)
        g.set_fused_time_stepping(True, 1.0)
        g.step(steps)
        out.append((g.solution(), g.predictor_reruns()))
        g.close()
    l2 = np.linalg.norm(out[1][0] - out[0][0]) / np.linalg.norm(out[0][0])
    print(f"  L2 {l2:.2e} reruns {out[0][1]} {out[1][1]}", flush=True)
    assert out[0][1] > 0 and out[0][1] == out[1][1]
    assert l2 < 1e-6


def test_fused_time_stepping_refusals(exa):
    L = exa.lib()
    g = exa.Grid(exa.euler_solver(3), 2)
    assert L.exa_grid_set_fused_time_stepping(g._h, 1, 0.0) == exa.EXA_E_INVALID
    assert L.exa_grid_set_fused_time_stepping(g._h, 1, 1.5) == exa.EXA_E_INVALID
    assert L.exa_grid_set_fused_time_stepping(g._h, 1, 0.9) == exa.EXA_OK
    assert L.exa_grid_set_fused_time_stepping(g._h, 0, 0.0) == exa.EXA_OK
    g.close()
    lin = exa.Grid(exa.elastic_solver(3, linear=True), 2)
    assert L.exa_grid_set_fused_time_stepping(lin._h, 1, 0.9) == exa.EXA_E_UNSUPPORTED
    lin.close()
