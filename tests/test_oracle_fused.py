"""The oracle's fused time stepping (fuse-algorithmic-steps; Runner.cpp:
1053-1110, ADERDGSolver::fusedTimeStep ADERDGSolver.cpp:1853-1879,
TimeStepSizeComputation.cpp:167-204 and 226-237, ADERDGSolver.cpp:560-616).

Pins:
  * the fused scheme equals the standard scheme run with the fused scheme's
    corrector dt sequence, bit for bit (every prediction used the dt of the
    correction that consumes it; a missing or misplaced re-run breaks this);
  * a uniform flow: the dt sequence in closed form, dt_k - w s halving each
    step towards the fixed point w s (or w s from the first step when the
    first stability check fails by rounding);
  * a uniform state accelerated by an adjustPointSolution hook: the stable dt
    shrinks in closed form (s = C / sum_d (|u_d| + c)), the re-runs and the
    dt sequence follow from it, the time stamp is the sum of the corrector dts.
"""
import numpy as np
import pytest

from refutil import euler_state, node_coords

GAMMA = 1.4


def _grid(oracle, order, n, luh0, weight, fused=True):
    g = oracle.Grid(oracle.make_solver(order, 5, 0), n)
    g.set_solution(luh0)
    if fused:
        g.set_fused(True, weight)
    return g


def _wave(order, n, seed=0):
    X = node_coords(order, n)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * X.sum(-1))
    for _ in range(4):
        k = rng.integers(1, 3, size=3)
        rho = rho + 0.05 * np.sin(2 * np.pi * (X @ k) + rng.uniform(0, 2 * np.pi))
    return euler_state(rho, (1.0, 1.0, 1.0), 1.0).ravel()


@pytest.mark.parametrize("weight", [0.9, 0.99, 1.0])
def test_fused_equals_standard_with_its_dt_sequence(oracle, weight):
    order, n, steps = 3, 3, 8
    luh0 = _wave(order, n)
    f = _grid(oracle, order, n, luh0, weight)
    dts, _ = f.step(steps)
    s = _grid(oracle, order, n, luh0, weight, fused=False)
    s.step_dts(dts)
    assert np.array_equal(f.solution(), s.solution())
    assert f.time() == s.time()
    # the fused dts are smaller than the standard scheme's (lagged and weighted)
    std = _grid(oracle, order, n, luh0, weight, fused=False)
    sdts, _ = std.step(steps)
    assert dts[0] == sdts[0]
    if weight < 1.0:
        assert np.all(dts[2:] < sdts[2:])


def _speed_sum(Q):
    rho = Q[0]
    u = Q[1:4] / rho
    p = (GAMMA - 1.0) * (Q[4] - 0.5 * rho * (u ** 2).sum())
    c = np.sqrt(GAMMA * p / rho)
    return (np.abs(u) + c).sum()


@pytest.mark.parametrize("weight", [0.9, 0.5])
def test_fused_uniform_flow_dt_closed_form(oracle, weight):
    order, n, steps = 2, 2, 10
    B = order + 1
    Q = euler_state(1.0, (0.5, -0.25, 0.125), 1.0)
    luh0 = np.tile(Q, n ** 3 * B ** 3)
    g = _grid(oracle, order, n, luh0, weight)
    s = g.stable_dt()
    dts, _ = g.step(steps)
    reruns = g.predictor_reruns()
    assert reruns in (0, 1)
    if reruns == 0:  # the first check passed: dt_k = w s + (1 - w) s / 2^(k-1), k >= 1
        exp = np.array([s] + [weight * s + (1 - weight) * s / 2 ** (k - 1) for k in range(1, steps)])
    else:            # the first check failed by rounding: w s from the first step on
        exp = np.array([s] + [weight * s] * (steps - 1))
    assert np.abs(dts - exp).max() <= 1e-13 * s  # s itself moves by rounding each step
    assert abs(g.time() - dts.sum()) <= 1e-15 * steps
    assert np.abs(g.solution() - luh0).max() < 1e-12 * np.abs(luh0).max()  # rounding of the DG update


@pytest.mark.parametrize("weight", [1.0, 0.99, 0.9])
def test_fused_reruns_when_an_adjustment_accelerates_the_flow(oracle, weight):
    """Uniform rest state; after every update the hook relaxes Q towards a
    fast uniform flow, Q += (dt/2)(Q_t - Q), so the stable dt shrinks."""
    order, n, steps = 2, 2, 12
    B = order + 1
    Q0 = euler_state(1.0, (0.0, 0.0, 0.0), 1.0)
    Qt = euler_state(1.0, (2.0, 1.0, 0.5), 1.0)

    def adj(_s, x, t, dt, q):
        for m in range(5):
            q[m] += 0.5 * dt * (Qt[m] - q[m])

    cb = oracle.ADJ_FN(adj)
    luh0 = np.tile(Q0, n ** 3 * B ** 3)
    g = _grid(oracle, order, n, luh0, weight)
    g.set_hooks(None, cb)
    g.set_fused(True, weight)
    C = g.stable_dt() * _speed_sum(Q0)  # s(Q) = C / sum_d (|u_d| + c) on this grid
    dts, _ = g.step(steps)
    # the rule with the closed-form stable dt of the uniform state
    Q, dtc, dtp, reruns, exp = Q0.copy(), C / _speed_sum(Q0), C / _speed_sum(Q0), 0, []
    for _ in range(steps):
        exp.append(dtc)
        Q = Q + 0.5 * dtc * (Qt - Q)
        stable = C / _speed_sum(Q)
        if dtp > stable:
            dtp = weight * stable
            nxt = weight * stable
            reruns += 1
        else:
            nxt = 0.5 * (dtp + weight * stable)
        dtc, dtp = dtp, nxt
    exp = np.array(exp)
    assert np.abs(dts - exp).max() <= 1e-13 * exp.max()
    assert g.predictor_reruns() == reruns
    if weight == 1.0:
        assert reruns >= steps - 1  # the shrinking dt fails every check
    got = g.solution().reshape(-1, 5)
    assert np.abs(got - Q).max() < 1e-12 * np.abs(Q).max()
    # and the fused run equals the standard run with its dts
    s = _grid(oracle, order, n, luh0, weight, fused=False)
    s.set_hooks(None, cb)
    s.step_dts(dts)
    assert np.array_equal(g.solution(), s.solution())
