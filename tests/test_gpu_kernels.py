"""Parity of the HIP kernels (through the C-ABI) with the reference's golden
vectors and with the CPU oracle on the same inputs.

Bar: <= 1e-12 relative (FP64) against the oracle; the reference's eps = 1e-10
against its golden vectors; Picard iteration counts identical.
"""
import numpy as np
import pytest

from refutil import golden, random_euler_cells, rel_max, smooth_euler_cells

pytestmark = pytest.mark.gpu

G = golden()
EPS = 1e-10
TOL = 1e-12
MOCK_TERMS = 1 | 2 | 4  # flux | source | ncp, the predictor test's template flags


def mock(exa, terms=1):
    return exa.euler_solver(3, npar=1, terms=terms)


def test_golden_space_time_predictor(exa, oracle):
    s = mock(exa, MOCK_TERMS)
    luh = np.zeros(384)
    luh.reshape(64, 6)[:, :5] = np.array(G["testSpaceTimePredictorNonlinear"]["luh"]).reshape(64, 5)
    print("case: golden predictor, N=3 P=1", flush=True)
    r = exa.space_time_predictor_nonlinear(s, luh, 0.05, 1.083937460199773e-3, temps=True)
    g = G["testSpaceTimePredictorNonlinear"]
    assert np.abs(r["lQi"].reshape(256, 6)[:, :5].reshape(-1) - g["lQi"]).max() < EPS
    assert np.abs(r["lFi"].reshape(256, 4, 5)[:, :3, :].reshape(-1) - g["lFi"]).max() < EPS
    assert np.abs(r["lQhbnd"].reshape(96, 6)[:, :5].reshape(-1) - g["lQhbnd"]).max() < EPS
    assert np.abs(r["lFhbnd"] - g["lFhbnd"]).max() < EPS
    os_ = oracle.make_solver(3, 5, 1, use_source=True, use_ncp=True)
    o = oracle.space_time_predictor(os_, luh, 0.05, 1.083937460199773e-3)
    assert r["iters"] == o["iters"]
    for k in ("lQhbnd", "lFhbnd", "lQi", "lFi", "lQhi", "lFhi"):
        assert rel_max(r[k], o[k]) < TOL, k


def test_golden_surface_integral(exa):
    s = mock(exa)
    lF = np.zeros(480)
    for i in range(0, 80, 5):
        lF[0 + i + 1] = lF[80 + i + 1] = 1.0
        lF[160 + i + 2] = lF[240 + i + 2] = 1.0
        lF[320 + i + 3] = lF[400 + i + 3] = 1.0
    out = exa.surface_integral_nonlinear(s, np.arange(320) / 10.0, lF, 0.1)
    assert np.abs(out - G["testSurfaceIntegralNonlinear"]["lduh"]).max() < EPS


def test_golden_riemann(exa):
    s = mock(exa)
    QL = np.zeros(96)
    QR = np.zeros(96)
    for i in range(16):
        QL[i * 6:i * 6 + 5] = G["testRiemannSolver"]["QL"][i * 5:i * 5 + 5]
        QR[i * 6:i * 6 + 5] = G["testRiemannSolver"]["QR"][i * 5:i * 5 + 5]
    FL, FR = exa.riemann_solver_nonlinear(s, np.arange(80) + 1.0, np.arange(80) - 1.0, QL, QR,
                                          1.40831757919882352703e-03, 1)
    assert np.abs(FL - G["testRiemannSolverNonlinear"]["FL"]).max() < EPS
    assert np.abs(FR - G["testRiemannSolverNonlinear"]["FR"]).max() < EPS


def test_golden_solution_update(exa):
    s = mock(exa)
    luh = np.zeros(384)
    luh[0::6] = 1.0
    luh[4::6] = 2.5
    out = exa.solution_update(s, luh, np.arange(320) * 1.0, 1.40831757919882352703e-03)
    assert np.abs(out.reshape(64, 6)[:, :5].reshape(-1) - G["testSolutionUpdate"]["luh"]).max() < EPS
    assert np.all(out.reshape(64, 6)[:, 5] == 0.0)


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 7])
def test_batch_predictor_volume_vs_oracle(exa, oracle, order):
    """performPredictionAndVolumeIntegral over 48 cells with smooth and rough data."""
    rng = np.random.default_rng(100 + order)
    n = 24
    luh = np.concatenate([smooth_euler_cells(rng, n, order), random_euler_cells(rng, n, order)])
    dx = np.tile([0.05, 0.04, 0.06], (2 * n, 1))
    s = oracle.make_solver(order, 5, 0)
    B = order + 1
    nb3, nbf = B ** 3 * 5, 6 * B * B * 5
    # the time step the reference would take for these cells (CFL 0.9): beyond it the
    # Picard iteration of rough data diverges and only rounding noise is compared
    dt = min(oracle.stable_dt(s, luh[c * nb3:(c + 1) * nb3], dx[c]) for c in range(2 * n))
    print(f"case: batch predictor N={order} dt={dt:.3e}", flush=True)
    r = exa.batch_predictor_volume(exa.euler_solver(order), luh, dx, dt)
    for c in range(2 * n):
        o = oracle.space_time_predictor(s, luh[c * nb3:(c + 1) * nb3], dx[c], dt)
        lduh = oracle.volume_integral(s, o["lFhi"], dx[c])
        assert r["iters"][c] == o["iters"], c
        assert rel_max(r["lQhbnd"][c * nbf:(c + 1) * nbf], o["lQhbnd"]) < TOL, c
        assert rel_max(r["lFhbnd"][c * nbf:(c + 1) * nbf], o["lFhbnd"]) < TOL, c
        assert rel_max(r["lduh"][c * nb3:(c + 1) * nb3], lduh) < TOL, c


def test_golden_volume_integral(exa):
    """testVolumeIntegralNonlinear::lduh (GenericEulerKernelTest3D.cpp:204-215
    inputs, dx = 0.05: the mesh size the golden data was made with)."""
    print("case: golden volume integral", flush=True)
    lFhi = np.zeros(1280)
    for i in range(0, 320, 5):
        lFhi[i + 1] = lFhi[320 + i + 2] = lFhi[640 + i + 3] = 1.0
    out = exa.volume_integral_nonlinear(mock(exa), lFhi, np.full(3, 0.05))
    assert np.abs(out - np.array(G["testVolumeIntegralNonlinear"]["lduh"])).max() < EPS


@pytest.mark.parametrize("order", [2, 3, 5, 7])
def test_single_cell_kernels_vs_oracle(exa, oracle, order):
    rng = np.random.default_rng(7 + order)
    B = order + 1
    s = exa.euler_solver(order)
    so = oracle.make_solver(order, 5, 0)
    dx = np.array([0.1, 0.07, 0.05])
    luh = random_euler_cells(rng, 1, order)
    # volume integral from an arbitrary lFhi
    lFhi = rng.uniform(-1, 1, 4 * B ** 3 * 5)
    assert rel_max(exa.volume_integral_nonlinear(s, lFhi, dx), oracle.volume_integral(so, lFhi, dx)) < TOL
    # surface integral
    lduh = rng.uniform(-1, 1, B ** 3 * 5)
    lF = rng.uniform(-1, 1, 6 * B * B * 5)
    assert rel_max(exa.surface_integral_nonlinear(s, lduh, lF, dx), oracle.surface_integral(so, lduh, lF, dx)) < TOL
    # solution update
    assert rel_max(exa.solution_update(s, luh, lduh, 1e-3), oracle.solution_update(so, luh, lduh, 1e-3)) < TOL
    # stable time step
    a = exa.stable_time_step_size(s, luh, dx)
    b = oracle.stable_dt(so, luh, dx)
    assert abs(a - b) <= TOL * b
    # Riemann in every direction
    QL = random_euler_cells(rng, 1, order)[: B * B * 5]
    QR = random_euler_cells(rng, 1, order)[: B * B * 5]
    FL0 = rng.uniform(-1, 1, B * B * 5)
    FR0 = rng.uniform(-1, 1, B * B * 5)
    for d in range(3):
        FL, FR = exa.riemann_solver_nonlinear(s, FL0, FR0, QL, QR, 1e-3, d)
        oFL, oFR = oracle.riemann(so, FL0, FR0, QL, QR, 1e-3, d)
        assert rel_max(FL, oFL) < TOL and rel_max(FR, oFR) < TOL


def test_predictor_temporaries_vs_oracle(exa, oracle):
    """lQi / lFi / lQhi / lFhi exports (the reference's scratch arrays) at N=5."""
    rng = np.random.default_rng(3)
    luh = smooth_euler_cells(rng, 1, 5, scale=0.2)
    r = exa.space_time_predictor_nonlinear(exa.euler_solver(5), luh, 0.02, 1e-3, temps=True)
    o = oracle.space_time_predictor(oracle.make_solver(5, 5, 0), luh, 0.02, 1e-3)
    assert r["iters"] == o["iters"]
    for k in ("lQi", "lQhi", "lQhbnd", "lFhbnd"):
        assert rel_max(r[k], o[k]) < TOL, k
    B = 6
    assert rel_max(r["lFi"].reshape(B ** 4, 4, 5)[:, :3], o["lFi"].reshape(B ** 4, 4, 5)[:, :3]) < TOL
    assert rel_max(r["lFhi"][: 3 * B ** 3 * 5], o["lFhi"][: 3 * B ** 3 * 5]) < TOL


@pytest.mark.parametrize("order", [3, 5])
def test_no_time_averaging_vs_oracle(exa, oracle, order):
    """noTimeAveraging: the device averages in time first; the reference's
    branch integrates the space-time data -- the same linear maps, parity to
    the usual bar; the volume integral takes lFi in this form."""
    print(f"case: noTimeAveraging single-cell kernels N={order}", flush=True)
    rng = np.random.default_rng(90 + order)
    luh = smooth_euler_cells(rng, 1, order, scale=0.2)
    dx, dt = np.array([0.05, 0.04, 0.03]), 1e-3
    s = exa.euler_solver(order)
    s.no_time_averaging = 1
    so = oracle.make_solver(order, 5, 0, no_time_averaging=True)
    r = exa.space_time_predictor_nonlinear(s, luh, dx, dt, temps=True)
    o = oracle.space_time_predictor(so, luh, dx, dt)
    assert r["iters"] == o["iters"]
    for k in ("lQhbnd", "lFhbnd", "lQi"):
        assert rel_max(r[k], o[k]) < TOL, k
    lFi = o["lFi"]
    assert rel_max(exa.volume_integral_nonlinear(s, lFi, dx), oracle.volume_integral(so, o["lFhi"], dx, lFi=lFi)) < TOL


@pytest.mark.parametrize("order,seed,cfl", [(3, 1, 0.05), (5, 4, 0.07), (6, 6, 0.04)])
def test_picard_iterates_in_closed_form(exa, order, seed, cfl):
    """The device predictor on entropy-wave data, where every Picard iterate
    is known in closed form (tests/test_oracle_picard.py): the exit
    iteration, the last iterate in lQi/lQhi/lQhbnd and the one-iterate-stale
    fluxes in lFi/lFhi/lFhbnd."""
    print(f"case: Picard closed form N={order} seed={seed}", flush=True)
    from test_oracle_picard import check_against_closed_form
    s = exa.euler_solver(order)
    check_against_closed_form(lambda luh, h, dt: exa.space_time_predictor_nonlinear(s, luh, h, dt, temps=True),
                              order, seed, cfl)
