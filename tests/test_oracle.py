"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

1. Golden vectors of the reference's own kernel tests
   (ExaHyPE/exahype/tests/kernels/c/GenericEulerKernelTest3D.cpp, data from
   tests/kernels/testdata/generic_euler_testdata.cpp), with the reference's
   tolerance eps = 1e-10 (GenericEulerKernelTest.h:69).
2. Closed forms the mathematics fixes where no golden vector exists:
   volume integral of a polynomial flux, stable time step of a constant state,
   Rusanov flux of equal states, exactness of the space-time predictor on
   an entropy wave whose density is a polynomial, conservation, and
   convergence of the driver to the exact entropy-wave solution.
"""
import numpy as np
import pytest

from refutil import (GAMMA, random_euler_cells, cell_weights, euler_flux, euler_state, gauss_legendre, golden, lagrange,
                     node_coords)

G = golden()
EPS = 1e-10  # GenericEulerKernelTest.h:69


def mock_solver(O, **kw):
    """GenericEulerKernelTest: V=5, P=1, Order=3 (GenericEulerKernelTest.h:29-32)."""
    return O.make_solver(3, 5, 1, **kw)


def test_pde_flux_kat(oracle):
    """testPDEFluxes (GenericEulerKernelTest3D.cpp:133-159)."""
    s = mock_solver(oracle)
    F = oracle.flux(s, [1.0, 0.1, 0.2, 0.3, 3.5, 0.0])
    g = G["testPDEFluxes"]
    for d, key in enumerate("fgh"):
        assert np.abs(F[d] - g[key]).max() < EPS


def test_surface_integral_golden(oracle):
    """testSurfaceIntegralNonlinear (GenericEulerKernelTest3D.cpp:274-316)."""
    s = mock_solver(oracle)
    lF = np.zeros(480)
    for i in range(0, 80, 5):
        lF[0 + i + 1] = lF[80 + i + 1] = 1.0
        lF[160 + i + 2] = lF[240 + i + 2] = 1.0
        lF[320 + i + 3] = lF[400 + i + 3] = 1.0
    out = oracle.surface_integral(s, np.arange(320) / 10.0, lF, 0.1)
    assert np.abs(out - G["testSurfaceIntegralNonlinear"]["lduh"]).max() < EPS


def volume_integral_golden_inputs():
    """GenericEulerKernelTest3D.cpp:204-215: lFhi = [x | y | z | src] blocks of
    64 nodes x 5 variables; variable d+1 of block d set to 1, source zero."""
    lFhi = np.zeros(1280)
    for i in range(0, 320, 5):
        lFhi[i + 1] = lFhi[320 + i + 2] = lFhi[640 + i + 3] = 1.0
    return lFhi


def test_volume_integral_golden(oracle):
    """testVolumeIntegralNonlinear::lduh (generic_euler_testdata.cpp).  The
    reference test body is commented out (GenericEulerKernelTest3D.cpp:217-225)
    and names no mesh size; the golden data matches the seeded lFhi with
    dx = 0.05 to 8e-13, and any other dx misses by O(1), which pins the 1/dx
    scaling as well as the Kxi orientation and the y/z block index orders."""
    s = mock_solver(oracle)
    lFhi = volume_integral_golden_inputs()
    ref = np.array(G["testVolumeIntegralNonlinear"]["lduh"])
    assert np.abs(oracle.volume_integral(s, lFhi, 0.05) - ref).max() < EPS
    assert np.abs(oracle.volume_integral(s, lFhi, 0.1) - ref).max() > 1.0


def riemann_golden_inputs():
    QL = np.zeros(96)
    QR = np.zeros(96)
    for i in range(16):
        QL[i * 6:i * 6 + 5] = G["testRiemannSolver"]["QL"][i * 5:i * 5 + 5]
        QR[i * 6:i * 6 + 5] = G["testRiemannSolver"]["QR"][i * 5:i * 5 + 5]
    return np.arange(80) + 1.0, np.arange(80) - 1.0, QL, QR


def test_riemann_golden(oracle):
    """testRiemannSolverNonlinear (GenericEulerKernelTest3D.cpp:408-482), normal 1."""
    s = mock_solver(oracle)
    FL, FR, QL, QR = riemann_golden_inputs()
    FL, FR = oracle.riemann(s, FL, FR, QL, QR, 1.40831757919882352703e-03, 1)
    assert np.abs(FL - G["testRiemannSolverNonlinear"]["FL"]).max() < EPS
    assert np.abs(FR - G["testRiemannSolverNonlinear"]["FR"]).max() < EPS


def test_solution_update_golden(oracle):
    """testSolutionUpdate (GenericEulerKernelTest3D.cpp:491-528); all 64 nodes checked."""
    s = mock_solver(oracle)
    luh = np.zeros(384)
    luh[0::6] = 1.0
    luh[4::6] = 2.5
    out = oracle.solution_update(s, luh, np.arange(320) * 1.0, 1.40831757919882352703e-03)
    got = out.reshape(64, 6)[:, :5].reshape(-1)
    assert np.abs(got - G["testSolutionUpdate"]["luh"]).max() < EPS
    assert np.all(out.reshape(64, 6)[:, 5] == 0.0)  # parameters untouched


def predictor_golden_run(O):
    s = mock_solver(O, use_source=True, use_ncp=True)
    luh = np.zeros(384)
    luh.reshape(64, 6)[:, :5] = np.array(G["testSpaceTimePredictorNonlinear"]["luh"]).reshape(64, 5)
    return s, luh, O.space_time_predictor(s, luh, 0.05, 1.083937460199773e-3)


def test_space_time_predictor_golden(oracle):
    """testSpaceTimePredictorNonlinear (GenericEulerKernelTest3D.cpp:611-766):
    <useSource,useFlux,useNCP>=true, P=1, dx=0.05, dt=1.0839e-3."""
    _, _, r = predictor_golden_run(oracle)
    g = G["testSpaceTimePredictorNonlinear"]
    assert r["iters"] >= 1
    assert np.abs(r["lQi"].reshape(256, 6)[:, :5].reshape(-1) - g["lQi"]).max() < EPS
    assert np.abs(r["lFi"].reshape(256, 4, 5)[:, :3, :].reshape(-1) - g["lFi"]).max() < EPS
    assert np.abs(r["lQhbnd"].reshape(96, 6)[:, :5].reshape(-1) - g["lQhbnd"]).max() < EPS
    assert np.abs(r["lFhbnd"] - g["lFhbnd"]).max() < EPS
    # lQhi / lFhi are in the data file too (their checks are commented out there)
    assert np.abs(r["lQhi"].reshape(64, 6)[:, :5].reshape(-1) - g["lQhi"]).max() < EPS
    assert np.abs(r["lFhi"][:960] - g["lFhi"]).max() < EPS


@pytest.mark.parametrize("order", [2, 3, 5, 7])
def test_volume_integral_closed_form(oracle, order):
    """sum_m Kxi[k][m] F(x_m) = int_0^1 phi_k'(x) F(x) dx for F of degree <= N (GL exact),
    scaled by w_a w_b / dx in each direction; lFhi_y is stored (z,x,y), lFhi_z (y,x,z)."""
    B = order + 1
    V = 5
    s = oracle.make_solver(order, V, 0)
    x, w = gauss_legendre(B)
    phi, dphi = lagrange(x)
    rng = np.random.default_rng(order)
    dx = np.array([0.3, 0.2, 0.1])
    # separable polynomial flux fields f_d(x,y,z) = a(x) b(y) c(z) per var, degree <= N
    coef = rng.uniform(-1, 1, (3, V, 3, B))
    P = lambda c, t: np.polyval(c[::-1], t)
    lFhi = np.zeros((4, B, B, B, V))
    expected = np.zeros((B, B, B, V))
    xq, wq = np.polynomial.legendre.leggauss(2 * B)
    xq = (xq + 1) / 2
    wq = wq / 2
    for v in range(V):
        for d in range(3):
            cx, cy, cz = coef[d, v]
            f = lambda X, Y, Z: P(cx, X) * P(cy, Y) * P(cz, Z)
            # stored blocks: x (z,y,x), y (z,x,y), z (y,x,z)
            for i in range(B):
                for j in range(B):
                    for k in range(B):
                        if d == 0:
                            lFhi[0, i, j, k, v] = f(x[k], x[j], x[i])
                        elif d == 1:
                            lFhi[1, i, j, k, v] = f(x[j], x[k], x[i])
                        else:
                            lFhi[2, i, j, k, v] = f(x[j], x[i], x[k])
            for iz in range(B):
                for iy in range(B):
                    for ix in range(B):
                        node = (x[ix], x[iy], x[iz])
                        if d == 0:
                            integ = sum(wq[q] * dphi(ix, xq[q]) * f(xq[q], node[1], node[2]) for q in range(2 * B))
                            expected[iz, iy, ix, v] += w[iz] * w[iy] / dx[0] * integ
                        elif d == 1:
                            integ = sum(wq[q] * dphi(iy, xq[q]) * f(node[0], xq[q], node[2]) for q in range(2 * B))
                            expected[iz, iy, ix, v] += w[iz] * w[ix] / dx[1] * integ
                        else:
                            integ = sum(wq[q] * dphi(iz, xq[q]) * f(node[0], node[1], xq[q]) for q in range(2 * B))
                            expected[iz, iy, ix, v] += w[iy] * w[ix] / dx[2] * integ
    got = oracle.volume_integral(s, lFhi.reshape(-1), dx)
    scale = np.abs(expected).max()
    assert np.abs(got - expected.reshape(-1)).max() < 1e-10 * scale


@pytest.mark.parametrize("order", [1, 3, 5, 9])
def test_stable_dt_constant_state(oracle, order):
    """dt = CFL PNPM[N] / sum_d (|u_d| + c)/dx_d for a constant state (stableTimeStepSize.cpph)."""
    PNPM = [1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015]
    s = oracle.make_solver(order, 5, 0)
    B = order + 1
    u = (0.3, -0.7, 0.1)
    Q = euler_state(np.full(B ** 3, 1.3), u, 0.9)
    dx = np.array([0.1, 0.2, 0.05])
    c = np.sqrt(GAMMA * 0.9 / 1.3)
    expected = 0.9 * PNPM[order] / sum((abs(u[d]) + c) / dx[d] for d in range(3))
    assert abs(oracle.stable_dt(s, Q.reshape(-1), dx) - expected) < 1e-13 * expected


def test_riemann_equal_states(oracle):
    """Rusanov with QL = QR and FL = FR = F(Q) returns F(Q) on both sides."""
    s = oracle.make_solver(4, 5, 0)
    B = 5
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.5, 2, B * B)
    Q = euler_state(rho, (0.2, 0.4, -0.3), 1.1)
    for d in range(3):
        F = euler_flux(Q, d).reshape(-1)
        FL, FR = oracle.riemann(s, F, F, Q.reshape(-1), Q.reshape(-1), 0.0, d)
        assert np.abs(FL - F).max() < 1e-14 and np.abs(FR - F).max() < 1e-14


def test_transmissive_boundary_riemann_is_exact_identity(oracle):
    """boundaryConditions with transmissive boundaryValues (outside = inside,
    boundaryConditions.cpph:24-85) followed by the Riemann solve of
    ADERDGSolver.cpp:2952-2960 returns the inner flux bit for bit: Rusanov of
    equal states is 0.5 (F + F) - 0.5 s (Q - Q) = F exactly."""
    s = oracle.make_solver(3, 5, 0)
    B = 4
    rng = np.random.default_rng(2)
    for d in range(3):
        Q = random_euler_cells(rng, 1, 3)[: B * B * 5]
        F = rng.normal(size=B * B * 5)
        FL, FR = oracle.riemann(s, F, F, Q, Q, 0.0, d)
        assert np.array_equal(FL, F) and np.array_equal(FR, F)


def test_transmissive_grid_keeps_a_uniform_state(oracle):
    order, n = 3, 3
    s = oracle.make_solver(order, 5, 0)
    g = oracle.Grid(s, n, transmissive=True)
    luh = np.tile(euler_state(np.array([1.3]), (0.3, -0.2, 0.1), 2.0).ravel(), n ** 3 * (order + 1) ** 3)
    g.set_solution(luh)
    g.step(3)
    # table rounding (sum_n Kxi[n][l] ~ 1e-13, 13-digit literals) moves it by
    # ~1e-12; a wrong boundary flux would move it by O(dt/dx)
    assert np.abs(g.solution() - luh).max() < 1e-10 * np.abs(luh).max()


def entropy_wave_cell(order, u, centre, h, dt):
    """A cell whose density is a cubic polynomial, advected with constant u and p:
    the exact Euler solution rho(x - u t) lies in the space-time polynomial space."""
    B = order + 1
    x, w = gauss_legendre(B)

    def rho(X, Y, Z):
        a, b, c = (X - centre[0]) / h, (Y - centre[1]) / h, (Z - centre[2]) / h
        return 1.0 + 0.2 * a + 0.1 * b * b - 0.05 * a * c + 0.03 * a * b * c

    loc = centre[0] + h * (x - 0.5), centre[1] + h * (x - 0.5), centre[2] + h * (x - 0.5)
    Z, Y, X = np.meshgrid(loc[2], loc[1], loc[0], indexing="ij")
    luh = euler_state(rho(X, Y, Z), u, 1.0)
    return rho, luh, x, w, loc


@pytest.mark.parametrize("order", [3, 4, 5])
def test_predictor_exact_on_polynomial_entropy_wave(oracle, order):
    """The converged space-time predictor equals the exact solution (to the Picard
    tolerance): lQi(t) = Q(x - u tau_t dt), lQhbnd = time-averaged face traces,
    lFhbnd = time-averaged normal fluxes.  Wrong signs, directions, time-node
    mappings or dx scalings fail by O(1e-2)."""
    s = oracle.make_solver(order, 5, 0)
    B = order + 1
    u = (0.7, -0.4, 0.3)
    h = 0.1
    centre = (0.35, 0.55, 0.45)
    dt = 0.2 * h / 2.0
    rho, luh, x, w, loc = entropy_wave_cell(order, u, centre, h, dt)
    r = oracle.space_time_predictor(s, luh.reshape(-1), h, dt)
    assert 1 <= r["iters"] <= 2 * B
    # lQi (z,y,x,t,var)
    lQi = r["lQi"].reshape(B, B, B, B, 5)
    Z, Y, X = np.meshgrid(loc[2], loc[1], loc[0], indexing="ij")
    for t in range(B):
        tt = x[t] * dt
        exact = euler_state(rho(X - u[0] * tt, Y - u[1] * tt, Z - u[2] * tt), u, 1.0)
        assert np.abs(lQi[:, :, :, t, :] - exact).max() < 1e-7
    # faces: x (z,y), y (z,x), z (y,x); side 0 at the lower face
    lQhbnd = r["lQhbnd"].reshape(6, B, B, 5)
    lFhbnd = r["lFhbnd"].reshape(6, B, B, 5)
    for f in range(6):
        d, side = f // 2, f % 2
        face = centre[d] + h * (side - 0.5)
        qa = np.zeros((B, B, 5))
        fa = np.zeros((B, B, 5))
        for t in range(B):
            tt = x[t] * dt
            A, Bm = np.meshgrid(loc[2] if d != 2 else loc[1], loc[1] if d == 0 else loc[0], indexing="ij")
            if d == 0:
                Xs, Ys, Zs = np.full_like(A, face), Bm, A
            elif d == 1:
                Xs, Ys, Zs = Bm, np.full_like(A, face), A
            else:
                Xs, Ys, Zs = Bm, A, np.full_like(A, face)
            q = euler_state(rho(Xs - u[0] * tt, Ys - u[1] * tt, Zs - u[2] * tt), u, 1.0)
            qa += w[t] * q
            fa += w[t] * euler_flux(q, d)
        assert np.abs(lQhbnd[f] - qa).max() < 1e-7, f
        assert np.abs(lFhbnd[f] - fa).max() < 1e-7, f


def grid_entropy_wave(O, order, n, steps, amp=0.2, u=(1.0, 1.0, 1.0)):
    s = O.make_solver(order, 5, 0)
    g = O.Grid(s, n)
    X = node_coords(order, n)
    rho0 = 1.0 + amp * np.sin(2 * np.pi * (X[..., 0] + X[..., 1] + X[..., 2]))
    g.set_solution(euler_state(rho0, u, 1.0).reshape(-1))
    dts, _ = g.step(steps, nthreads=4)
    t = g.time()
    sol = g.solution().reshape(-1, order + 1, order + 1, order + 1, 5)
    exact = euler_state(1.0 + amp * np.sin(2 * np.pi * ((X[..., 0] - u[0] * t) + (X[..., 1] - u[1] * t)
                                                         + (X[..., 2] - u[2] * t))), u, 1.0)
    return g, sol, exact, dts, t


def test_driver_conserves_and_converges(oracle):
    """Periodic entropy wave: mass, momentum and energy are conserved to rounding, and
    the error to the exact solution rho(x - u t) drops at least 2^3.5 when h halves (N=3)."""
    errs = []
    for n in (4, 8):
        g, sol, exact, dts, t = grid_entropy_wave(oracle, 3, n, steps=int(12 * n / 4))
        errs.append(np.abs(sol - exact).max())
        # conservation: sum over cells of w^3 luh (all cells have the same volume)
        w3 = cell_weights(3)[None, ..., None]
        X = node_coords(3, n)
        rho0 = 1.0 + 0.2 * np.sin(2 * np.pi * X.sum(-1))
        init = euler_state(rho0, (1.0, 1.0, 1.0), 1.0)
        tot0 = (w3 * init).sum(axis=(0, 1, 2, 3))
        tot1 = (w3 * sol).sum(axis=(0, 1, 2, 3))
        assert np.abs(tot1 - tot0).max() < 1e-12 * np.abs(tot0).max()
        assert np.all(dts > 0)
    assert errs[0] < 2e-2
    assert errs[0] / errs[1] > 2 ** 3.5, errs


@pytest.mark.parametrize("order", [2, 3, 5])
def test_no_time_averaging_branch_matches_the_averaged_one(oracle, order):
    """noTimeAveraging (spaceTimePredictorNonlinear.cpph:690-713 and the lFi
    volume integral, volumeIntegralNonlinear.cpph:44-135) applies the same
    linear maps as the time-averaged branch in another order: the face data
    and the volume integral agree to rounding, the Picard loop is shared."""
    rng = np.random.default_rng(40 + order)
    luh = random_euler_cells(rng, 1, order)
    dx, dt = np.array([0.1, 0.08, 0.06]), 2e-3
    a = oracle.space_time_predictor(oracle.make_solver(order, 5, 0), luh, dx, dt)
    s_nt = oracle.make_solver(order, 5, 0, no_time_averaging=True)
    b = oracle.space_time_predictor(s_nt, luh, dx, dt)
    assert a["iters"] == b["iters"]
    assert np.array_equal(a["lQi"], b["lQi"]) and np.array_equal(a["lFi"], b["lFi"])
    for k in ("lQhbnd", "lFhbnd"):
        assert np.abs(a[k] - b[k]).max() <= 1e-13 * np.abs(a[k]).max(), k
    va = oracle.volume_integral(oracle.make_solver(order, 5, 0), a["lFhi"], dx)
    vb = oracle.volume_integral(s_nt, np.zeros_like(a["lFhi"]), dx, lFi=a["lFi"])
    assert np.abs(va - vb).max() <= 1e-13 * np.abs(va).max()
