"""The linear-PDE part of the oracle (Cauchy-Kowalevski predictor, linear
Riemann solver, linear volume and surface integrals, coefficientMatrix),
pinned before it is trusted:

* the reference's 3-D Euler golden vectors (GenericEulerKernelTest3D.cpp:
  testVolumeIntegralLinear, testSurfaceIntegralLinear,
  testRiemannSolverLinear -- the test solver's coefficient matrix is zero),
  at the reference test's tolerance (GenericEulerKernelTest.h:69, 1e-10)
  and tighter where the printed digits allow;
* the reference's 2-D elasticity golden vectors (ElasticityKernelTest2D.cpp:
  the same three kernels, order 4, 9 variables + 3 parameters; the Riemann
  case exercises a non-zero coefficient matrix) through a z-invariant
  embedding: 3-D data constant along z reproduces the 2-D result, scaled by
  the z weight where the 3-D kernel integrates over z, up to rounding;
* the predictor on elastic plane waves along x, y and z with polynomial
  profiles of degree N in a homogeneous material, where Cauchy-Kowalevski is
  exact: the time-averaged state, the time-averaged B_d grad_d Q and the
  face values equal the exact solution computed independently by
  characteristic decomposition (numpy's eigenvectors of B_d) and polynomial
  integration in time;
* point sources: the time-derivative and time-average injection of
  STPL:79-105, 292-309 in closed form (read, not independently pinned).
"""
import numpy as np
import pytest

from refutil import golden

GE = golden("generic_euler_3d.json")
EL = golden("elasticity_2d.json")
EPS_EULER = 1e-10     # GenericEulerKernelTest.h:69
EPS_ELASTIC = 1e-9    # ElasticityKernelTest.h:69


def _w(oracle, order):
    return np.array([oracle.table(oracle.T_WEIGHTS, order, i) for i in range(order + 1)])


def _nodes(oracle, order):
    return np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])


# ---------------------------------------------------------------- Euler 3-D

def test_volume_integral_linear_euler_golden(oracle):
    s = oracle.make_solver(3, npar=1)
    lFhi = np.zeros((3, 64, 5))
    lFhi[0, :, 1] = 1.0
    lFhi[1, :, 2] = 2.0
    lFhi[2, :, 3] = 3.0
    lduh = oracle.volume_integral_linear(s, lFhi)
    ref = np.array(GE["testVolumeIntegralLinear"]["lduh"])
    assert np.abs(lduh - ref).max() <= 1e-14 * np.abs(ref).max()


def test_surface_integral_linear_euler_golden(oracle):
    s = oracle.make_solver(3, npar=1)
    F = np.zeros((6, 16, 5))
    F[0, :, 1] = F[1, :, 1] = 1.0
    F[2, :, 2] = F[3, :, 2] = 1.0
    F[4, :, 3] = F[5, :, 3] = 1.0
    lduh = oracle.surface_integral_linear(s, np.arange(320) / 10.0, F, 0.1)
    ref = np.array(GE["testSurfaceIntegralLinear"]["lduh"])
    assert np.abs(lduh - ref).max() <= 1e-13 * np.abs(ref).max()


def test_riemann_solver_linear_euler_golden(oracle):
    s = oracle.make_solver(3, npar=1)
    QL = np.zeros((16, 6))
    QR = np.zeros((16, 6))
    QL[:, :5] = np.array(GE["testRiemannSolver"]["QL"]).reshape(16, 5)
    QR[:, :5] = np.array(GE["testRiemannSolver"]["QR"]).reshape(16, 5)
    FL, FR = oracle.riemann_linear(s, QL, QR, 1.40831757919882352703e-03, 1)
    for got, key in ((FL, "FL"), (FR, "FR")):
        ref = np.array(GE["testRiemannSolverLinear"][key])
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), key


# ---------------------------------------------------------- elastic 2-D golden

def _elastic(order=4):
    from oracle import oracle as O
    return O.elastic_solver(order)


def test_riemann_solver_linear_elastic_golden_embedded(oracle):
    """2-D face of B nodes -> 3-D face of B x B nodes constant along the new
    axis; the face averages carry sum_i w_i = 1 (to the tables' rounding)."""
    s = _elastic(4)
    g = EL["testRiemannSolverLinear"]
    B = 5
    q2L = np.concatenate([np.array(g["qL_IN"]).reshape(B, 9), np.array(g["paramL_IN"]).reshape(B, 3)], 1)
    q2R = np.concatenate([np.array(g["qR_IN"]).reshape(B, 9), np.array(g["paramR_IN"]).reshape(B, 3)], 1)
    QL = np.broadcast_to(q2L, (B, B, 12)).copy()
    QR = np.broadcast_to(q2R, (B, B, 12)).copy()
    FL, FR = oracle.riemann_linear(s, QL, QR, 1.916666666666667e-4, 1)
    FL, FR = FL.reshape(B, B, 9), FR.reshape(B, B, 9)
    for got, key in ((FL, "FL_OUT"), (FR, "FR_OUT")):
        ref = np.array(g[key]).reshape(B, 9)
        scale = np.abs(ref).max()
        for i in range(B):
            # the reference test's own bar (absolute 1e-9): the vectors' inputs are
            # printed to 15 digits and their outputs agree with them to ~1e-10
            assert np.abs(got[i] - ref).max() <= EPS_ELASTIC, key
    # the coefficient matrix carries the data: zero it and the result moves
    assert np.abs(FR - FL).max() > 1e-3 * np.abs(FL).max()


def test_volume_integral_linear_elastic_golden_embedded(oracle):
    s = _elastic(4)
    g = EL["testVolumeIntegralLinear"]
    B = 5
    w = _w(oracle, 4)
    l2 = np.array(g["lFhi_IN"]).reshape(2, B, B, 9)
    lFhi = np.zeros((3, B, B, B, 9))
    lFhi[0] = l2[0][None]
    lFhi[1] = l2[1][None]
    lduh = oracle.volume_integral_linear(s, lFhi).reshape(B, B, B, 9)
    ref = np.array(g["lduh_OUT"]).reshape(B, B, 9)
    scale = np.abs(ref).max()
    for i in range(B):
        assert np.abs(lduh[i] / w[i] - ref).max() <= 1e-12 * scale


def test_surface_integral_linear_elastic_golden_embedded(oracle):
    s = _elastic(4)
    g = EL["testSurfaceIntegralLinear"]
    B = 5
    w = _w(oracle, 4)
    F2 = np.array(g["lFbnd_IN"]).reshape(4, B, 9)
    F = np.zeros((6, B, B, 9))
    F[0], F[1] = F2[0][None], F2[1][None]        # x faces (z, y): constant in z
    F[2], F[3] = F2[2][None], F2[3][None]        # y faces (z, x)
    lduh_in = w[:, None, None, None] * np.array(g["lduh_IN"]).reshape(B, B, 9)[None]
    dx = (38.4615384615385, 35.7142857142857, 1.0)
    out = oracle.surface_integral_linear(s, lduh_in, F, dx).reshape(B, B, B, 9)
    ref = np.array(g["lduh_OUT"]).reshape(B, B, 9)
    scale = np.abs(ref).max()
    for i in range(B):
        assert np.abs(out[i] / w[i] - ref).max() <= 1e-12 * scale

def _check_stp_linear_elastic_golden(g, lQhi, lFhi, lQhbnd, lFhbnd):
    """Every z slice of the 3-D outputs against the 2-D golden: lQhi (y,x,var),
    lFhi x block (y,x) and y block (x,y), face blocks 0-3 (x faces by y, y
    faces by x).  Bar: the reference test's absolute eps 1e-9, and 1e-9 of
    each array's scale (the outputs are ~1e-4, so the absolute bar alone
    would let a dropped Taylor term through)."""
    B = 5
    checks = [(lQhi.reshape(B, B, B, 12)[..., :9], np.array(g["lQhi_OUT"]).reshape(B, B, 9), "lQhi")]
    F = lFhi.reshape(4, B, B, B, 9)
    rF = np.array(g["lFhi_OUT"]).reshape(2, B, B, 9)
    checks += [(F[0], rF[0], "lFhi_x"), (F[1], rF[1], "lFhi_y")]
    Q = lQhbnd.reshape(6, B, B, 12)[..., :9]
    Fb = lFhbnd.reshape(6, B, B, 9)
    rQ, rFb = np.array(g["lQbnd_OUT"]).reshape(4, B, 9), np.array(g["lFbnd_OUT"]).reshape(4, B, 9)
    for f in range(4):
        checks += [(Q[f], rQ[f], f"lQbnd{f}"), (Fb[f], rFb[f], f"lFbnd{f}")]
    for got, ref, key in checks:
        scale = np.abs(ref).max()
        assert scale > 0, key
        for z in range(B):
            err = np.abs(got[z] - ref).max()
            assert err <= EPS_ELASTIC and err <= 1e-9 * scale, (key, z, err / scale)


def test_space_time_predictor_linear_elastic_golden_embedded(oracle):
    """testSpaceTimePredictorLinear (ElasticityKernelTest2D.cpp:315-420; order
    4, spaceTimePredictorLinear<false,false,false,true>, i.e. NCP only; the
    reference flags the test as needing a fix and does not run it).  The
    z-invariant embedding of luh (constant along z, so every z derivative
    vanishes up to the tables' rounding) reproduces the 2-D time-averaged
    state, time-averaged fluxes and face values at every z slice: this pins
    the Cauchy-Kowalevski recursion, the Taylor factors dt^i/(i+1)!, the
    directional B_d grad_d Q with non-constant material and the face
    extrapolation on data that is not a plane wave."""
    s = _elastic(4)
    g = EL["testSpaceTimePredictorLinear"]
    B = 5
    luh2 = np.concatenate([np.array(g["luh_IN"]).reshape(B, B, 9), np.array(g["param_IN"]).reshape(B, B, 3)], 2)
    luh = np.broadcast_to(luh2, (B, B, B, 12)).copy()
    dx = (38.4615384615385, 35.7142857142857, 1.0)
    dt = 0.813172798364530
    o = oracle.space_time_predictor_linear(s, luh, dx, dt)
    _check_stp_linear_elastic_golden(g, o["lQhi"], o["lFhi"], o["lQhbnd"], o["lFhbnd"])


def test_coefficient_matrix_matches_directional_ncp(oracle):
    """Bn[l][k] = B_d(k, l): applying Bn^T to a gradient gives the directional NCP."""
    s = _elastic(3)
    rng = np.random.default_rng(5)
    Q = np.concatenate([rng.normal(size=9), [2.0, 1.5, 2.5]])
    for d in range(3):
        g = np.zeros((3, 9))
        g[d] = rng.normal(size=9)
        Bn = oracle.coefficient_matrix(s, Q, d)
        ncp = oracle.ncp_directional(s, Q, g)
        assert np.allclose(ncp[d], Bn.T @ g[d], rtol=0, atol=1e-14)
        assert np.abs(np.delete(ncp, d, 0)).max() == 0.0
        # and the summed NCP of the nonlinear kernels is the sum of the blocks
        assert np.allclose(oracle.ncp(s, Q, g), ncp.sum(0), rtol=0, atol=1e-14)


# ------------------------------------------------- predictor: exact plane waves

def _plane_wave(oracle, s, order, d, dx, dt, rng):
    """Elastic plane wave along axis d with polynomial profiles of degree N;
    returns (luh, expected lQhi, expected lFhi_d, expected face values)."""
    B = order + 1
    lam, mu, rho = 2.0, 1.3, 2.4
    par = np.array([lam, mu, rho])
    Bn = oracle.coefficient_matrix(s, np.concatenate([np.zeros(9), par]), d)
    A = Bn.T                          # dq/dt + A dq/dx_d = 0
    ev, R = np.linalg.eig(A)
    ev, R = ev.real, R.real
    Rinv = np.linalg.inv(R)
    coeffs = rng.normal(size=(9, B)) / np.arange(1, B + 1)  # f_p(s) = sum c_pk s^k, s = local coordinate
    polys = [np.polynomial.Polynomial(c) for c in coeffs]
    xi = _nodes(oracle, order)
    svals = dx[d] * xi

    def avg(p, lamb, x):   # (1/dt) int_0^dt f(x - lamb t) dt
        if abs(lamb) < 1e-14:
            return p(x)
        P = p.integ()
        return (P(x) - P(x - lamb * dt)) / (lamb * dt)

    W0 = np.array([polys[p](svals) for p in range(9)])                     # w_p at the nodes, t = 0
    Wa = np.array([avg(polys[p], ev[p], svals) for p in range(9)])         # time averages
    Wd = np.array([avg(polys[p].deriv(), ev[p], svals) for p in range(9)])  # of d/dx
    q0 = R @ W0                     # (9, B) along axis d
    qa = R @ Wa
    fa = A @ (R @ Wd)               # time average of A dq/dx
    face = np.array([R @ np.array([avg(polys[p], ev[p], x) for p in range(9)]) for x in (0.0, dx[d])])

    def spread(v):  # (9 or 12, B along d) -> (z,y,x,var)
        out = np.zeros((B, B, B, v.shape[0]))
        for n in range(B):
            idx = [slice(None)] * 3
            idx[2 - d] = n
            out[tuple(idx)] = v[:, n]
        return out

    luh = np.concatenate([spread(q0), np.broadcast_to(par, (B, B, B, 3))], -1)
    qa12 = np.concatenate([spread(qa), np.broadcast_to(par, (B, B, B, 3))], -1)
    return luh, qa12, spread(fa), face, par


@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_predictor_linear_exact_on_plane_waves(oracle, order, d):
    s = _elastic(order)
    B = order + 1
    rng = np.random.default_rng(100 * order + d)
    dx = (0.7, 0.9, 1.1)
    dt = 0.05
    luh, qa, fa, face, par = _plane_wave(oracle, s, order, d, dx, dt, rng)
    out = oracle.space_time_predictor_linear(s, luh, dx, dt)
    lQhi = out["lQhi"].reshape(B, B, B, 12)
    lFhi = out["lFhi"].reshape(4, B, B, B, 9)
    scale = np.abs(qa[..., :9]).max()
    # the derivative operator dudx carries the reference's 13-digit literals:
    # terms built on derivatives agree to ~3e-11 of their scale
    assert np.abs(lQhi - qa).max() <= 1e-11 * scale
    assert np.abs(lFhi[d] - fa).max() <= 1e-10 * np.abs(fa).max()
    for e in range(3):
        if e != d:
            assert np.abs(lFhi[e]).max() <= 1e-10 * np.abs(fa).max()
    assert np.abs(lFhi[3]).max() == 0.0
    # faces 2d (s = 0) and 2d+1 (s = dx): the time average extrapolated
    lQb = out["lQhbnd"].reshape(6, B, B, 12)
    for side in (0, 1):
        got = lQb[2 * d + side][..., :9]
        assert np.abs(got - face[side][None, None, :]).max() <= 1e-10 * scale
        assert np.allclose(lQb[2 * d + side][..., 9:], par, rtol=1e-14, atol=0)
    # the (N+1)-th time derivative of degree-N data vanishes (up to the table
    # rounding amplified by N+1 applications of the derivative operator)
    lQi = out["lQi"].reshape(B + 1, B, B, B, 12)
    assert np.abs(lQi[B, ..., :9]).max() <= 1e-7 * np.abs(lQi[:B, ..., :9]).max()


def test_predictor_linear_point_sources(oracle):
    s = _elastic(3)
    B = 4
    dt = 0.02
    luh = np.zeros((B, B, B, 12))
    luh[..., 9:] = (2.0, 1.0, 2.0)
    g = np.linspace(0.1, 0.9, 9)
    PSi = np.zeros((B + 1, B, B, B, 9))
    PSi[0] = g                                   # first time derivative (STPL:96-105)
    out = oracle.space_time_predictor_linear(s, luh, (0.3, 0.3, 0.3), dt, PSi)
    lQhi = out["lQhi"].reshape(B, B, B, 12)
    assert np.allclose(lQhi[..., :9], 0.5 * dt * g, rtol=1e-11, atol=0)  # derivative tables: 13 digits
    assert np.abs(out["lFhi"]).max() <= 1e-12
    h = np.linspace(-1.0, 1.0, 9)
    PSi = np.zeros((B + 1, B, B, B, 9))
    PSi[1:] = h                                  # time nodes (STPL:82-94, 294-307)
    out = oracle.space_time_predictor_linear(s, luh, (0.3, 0.3, 0.3), dt, PSi)
    lQhi = out["lQhi"].reshape(B, B, B, 12)
    assert np.allclose(lQhi[..., :9], dt * h, rtol=1e-11, atol=1e-14)
    lFhi = out["lFhi"].reshape(4, B, B, B, 9)
    for e in range(3):
        assert np.allclose(lFhi[e], h / 3, rtol=1e-11, atol=1e-14)


# ------------------------------------------------------- linear grid driver

def test_linear_grid_matches_nonlinear_grid_homogeneous(oracle):
    """In a homogeneous material the two ADER-DG forms of the same elastic
    problem agree: the converged Picard predictor and the Cauchy-Kowalevski
    predictor expand the same linear equation, the Rusanov solver's NCP form
    0.5 (B dq -+ smax dq) equals the linear solver's fluctuations, and the
    nonlinear surface integral's F_{2d} sign matches the fluctuation of the
    opposite orientation.  So the oracle's linear grid driver (a different
    code path at every stage) must follow the nonlinear driver to the Picard
    tolerance; a wrong sign, index or Taylor factor anywhere in it does not."""
    from refutil import elastic_cells
    order, n, steps = 3, 3, 3
    nodes = _nodes(oracle, order)
    luh = elastic_cells(order, n, seed=5, nodes=nodes, width=0.2).reshape(-1, 12)
    luh[:, 9], luh[:, 10], luh[:, 11] = 4.0, 2.5, 1.7
    runs = {}
    for linear in (False, True):
        s = oracle.elastic_solver(order)
        s.linear = int(linear)
        g = oracle.Grid(s, n)
        g.set_solution(luh.reshape(-1))
        dts, _ = g.step(steps, nthreads=1)
        runs[linear] = (dts, g.solution().reshape(-1, 12))
    assert np.allclose(runs[False][0], runs[True][0], rtol=1e-14, atol=0)
    a, b = runs[True][1], runs[False][1]
    err = max(np.abs(a[:, v] - b[:, v]).max() / np.abs(b[:, v]).max() for v in range(9))
    # Picard stops once the squared change summed over the cell is below 1e-14
    # (entries moving by < 1e-7): the two forms agree to that level, measured
    # 4e-7 of the smallest variable's scale here
    assert err <= 2e-6, err
    assert np.array_equal(a[:, 9:], b[:, 9:])
    # the step moved the state by far more than that difference
    assert np.abs(a[:, :9] - luh[:, :9]).max() > 1e3 * err * np.abs(a[:, :9]).max()


# ----------------------------------------------------------- pointSource projection

@pytest.mark.parametrize("order", [1, 3, 5])
def test_point_source_projection_interpolates_the_delta(oracle, order):
    """pointSource.cpph:40-126 projects a point force at x0 onto the nodal
    basis: quadrature of the projected field against any polynomial p of
    degree <= N per direction returns F p(x0) (Lagrange interpolation on the
    Gauss nodes), for every time node; sources outside the cell give zero."""
    rng = np.random.default_rng(order)
    s = oracle.elastic_solver(order)
    B, V = order + 1, 9
    nodes, w = _nodes(oracle, order), _w(oracle, order)
    center, dx = np.array([0.3, -0.2, 1.1]), np.array([0.05, 0.07, 0.04])
    xref0 = rng.random(3)
    x0 = center - dx / 2 + xref0 * dx
    F = rng.standard_normal((B + 1, V))
    PSi = np.repeat(F[:, None, :], B ** 3, 1)                  # same force at every space node of a time node
    X0 = np.tile(x0, ((B + 1) * B ** 3, 1))
    out = oracle.point_source_projection(s, PSi, X0, center, dx).reshape(B + 1, B, B, B, V)
    coef = rng.standard_normal((B, B, B))
    def p(x, y, z):
        return sum(coef[a, b, c] * x ** a * y ** b * z ** c for a in range(B) for b in range(B) for c in range(B))
    quad = np.zeros((B + 1, V))
    for i in range(B):
        for j in range(B):
            for k in range(B):
                wt = (w[i] * dx[2]) * (w[j] * dx[1]) * (w[k] * dx[0])
                quad += wt * out[:, i, j, k, :] * p(nodes[k], nodes[j], nodes[i])
    exact = F * p(*xref0)
    assert np.abs(quad - exact).max() <= 1e-10 * np.abs(exact).max()
    # outside the cell: zero
    X1 = np.tile(center + 0.6 * dx, ((B + 1) * B ** 3, 1))
    assert not np.any(oracle.point_source_projection(s, PSi, X1, center, dx))
