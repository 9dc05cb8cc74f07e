"""The oracle's restatement of the user-hook kernels: boundaryConditions
(K/c/3d/boundaryConditions.cpph:24-85) and solutionAdjustment
(K/c/3d/solutionAdjustment.cpph:17-47), the two test hooks (the Euler
demonstrator's reflecting wall, MyEulerSolver_ADERDG.cpp:63-88, and an
entropy-wave adjustPointSolution) and where the grid driver calls them
(applyBoundaryConditions, ADERDGSolver.cpp:2893-2966; adjustSolution after the
update at t + dt, ADERDGSolver.cpp:2231-2236).

Pins: the coordinates and node pointers handed to the callbacks against the
Gauss-Legendre nodes from numpy and the face/volume block layouts; the wall
state and flux against the textbook Euler flux; the first step of a uniform
flow in a reflecting box in closed form (only the wall faces act, through the
Rusanov flux of the inside and mirrored states); the adjustment's
place in the step (one step with it equals one step without it followed by
the adjustment at t + dt).
"""
import numpy as np

from refutil import euler_flux, euler_state, gauss_legendre, node_coords


def _euler(oracle, order):
    return oracle.make_solver(order, 5, 0)


def test_boundary_conditions_hands_each_face_node_to_the_callback(oracle):
    order = 4
    B = order + 1
    s = _euler(oracle, order)
    nodes, _ = gauss_legendre(B)
    centre, size, t, dt = (0.3, -0.2, 0.7), (0.1, 0.2, 0.4), 0.25, 0.01
    rec = {}

    def bv(_s, x, tt, ddt, face, normal, FIn, QIn, FOut, QOut):
        for k in range(3):
            QOut[k] = x[k]
        QOut[3] = QIn[0]
        QOut[4] = tt + 10 * ddt + 100 * face + 1000 * normal
        FOut[0] = FIn[0] + 0.5
        rec["n"] = rec.get("n", 0) + 1

    cb = oracle.BV_FN(bv)
    for face in range(6):
        d, side = face // 2, face % 2
        d1, d2 = [e for e in range(3) if e != d]
        QIn = np.zeros((B, B, 5))
        QIn[..., 0] = np.arange(B * B).reshape(B, B)  # (i, j) -> i*B + j
        FIn = np.zeros((B, B, 5))
        FIn[..., 0] = -np.arange(B * B).reshape(B, B)
        rec.clear()
        FOut, QOut = oracle.boundary_conditions(s, cb, FIn.ravel(), QIn.ravel(), centre, size, t, dt, face)
        FOut, QOut = FOut.reshape(B, B, 5), QOut.reshape(B, B, 5)
        assert rec["n"] == B * B
        # the face block is indexed (i, j) with i along d2 and j along d1,
        # the layout of the extrapolated boundary blocks
        exp = np.zeros((B, B, 3))
        exp[..., d] = centre[d] + (side - 0.5) * size[d]
        exp[..., d2] = (centre[d2] + size[d2] * (nodes - 0.5))[:, None]
        exp[..., d1] = (centre[d1] + size[d1] * (nodes - 0.5))[None, :]
        assert np.abs(QOut[..., :3] - exp).max() < 1e-15
        assert np.array_equal(QOut[..., 3], QIn[..., 0])
        assert np.all(QOut[..., 4] == t + 10 * dt + 100 * face + 1000 * d)
        assert np.array_equal(FOut[..., 0], FIn[..., 0] + 0.5)


def test_reflecting_wall_state_and_flux(oracle):
    order = 3
    B = order + 1
    s = _euler(oracle, order)
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.5, 2.0, (B, B))
    u = rng.uniform(-1, 1, (B, B, 3))
    p = rng.uniform(0.5, 2.0, (B, B))
    QIn = np.concatenate([rho[..., None], rho[..., None] * u,
                          (p / 0.4 + 0.5 * rho * (u ** 2).sum(-1))[..., None]], axis=-1)
    FIn = rng.normal(size=(B, B, 5))
    for face in range(6):
        d = face // 2
        FOut, QOut = oracle.boundary_conditions(s, None, FIn.ravel(), QIn.ravel(), (0.5,) * 3, (0.1,) * 3,
                                                0.0, 0.01, face)
        QOut, FOut = QOut.reshape(B, B, 5), FOut.reshape(B, B, 5)
        Qm = QIn.copy()
        Qm[..., 1 + d] *= -1
        assert np.array_equal(QOut, Qm)
        assert np.abs(FOut - euler_flux(Qm, d)).max() < 1e-14
        # no mass through the wall; the wall flux's mass entry is minus the inside one
        assert np.abs(FOut[..., 0] + euler_flux(QIn, d)[..., 0]).max() < 1e-15


def test_solution_adjustment_hands_each_node_to_the_callback(oracle):
    order = 3
    B = order + 1
    s = _euler(oracle, order)
    nodes, _ = gauss_legendre(B)
    centre, dx = (0.5, 0.25, -0.75), (0.1, 0.2, 0.3)

    def adj(_s, x, t, dt, Q):
        Q[0], Q[1], Q[2] = x[0], x[1], x[2]
        Q[3] += t
        Q[4] = dt

    cb = oracle.ADJ_FN(adj)
    luh = np.arange(B ** 3 * 5, dtype=float)
    out = oracle.solution_adjustment(s, cb, luh, centre, dx, 0.5, 0.125).reshape(B, B, B, 5)
    zz, yy, xx = np.meshgrid(*(nodes,) * 3, indexing="ij")
    for k, grid in enumerate((xx, yy, zz)):
        assert np.abs(out[..., k] - (centre[k] + dx[k] * (grid - 0.5))).max() < 1e-15
    assert np.array_equal(out[..., 3], luh.reshape(B, B, B, 5)[..., 3] + 0.5)
    assert np.all(out[..., 4] == 0.125)


def test_entropy_wave_adjustment(oracle):
    """t = 0: the entropy wave at every node; t > 0: Q + dt/2 (Q_0 - Q)."""
    order, n = 3, 2
    B = order + 1
    s = _euler(oracle, order)
    X = node_coords(order, n)
    exact = euler_state(1.0 + 0.2 * np.sin(2 * np.pi * X.sum(-1)), (1.0, 1.0, 1.0), 1.0)
    dx = 1.0 / n
    rng = np.random.default_rng(1)
    for c in range(n ** 3):
        i, j, k = c % n, (c // n) % n, c // (n * n)
        centre = ((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx)
        q0 = oracle.solution_adjustment(s, None, np.zeros(B ** 3 * 5), centre, (dx,) * 3, 0.0, 0.0)
        assert np.abs(q0 - exact[c].ravel()).max() < 1e-14
        q = rng.uniform(0.5, 2.0, B ** 3 * 5)
        q1 = oracle.solution_adjustment(s, None, q, centre, (dx,) * 3, 0.3, 0.02)
        assert np.abs(q1 - (q + 0.5 * 0.02 * (exact[c].ravel() - q))).max() < 1e-14


def _rusanov(QL, QR, d):
    """Rusanov flux of two constant states (the smax of riemannSolverNonlinear
    over both states' eigenvalues |u_d| + c), from the textbook flux."""
    def smax(Q):
        rho, u = Q[0], Q[1:4] / Q[0]
        p = 0.4 * (Q[4] - 0.5 * rho * (u ** 2).sum())
        return abs(u[d]) + np.sqrt(1.4 * p / rho)
    s = max(smax(QL), smax(QR))
    return 0.5 * (euler_flux(QL, d) + euler_flux(QR, d)) - 0.5 * s * (QR - QL)


def test_reflecting_box_first_step_closed_form(oracle):
    """A uniform flow in a reflecting box: the predictor is exact, interior
    faces and the volume integral cancel, and one step changes the wall cells
    only, by dt/(dx w_i) phi_i(0) (F*_L - F) at a left wall and
    dt/(dx w_i) phi_i(1) (F - F*_R) at a right wall, F* the Rusanov flux of
    the inside state and its mirror (outside state on the left of faces 0, 2, 4)."""
    from refutil import lagrange
    order, n = 3, 3
    B = order + 1
    Q = euler_state(1.0, (0.5, -0.3, 0.2), 1.0)
    luh0 = np.tile(Q, n ** 3 * B ** 3)
    g = oracle.Grid(_euler(oracle, order), n, transmissive=True)
    g.set_user(True, False)
    g.set_solution(luh0)
    dts, _ = g.step(1)
    dt, dx = dts[0], 1.0 / n
    nodes, w = gauss_legendre(B)
    phi, _ = lagrange(nodes)
    phi0 = np.array([phi(i, 0.0) for i in range(B)])
    phi1 = np.array([phi(i, 1.0) for i in range(B)])
    exp = np.broadcast_to(Q, (n, n, n, B, B, B, 5)).copy()  # (k, j, i, z, y, x, var)
    for d in range(3):
        Qm = Q.copy()
        Qm[1 + d] *= -1
        F = euler_flux(Q, d)
        left = dt / (dx * w) * phi0  # per node index along d
        right = dt / (dx * w) * phi1
        dL = np.multiply.outer(left, _rusanov(Qm, Q, d) - F)
        dR = np.multiply.outer(right, F - _rusanov(Q, Qm, d))
        cell_ax, node_ax = 2 - d, 5 - d
        for first, dq in ((True, dL), (False, dR)):
            idx = [slice(None)] * 6
            idx[cell_ax] = 0 if first else n - 1
            shape = [1] * 6 + [5]
            shape[node_ax] = B
            shape[cell_ax] = 1
            sub = exp[tuple(idx)]
            shp = shape[:cell_ax] + shape[cell_ax + 1:]
            exp[tuple(idx)] = sub + dq.reshape(shp)
    got = g.solution().reshape(n, n, n, B, B, B, 5)
    assert np.abs(got - exp).max() < 1e-12 * np.abs(exp).max()  # rounding of O(1) sums scaled by dt/(dx w)
    # the walls act: the wall cells moved
    assert np.abs(got - Q).max() > 1e-3
    # transmissive boundaries leave a uniform flow unchanged
    t = oracle.Grid(_euler(oracle, order), n, transmissive=True)
    t.set_solution(luh0)
    t.step(1)
    assert np.abs(t.solution() - luh0).max() < 1e-12 * np.abs(luh0).max()


def test_adjustment_runs_after_the_update_at_t_plus_dt(oracle):
    order, n = 2, 3
    s = _euler(oracle, order)
    X = node_coords(order, n)
    luh0 = euler_state(1.0 + 0.1 * np.cos(2 * np.pi * X[..., 0]), (0.3, 0.2, 0.1), 1.0).ravel()
    a = oracle.Grid(s, n)
    a.set_user(False, True)
    a.set_solution(luh0)
    dts, _ = a.step(1)
    b = oracle.Grid(s, n)
    b.set_solution(luh0)
    b.step(1)
    ub = b.solution().reshape(n ** 3, -1)
    dx = 1.0 / n
    for c in range(n ** 3):
        i, j, k = c % n, (c // n) % n, c // (n * n)
        centre = ((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx)
        ub[c] = oracle.solution_adjustment(s, None, ub[c], centre, (dx,) * 3, dts[0], dts[0])
    assert np.array_equal(a.solution(), ub.ravel())


def test_grid_adjust_sets_the_initial_condition(oracle):
    order, n = 3, 2
    g = oracle.Grid(_euler(oracle, order), n)
    g.set_user(False, True)
    g.set_solution(np.zeros(n ** 3 * (order + 1) ** 3 * 5))
    g.adjust()
    X = node_coords(order, n)
    exact = euler_state(1.0 + 0.2 * np.sin(2 * np.pi * X.sum(-1)), (1.0, 1.0, 1.0), 1.0)
    assert np.abs(g.solution() - exact.ravel()).max() < 1e-14
