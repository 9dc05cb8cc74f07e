"""The limiter part of the oracle (CPU restatement of the reference's limiter,
finite-volume and limiting-driver code) pinned before it is trusted:

* projection tables and DG<->FV projections against the reference's limiter
  golden vectors (limiter_testdata.cpp, 3-D block; B=4, L=7, V=5) at the
  reference test's own tolerance (LimiterKernelTest.cpp:30, eps = 1e-13);
* closed forms of the tables (subcell means, Lobatto interpolation, partition
  of unity, mean preservation, exact inverse on P_N);
* DMP / min-max semantics on hand-made cases;
* MUSCL-Hancock and the FV Rusanov flux against closed forms (constant states,
  one-dimensional data, telescoping conservation);
* ghost-layer filling and the transmissive FV boundary (mirror, and the
  reference's lower y/z extraction that reads one layer for both ghosts);
* the limiting driver: with no troubled cell it reproduces the plain ADER-DG
  grid driver bit for bit; with a blast it conserves mass on a periodic grid
  and keeps neighbouring limiter statuses within one of each other.
"""
import json
import os

import numpy as np
import pytest

from refutil import GAMMA, blast_cells, euler_flux, gauss_legendre, lagrange

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "limiter_3d.json")))
EPS_REF = 1e-13  # LimiterKernelTest.cpp:30


def test_project_on_fv_golden(oracle):
    luh = np.array(GOLD["testFromLuhConversion"]["luh_in"])
    lim = oracle.project_on_fv(3, 5, 0, luh)
    ref = np.array(GOLD["testFromLuhConversion"]["lim_out"])
    assert np.abs(lim - ref).max() <= EPS_REF


def test_project_on_dg_golden(oracle):
    lim = np.array(GOLD["testToLuhConversion"]["lim_in"])
    luh = oracle.project_on_dg(3, 5, 0, lim)
    ref = np.array(GOLD["testToLuhConversion"]["luh_out"])
    assert np.abs(luh - ref).max() <= EPS_REF


def test_lobatto_values_golden(oracle):
    luh = np.array(GOLD["testFromLuhConversion"]["luh_in"]).reshape(4, 4, 4, 5)
    _, uh2lob, _ = oracle.limiter_tables(3)
    lob = np.einsum("zyxv,za,yb,xc->abcv", luh, uh2lob, uh2lob, uh2lob).ravel()
    ref = np.array(GOLD["testFromLuhConversion"]["lob_out"])
    assert np.abs(lob - ref).max() <= EPS_REF


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7])
def test_table_closed_forms(oracle, order):
    B, L = order + 1, 2 * order + 1
    uh2lim, uh2lob, lim2uh = oracle.limiter_tables(order)
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(B)])
    phi, _ = lagrange(nodes)
    # uh2lim[k][i] = mean of phi_k over subcell i (high-order quadrature, numpy nodes)
    xq, wq = gauss_legendre(B + 4)
    for i in range(L):
        xs = (i + xq) / L
        for k in range(B):
            assert abs(uh2lim[k, i] - np.dot(wq, [phi(k, x) for x in xs])) < 1e-12
    # partition of unity and mean preservation
    assert np.allclose(uh2lim.sum(0), 1.0, atol=1e-13)
    w = np.array([oracle.table(oracle.T_WEIGHTS, order, i) for i in range(B)])
    assert np.allclose(uh2lim.sum(1) / L, w, atol=1e-13)
    # Lobatto interpolation; nodes are stored in descending order
    lob = np.array([oracle.table(oracle.T_LOB_NODES, order, i) for i in range(B)])
    if B > 1:
        assert lob[0] == 1.0 and lob[-1] == 0.0
    for j in range(B):
        for i in range(B):
            assert abs(uh2lob[j, i] - phi(j, lob[i])) < 1e-12
    # the constrained least-squares fit reproduces P_N exactly
    assert np.abs(uh2lim @ lim2uh - np.eye(B)).max() < 1e-11
    # and it preserves the cell mean of any subcell data
    v = np.random.default_rng(order).normal(size=L)
    assert abs(w @ (lim2uh.T @ v) - v.mean()) < 1e-13


def _cell_poly(order, coeffs):
    B = order + 1
    x, _ = gauss_legendre(B)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([c[0] + c[1] * X + c[2] * Y + c[3] * Z for c in coeffs], -1)


def test_local_min_max_linear_data(oracle):
    """Linear data: the extrema over GL nodes, Lobatto nodes (which include the
    cell corners) and subcell means are attained at the corners."""
    order = 3
    coeffs = [(1.0, 0.1, -0.2, 0.05), (0.2, 0.0, 0.0, 0.0), (-0.3, 0.1, 0.1, 0.1), (0.0, 1.0, 0.0, 0.0),
              (2.5, -0.5, 0.2, 0.3)]
    luh = _cell_poly(order, coeffs)
    mn, mx = oracle.cell_local_min_max(order, 5, luh)
    for v, c in enumerate(coeffs):
        corners = [c[0] + c[1] * a + c[2] * b + c[3] * d for a in (0, 1) for b in (0, 1) for d in (0, 1)]
        assert abs(mn[v] - min(corners)) < 1e-13 and abs(mx[v] - max(corners)) < 1e-13


def test_dmp_semantics(oracle):
    order = 3
    B = order + 1
    luh = np.zeros((B, B, B, 5))
    luh[..., 0] = 1.0
    luh[..., 4] = 2.5
    # boundary extrema equal to the constant state: DMP holds
    bmin = np.tile(np.array([1.0, 0, 0, 0, 2.5]), 6)
    bmax = bmin.copy()
    ok, mn, mx = oracle.dmp(order, 5, luh.ravel(), bmin, bmax)
    assert ok
    # faces 0/1 are overwritten with local / boundary extrema, faces 2..5 kept
    bmin2 = bmin.copy()
    bmin2[2 * 5 + 0] = 0.5
    ok, mn, mx = oracle.dmp(order, 5, luh.ravel(), bmin2, bmax)
    assert ok and abs(mn[0] - 1.0) < 1e-14 and mn[5] == 0.5 and mn[10] == 0.5
    # an undershoot beyond max(relax, scaling*(max-min)) violates the DMP ...
    luh2 = luh.copy()
    luh2[0, 0, 0, 0] = 1.0 - 2e-3
    ok, _, _ = oracle.dmp(order, 5, luh2.ravel(), bmin, bmax)
    assert not ok
    # ... one inside the relaxation does not
    luh3 = luh.copy()
    luh3[1, 1, 1, 0] += 1e-6
    ok, _, _ = oracle.dmp(order, 5, luh3.ravel(), bmin, bmax + np.tile([1e-6, 0, 0, 0, 0], 6), relax=1e-4)
    assert ok


def test_fv_rusanov_closed_form(oracle):
    s = oracle.make_solver(3)
    rng = np.random.default_rng(3)
    for d in range(3):
        qL = np.array([1.0, 0.1, -0.2, 0.3, 2.5]) + 0.1 * rng.random(5)
        qR = np.array([0.8, -0.1, 0.2, 0.1, 2.0]) + 0.1 * rng.random(5)
        fL, fR, smax = oracle.fv_rusanov(s, qL, qR, d)

        def lam(q):
            u = q[d + 1] / q[0]
            p = (GAMMA - 1) * (q[4] - 0.5 * (q[1:4] ** 2).sum() / q[0])
            c = np.sqrt(GAMMA * p / q[0])
            return max(abs(u - c), abs(u), abs(u + c))

        sm = max(lam(qL), lam(qR))
        ref = 0.5 * sm * (qL - qR) + 0.5 * (euler_flux(qL, d) + euler_flux(qR, d))
        assert abs(smax - sm) < 1e-14 * sm
        assert np.allclose(fL, ref, rtol=1e-14, atol=1e-14) and np.array_equal(fL, fR)


def _patch(L, g, fn):
    """Patch (S,S,S,5) with fn(zc,yc,xc) -> state at subcell centres of [0,1]^3
    extended into the ghosts; edge and corner ghosts zero (never filled)."""
    S = L + 2 * g
    c = (np.arange(S) - g + 0.5) / L
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    P = fn(Z, Y, X)
    idx = np.arange(S)
    out = (idx < g) | (idx >= g + L)
    oz, oy, ox = np.meshgrid(out, out, out, indexing="ij")
    P[(oz.astype(int) + oy + ox) >= 2] = 0.0
    return P


def test_muscl_constant_state_is_steady(oracle):
    s = oracle.make_solver(3)
    L, g = 7, 2
    q = np.array([1.2, 0.3, -0.1, 0.2, 3.0])
    P = _patch(L, g, lambda Z, Y, X: np.broadcast_to(q, Z.shape + (5,)).copy())
    new, dtm = oracle.muscl_update(s, L, g, P, 0.1, 1e-3)
    new = new.reshape(P.shape)
    assert np.array_equal(new[g:g + L, g:g + L, g:g + L], P[g:g + L, g:g + L, g:g + L])
    # every face sees the uniform state (all minmod slopes vanish, also next to
    # the zero edge ghosts), so the admissible dt is 0.99/3 h / (|u_x| + c)
    rho, u = q[0], q[1:4] / q[0]
    p = 0.4 * (q[4] - 0.5 * rho * (u @ u))
    c = np.sqrt(1.4 * p / rho)
    assert abs(dtm - 0.99 / 3 * (0.1 / L) / (np.abs(u).max() + c)) <= 1e-14 * dtm
    # the face loops also apply the boundary-face flux to the first ghost layer
    # of luh_new (musclhancock.cpph:267-329): low ghost q - dt/h F_d(q), high
    # ghost q + dt/h F_d(q); the second layer and edge/corner ghosts are untouched
    from refutil import euler_flux
    c_dt = 1e-3 / (0.1 / L)
    inn = slice(g, g + L)
    for d in range(3):
        F = euler_flux(q, d)
        lo, hi = [slice(None)] * 3, [slice(None)] * 3
        for axis in range(3):  # array axes (z, y, x); direction d is axis 2 - d
            if axis != 2 - d:
                lo[axis] = hi[axis] = inn
        lo[2 - d], hi[2 - d] = g - 1, g + L
        assert np.allclose(new[tuple(lo)], q - c_dt * F, rtol=0, atol=1e-14)
        assert np.allclose(new[tuple(hi)], q + c_dt * F, rtol=0, atol=1e-14)
        lo[2 - d], hi[2 - d] = g - 2, g + L + 1
        assert np.array_equal(new[tuple(lo)], P[tuple(lo)]) and np.array_equal(new[tuple(hi)], P[tuple(hi)])
    assert np.array_equal(new[0, 0], P[0, 0]) and np.array_equal(new[g - 1, g - 1], P[g - 1, g - 1])


def test_muscl_one_dimensional_and_conservative(oracle):
    """x-only data made periodic over the patch interior (x ghost layers are
    copies of the opposite interior layers).  Rows away from the patch's y/z
    edges stay identical and conserve their total (both boundary faces see
    identical states, so the update telescopes to zero).  Rows touching a y/z
    edge do not: the edge ghosts are never filled (zero), and minmod slopes of
    the face-ghost cells next to them differ -- the reference's behaviour
    (FiniteVolumesSolver.cpp:477-481 relies on those zeros)."""
    s = oracle.make_solver(3)
    L, g = 7, 2
    S = L + 2 * g

    def fn(Z, Y, X):
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * X)
        p = 1.0 + 0.1 * np.cos(2 * np.pi * X)
        return np.stack([rho, 0.5 * rho, 0 * rho, 0 * rho, p / (GAMMA - 1) + 0.125 * rho], -1)

    P = _patch(L, g, fn)
    inn = slice(g, g + L)
    P[inn, inn, :g] = P[inn, inn, L:L + g]      # periodic in x
    P[inn, inn, g + L:] = P[inn, inn, g:2 * g]
    for ax in (0, 1):                           # y/z ghosts: copies of the rows
        Pm = np.moveaxis(P, ax, 0)
        Pm[:g, inn, inn] = Pm[g:2 * g, inn, inn]
        Pm[g + L:, inn, inn] = Pm[L:L + g, inn, inn]
    dx, dt = 0.1, 1e-3
    new, _ = oracle.muscl_update(s, L, g, P, dx, dt)
    new = new.reshape(S, S, S, 5)
    core = slice(g + 1, g + L - 1)
    inner = new[core, core, inn]
    assert np.abs(inner - inner[0:1, 0:1]).max() == 0.0
    change = inner - P[core, core, inn]
    assert np.abs(change).max() > 1e-6  # it did evolve
    assert np.abs(change[0, 0, :, 2:4]).max() == 0.0
    assert np.abs(change[0, 0].sum(0)).max() < 1e-15 * L
    # an edge row does pick up transverse momentum
    assert np.abs(new[g, g, inn, 2:4] - P[g, g, inn, 2:4]).max() > 0.0


def test_ghost_fill_and_transmissive_boundary(oracle):
    L, g, D = 7, 2, 5
    S = L + 2 * g
    rng = np.random.default_rng(5)
    A = rng.random((S, S, S, D))
    Bn = rng.random((S, S, S, D))
    inn = slice(g, g + L)
    # neighbour at +x: ghost layers pe, pe+1 get neighbour layers pb, pb+1
    out = oracle.fv_ghost_fill(L, g, D, A, Bn, 0, +1).reshape(S, S, S, D)
    assert np.array_equal(out[inn, inn, g + L:], Bn[inn, inn, g:2 * g])
    # neighbour at -z: ghost layers 0, 1 get neighbour layers pe-2, pe-1
    out = oracle.fv_ghost_fill(L, g, D, A, Bn, 2, -1).reshape(S, S, S, D)
    assert np.array_equal(out[:g, inn, inn], Bn[L:L + g, inn, inn])
    # everything else untouched
    mask = np.ones((S, S, S), bool)
    mask[:g, inn, inn] = False
    assert np.array_equal(out[mask], A[mask])
    # transmissive: mirror at -x and +x, +y, +z
    for face in (0, 1, 3, 5):
        out = oracle.fv_boundary_transmissive(L, g, D, A, face).reshape(S, S, S, D)
        d = face // 2
        o = np.moveaxis(out, 2 - d, 0)
        a = np.moveaxis(A, 2 - d, 0)
        if face % 2 == 0:
            assert np.array_equal(o[1, inn, inn], a[g, inn, inn]) and np.array_equal(o[0, inn, inn], a[g + 1, inn, inn])
        else:
            assert np.array_equal(o[g + L, inn, inn], a[g + L - 1, inn, inn])
            assert np.array_equal(o[g + L + 1, inn, inn], a[g + L - 2, inn, inn])
    # lower y and z faces: both ghost layers take interior layer pb
    for face in (2, 4):
        out = oracle.fv_boundary_transmissive(L, g, D, A, face).reshape(S, S, S, D)
        d = face // 2
        o = np.moveaxis(out, 2 - d, 0)
        a = np.moveaxis(A, 2 - d, 0)
        assert np.array_equal(o[0, inn, inn], a[g, inn, inn]) and np.array_equal(o[1, inn, inn], a[g, inn, inn])


def _smooth_ic(oracle, order, n):
    from refutil import node_coords
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    X = node_coords(order, n, nodes=nodes)
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * (X[..., 0] + X[..., 1] + X[..., 2]))
    Q = np.stack([rho, rho, rho, rho, 1.0 / (GAMMA - 1) + 1.5 * rho], -1)
    return np.ascontiguousarray(Q.reshape(-1))


@pytest.mark.parametrize("periodic", [True, False])
def test_limiting_driver_without_troubled_cells_is_plain_dg(oracle, periodic):
    """A smooth flow with a loose DMP never flags a cell: the limiting driver
    must then reproduce the plain ADER-DG driver exactly (periodic, and with
    transmissive boundaries on both)."""
    order, n = 3, 4
    s = oracle.make_solver(order)
    luh = _smooth_ic(oracle, order, n)
    lg = oracle.LimitingGrid(s, n, periodic=periodic, relax=1.0, scaling=1.0)
    lg.set_solution(luh)
    st, dts, irr = lg.step(4)
    assert (st == 0).all() and not irr.any()
    pg = oracle.Grid(s, n, transmissive=not periodic)
    pg.set_solution(luh)
    dts2, _ = pg.step(4)
    assert np.array_equal(dts, dts2)
    assert np.array_equal(lg.solution(), pg.solution())


def test_limiting_driver_blast_invariants(oracle):
    """Blast on a periodic grid: troubled cells appear and trigger rollback and
    recomputation; after every step the DG solution of an FV-active cell is the
    projection of its FV patch, a helper cell's patch is the projection of its
    DG solution, the density stays positive, and face neighbours' statuses
    differ by at most one.  (The DG/FV coupling is not flux-conservative in the
    reference scheme, so total mass is not an invariant here.)"""
    order, n = 3, 6
    s = oracle.make_solver(order)
    B, L, g = order + 1, 2 * order + 1, 2
    S = L + 2 * g
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(B)])
    luh = blast_cells(order, n, nodes=nodes, radius=0.3)
    lim_ic = blast_cells(order, n, subcells=True, radius=0.3)
    grid = oracle.LimitingGrid(s, n, periodic=True)
    grid.set_solution(luh, lim_ic)
    seen_troubled = seen_irregular = False
    for step in range(4):
        st, dts, irr = grid.step(1)
        st = st[0]
        seen_troubled |= bool((st == 3).any())
        seen_irregular |= bool(irr[0])
        sol = grid.solution().reshape(-1, B ** 3 * 5)
        assert np.isfinite(sol).all()
        assert sol.reshape(-1, 5)[:, 0].min() > 0
        for c in np.nonzero(st > 0)[0][:12]:
            lim = grid.limiter_solution(int(c))
            if st[c] >= 2:
                assert np.abs(oracle.project_on_dg(order, 5, g, lim) - sol[c]).max() < 1e-13
            else:
                proj = oracle.project_on_fv(order, 5, g, sol[c]).reshape(S, S, S, 5)
                inner = lim.reshape(S, S, S, 5)[g:g + L, g:g + L, g:g + L]
                assert np.abs(proj[g:g + L, g:g + L, g:g + L] - inner).max() < 1e-13
        S3 = st.reshape(n, n, n)
        for ax in range(3):
            assert np.abs(S3 - np.roll(S3, 1, axis=ax)).max() <= 1
    assert seen_troubled and seen_irregular


def injected_cell_case(order, n, c0):
    """A uniform state (rho=1, u=0, p=1) on an n^3 periodic grid in which cell
    c0's DG solution has rho = -0.01 at one Gauss-Legendre node; every FV
    patch starts uniform (the FV initial condition at subcell centres)."""
    B, L = order + 1, 2 * order + 1
    q = np.array([1.0, 0.0, 0.0, 0.0, 1.0 / (GAMMA - 1.0)])
    luh = np.tile(q, (n ** 3, B ** 3, 1))
    luh[c0, 5, 0] = -0.01
    lim_ic = np.tile(q, (n ** 3, L ** 3, 1))
    return luh.reshape(-1), lim_ic.reshape(-1)


def expected_injected_statuses(n, c0, helper_layers, cure, step):
    """Closed form of the reference's status machine for injected_cell_case
    (ADERDGSolver.cpp:461-462, 2544-2596; LimitingADERDGSolver.cpp:955-994,
    1031-1043): at set-up c0 is not admissible, so it is troubled (T = 2h+1)
    and spreading gives every cell max(0, T - d), d its periodic face-graph
    distance to c0.  Every later solution is uniform and admissible, so no
    cell becomes troubled again: c0 keeps T while its cure counter (cure + 1
    at set-up) runs down, then takes T - 1 and, like every other cell,
    max over its faces of max(own, neighbour) - 1, one less per step.
    Together: status = max(0, min(T - d, T - max(0, step - cure)))."""
    T = 2 * helper_layers + 1
    k = np.arange(n)
    z, y, x = np.meshgrid(k, k, k, indexing="ij")
    cz, cy, cx = c0 // (n * n), (c0 // n) % n, c0 % n
    per = lambda a, b: np.minimum(np.abs(a - b), n - np.abs(a - b))
    d = (per(x, cx) + per(y, cy) + per(z, cz)).ravel()
    return np.maximum(0, np.minimum(T - d, T - max(0, step - cure)))


@pytest.mark.parametrize("helper_layers,cure", [(1, 0), (2, 1), (1, 2)])
def test_limiter_status_machine_injected_cell(oracle, helper_layers, cure):
    """Hand case: one cell with negative density at one node on a uniform
    grid; statuses after set-up and after each step in closed form, no
    irregular change (nothing becomes newly troubled), the state stays
    uniform."""
    order, n = 3, 7
    c0 = (3 * n + 2) * n + 4
    luh, lim_ic = injected_cell_case(order, n, c0)
    g = oracle.LimitingGrid(oracle.make_solver(order), n, periodic=True, helper_layers=helper_layers,
                            steps_till_cured=cure)
    g.set_solution(luh, lim_ic)
    assert np.array_equal(g.status(), expected_injected_statuses(n, c0, helper_layers, cure, 0))
    T = 2 * helper_layers + 1
    for step in range(1, T + cure + 2):
        st, _, irr = g.step(1)
        assert not irr[0]
        assert np.array_equal(st[0], expected_injected_statuses(n, c0, helper_layers, cure, step)), step
    assert (g.status() == 0).all()
    sol = g.solution().reshape(-1, 5)
    assert np.abs(sol[:, 0] - 1.0).max() < 1e-12 and np.abs(sol[:, 1:4]).max() < 1e-12


@pytest.mark.parametrize("periodic", [True, False])
def test_irregular_steps_leave_quiet_cells_as_plain_dg(oracle, periodic):
    """Rollback and local recomputation (LS:1343-1436) give the DG cells their
    own DG update back: a cell that is a DG cell (status below A) before and
    after a step, whose face neighbours all carried a prediction (status below
    troubled) before it, ends the step -- regular or irregular -- with exactly
    the solution the plain ADER-DG driver computes from the same state and dt
    (bit for bit: the same kernels on the same data).  Status-0 cells are not
    rolled back at all; helper cells are rolled back and recomputed to the same
    values.  A rollback that also reached status-0 cells, a helper left at the
    previous level, or a merge that used other face data fails it."""
    order, n = 3, 8
    s = oracle.make_solver(order)
    B = order + 1
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(B)])
    luh = blast_cells(order, n, nodes=nodes, radius=0.2)
    lim_ic = blast_cells(order, n, subcells=True, radius=0.2)
    grid = oracle.LimitingGrid(s, n, periodic=periodic)
    grid.set_solution(luh, lim_ic)
    T = 3
    idx = np.arange(n ** 3).reshape(n, n, n)  # [k][j][i]
    checked_irregular = 0
    counts = []
    for step in range(5):
        luh0, st0, dt = grid.solution(), grid.status(), grid.dt()
        st, _, irr = grid.step(1)
        st1, luh1 = st[0], grid.solution().reshape(n ** 3, -1)
        pg = oracle.Grid(s, n, transmissive=not periodic)
        pg.set_solution(luh0)
        pg.step_dts(np.array([dt]))
        ref = pg.solution().reshape(n ** 3, -1)
        S0 = st0.reshape(n, n, n)
        quiet = np.ones((n, n, n), dtype=bool)
        for ax in range(3):
            for sh in (1, -1):
                nb = np.roll(S0, sh, axis=ax)
                if not periodic:  # a missing neighbour does not disqualify the cell
                    edge = [slice(None)] * 3
                    edge[ax] = 0 if sh == 1 else -1
                    nb = nb.copy()
                    nb[tuple(edge)] = 0
                quiet &= nb < T
        cells = idx[quiet & (S0 < 2) & (st1.reshape(n, n, n) < 2)]  # DG before and after (A = 2)
        assert len(cells) > 0
        assert np.array_equal(luh1[cells], ref[cells]), f"step {step}: {len(cells)} cells"
        counts.append(len(cells))
        checked_irregular += int(irr[0])
    assert checked_irregular > 0


def test_fv_cells_step_from_the_previous_level(oracle):
    """Every cell FV-active before a step that keeps a patch after it
    (regular or irregular, i.e. also after rollback and local recomputation;
    a cell back to DG then holds the projection of it), and every cell that an
    irregular change made FV-active, holds the MUSCL-Hancock
    update (pinned separately) of its previous patch, with its face ghost
    layers filled (pinned separately) from the neighbours' previous FV
    representations -- their previous patch when FV-active, else the
    projection of their previous DG solution.  A recomputation from the
    wrong time level, or with ghosts from the updated neighbours, fails it."""
    order, n = 3, 8
    s = oracle.make_solver(order)
    B, L, g, D, A = order + 1, 2 * order + 1, 2, 5, 2
    S = L + 2 * g
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(B)])
    luh = blast_cells(order, n, nodes=nodes, radius=0.2)
    lim_ic = blast_cells(order, n, subcells=True, radius=0.2)
    grid = oracle.LimitingGrid(s, n, periodic=True)
    grid.set_solution(luh, lim_ic)
    inner = (slice(g, g + L),) * 3

    def cell(i, j, k):
        return ((k % n) * n + (j % n)) * n + (i % n)

    checked = irregular = to_dg = newly = 0
    for step in range(4):
        st0, dt = grid.status(), grid.dt()
        luh0 = grid.solution().reshape(n ** 3, -1)
        prev = {}

        def rep(c):
            if c not in prev:
                if st0[c] >= A:
                    prev[c] = grid.limiter_solution(c).reshape(S, S, S, D).copy()
                else:
                    prev[c] = oracle.project_on_fv(order, D, g, luh0[c]).reshape(S, S, S, D)
            return prev[c]

        fv0 = [c for c in range(n ** 3) if st0[c] >= A]
        for c in fv0:
            rep(c)
            k, j, i = c // (n * n), (c // n) % n, c % n
            for d in range(3):
                for sd in (-1, 1):
                    off = [0, 0, 0]
                    off[d] = sd
                    rep(cell(i + off[0], j + off[1], k + off[2]))
        st, _, irr = grid.step(1)
        irregular += int(irr[0])
        luh1 = grid.solution().reshape(n ** 3, -1)
        # and, after an irregular change, the cells newly FV-active: rolled
        # back to their previous DG solution, projected, recomputed
        new_fv = [c for c in range(n ** 3) if irr[0] and st[0][c] >= A and st0[c] < A]
        newly += len(new_fv)
        for c in fv0 + new_fv:
            if st[0][c] == 0:  # no patch kept
                continue
            k, j, i = c // (n * n), (c // n) % n, c % n
            own = np.zeros((S, S, S, D))
            own[inner] = rep(c)[inner]
            own = own.ravel()
            for d in range(3):
                for sd in (-1, 1):
                    off = [0, 0, 0]
                    off[d] = sd
                    nb = rep(cell(i + off[0], j + off[1], k + off[2]))
                    own = oracle.fv_ghost_fill(L, g, D, own, nb, d, sd)
            exp, _ = oracle.muscl_update(s, L, g, own, 1.0 / n, dt)
            got = grid.limiter_solution(c).reshape(S, S, S, D)
            assert np.array_equal(got[inner], exp.reshape(S, S, S, D)[inner]), (step, c)
            if st[0][c] < A:  # back to DG: its solution is the projection of that update
                assert np.array_equal(luh1[c], oracle.project_on_dg(order, D, g, got.ravel())), (step, c)
                to_dg += 1
            checked += 1
    assert checked > 0 and to_dg > 0 and newly > 0 and irregular > 0


@pytest.mark.parametrize("periodic,helper_layers", [(True, 1), (False, 1), (True, 2), (False, 2)])
def test_statuses_after_spreading_in_closed_form(oracle, periodic, helper_layers):
    """After an irregular change the statuses are spread over T = 2h + 1
    sweeps (Runner.cpp:889-896, LS:392-405): then every cell's status is
    max(0, T - d), d its face-neighbour (L1) distance to the nearest troubled
    cell (status T) -- with periodic wrap-around, or none across a
    transmissive boundary.  Checked after every irregular step of a blast."""
    order, n = 3, 8
    s = oracle.make_solver(order)
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    luh = blast_cells(order, n, nodes=nodes, radius=0.2)
    lim_ic = blast_cells(order, n, subcells=True, radius=0.2)
    grid = oracle.LimitingGrid(s, n, periodic=periodic, helper_layers=helper_layers)
    grid.set_solution(luh, lim_ic)
    T = 2 * helper_layers + 1
    K, J, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    checked = 0
    for _ in range(6):
        st, _, irr = grid.step(1)
        if not irr[0]:
            continue
        S3 = st[0].reshape(n, n, n)
        d = np.full((n, n, n), 10 ** 9)
        for k, j, i in np.argwhere(S3 == T):
            dk, dj, di = np.abs(K - k), np.abs(J - j), np.abs(I - i)
            if periodic:
                dk, dj, di = np.minimum(dk, n - dk), np.minimum(dj, n - dj), np.minimum(di, n - di)
            d = np.minimum(d, dk + dj + di)
        assert np.array_equal(S3, np.maximum(0, T - d))
        checked += 1
    assert checked > 0
