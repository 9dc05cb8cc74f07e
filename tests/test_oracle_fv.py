"""The finite-volume part of the oracle added for the Godunov scheme and the
stand-alone finite-volumes solver, pinned before it is trusted:

* godunov::c::solutionUpdate (godunov.cpph:35-143) on a constant state
  (interior steady, the boundary-face fluxes in the first ghost layer, the
  admissible dt in closed form);
* against the separately pinned MUSCL-Hancock update on data where every
  minmod slope vanishes (a checkerboard), where MUSCL-Hancock reduces to the
  first-order scheme: identical new patch and, at the same CFL factor,
  identical dt_max_allowed;
* a single one-dimensional jump: only the two jump cells change, and their
  changes add up to -dt/h (F(qR) - F(qL)) with the Euler flux that the
  reference's testPDEFluxes vector pins (the Rusanov dissipation cancels);
* the stand-alone grid driver: splitting the domain into patches changes
  nothing (a 2x2x2 grid of 4^3-subcell patches against one 8^3 patch with
  the same periodic or transmissive ghosts, several steps, with the dt of
  every step), the first dt in closed form, and the time stamp.
"""
import numpy as np
import pytest

from refutil import GAMMA, euler_flux


def _state(rho, u, v, w, p):
    return np.stack([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def _centres(S, g, L):
    return (np.arange(S) - g + 0.5) / L


def _smooth(Z, Y, X):
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (X + 2 * Y)) * np.cos(2 * np.pi * Z)
    u = 0.3 + 0.1 * np.cos(2 * np.pi * Y)
    v = -0.2 + 0.1 * np.sin(2 * np.pi * Z)
    w = 0.1 * np.cos(2 * np.pi * X)
    p = 1.0 + 0.1 * np.sin(2 * np.pi * (X - Z))
    return _state(rho, u, v, w, p)


def test_godunov_constant_state(oracle):
    s = oracle.make_solver(3)
    L, g = 6, 1
    S = L + 2 * g
    q = _state(np.array(1.2), np.array(0.25), np.array(-0.1), np.array(0.15), np.array(2.0))
    P = np.zeros((S, S, S, 5))
    inn = slice(g, g + L)
    P[inn, inn, inn] = q
    for ax in range(3):  # face ghosts only; edges and corners stay zero (never read)
        Pm = np.moveaxis(P, ax, 0)
        Pm[0, inn, inn] = q
        Pm[S - 1, inn, inn] = q
    dx, dt = 0.1, 1e-3
    new, dtm = oracle.godunov_update(s, L, g, P, dx, dt, cfl=0.9)
    new = new.reshape(P.shape)
    # (q + c F) - c F: steady up to the rounding of the two additions
    assert np.abs(new[inn, inn, inn] - P[inn, inn, inn]).max() <= 1e-14
    rho, u = q[0], q[1:4] / q[0]
    pr = (GAMMA - 1) * (q[4] - 0.5 * rho * (u @ u))
    c = np.sqrt(GAMMA * pr / rho)
    assert abs(dtm - 0.9 / 3 * (dx / L) / (np.abs(u).max() + c)) <= 1e-14 * dtm
    c_dt = dt / (dx / L)
    for d in range(3):
        F = euler_flux(q, d)
        lo, hi = [inn] * 3, [inn] * 3
        lo[2 - d], hi[2 - d] = 0, S - 1
        assert np.allclose(new[tuple(lo)], q - c_dt * F, rtol=0, atol=1e-14)
        assert np.allclose(new[tuple(hi)], q + c_dt * F, rtol=0, atol=1e-14)
    assert np.array_equal(new[0, 0], P[0, 0])


def test_godunov_is_muscl_without_slopes(oracle):
    """A checkerboard deviation in every variable makes every minmod slope
    zero (neighbouring differences alternate in sign); MUSCL-Hancock's
    half-step then leaves the face states at the cell values and its update
    is the first-order one."""
    s = oracle.make_solver(3)
    L, g = 5, 2
    S = L + 2 * g
    c = _centres(S, g, L)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    base = _smooth(Z, Y, X)
    idx = np.arange(S)
    sign = (-1.0) ** (idx[:, None, None] + idx[None, :, None] + idx[None, None, :])
    P = base.mean(axis=(0, 1, 2)) + 0.05 * sign[..., None] * np.array([1.0, 0.3, -0.2, 0.25, 0.8])
    dx, dt = 0.2, 2e-3
    m_new, m_dt = oracle.muscl_update(s, L, g, P, dx, dt)
    g_new, g_dt = oracle.godunov_update(s, L, g, P, dx, dt, cfl=0.99)  # MUSCL-Hancock's cflFactor
    assert np.array_equal(m_new, g_new)
    assert m_dt == g_dt
    assert np.abs(g_new - P.ravel()).max() > 1e-4  # it did change


def test_godunov_single_jump(oracle):
    s = oracle.make_solver(3)
    L, g = 6, 1
    S = L + 2 * g
    qL = _state(np.array(1.0), np.array(0.2), np.array(0.0), np.array(0.0), np.array(1.0))
    qR = _state(np.array(0.5), np.array(-0.1), np.array(0.0), np.array(0.0), np.array(0.4))
    xs = np.arange(S)
    P = np.where((xs < g + L // 2)[None, None, :, None], qL, qR) * np.ones((S, S, 1, 1))
    dx, dt = 0.3, 1e-3
    new, _ = oracle.godunov_update(s, L, g, P, dx, dt)
    new = new.reshape(P.shape)
    inn = slice(g, g + L)
    change = (new - P)[inn, inn, inn]
    j = L // 2  # interior index of the first right cell
    others = np.delete(change, [j - 1, j], axis=2)
    assert np.abs(others).max() == 0.0
    assert np.abs(change - change[0:1, 0:1]).max() == 0.0  # every row alike
    fd = dt / (dx / L) * (euler_flux(qR, 0) - euler_flux(qL, 0))
    assert np.allclose(change[0, 0, j - 1] + change[0, 0, j], -fd, rtol=1e-12, atol=1e-15)
    # each side alone: -dt/h (F* - F(q)) with the (separately pinned) Rusanov flux
    fstar, _, _ = oracle.fv_rusanov(s, qL, qR, 0)
    assert np.allclose(change[0, 0, j - 1], -dt / (dx / L) * (fstar - euler_flux(qL, 0)), rtol=1e-12, atol=1e-15)


def _grid_interior(n, L, fn):
    """[ncells][L^3][5] at subcell centres of [0,1]^3, cell c = (k n + j) n + i."""
    c = (np.arange(n * L) + 0.5) / (n * L)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    G = fn(Z, Y, X)  # global (Z,Y,X,5)
    return G, G.reshape(n, L, n, L, n, L, 5).transpose(0, 2, 4, 1, 3, 5, 6).reshape(n ** 3, L ** 3 * 5)


def _to_global(interior, n, L):
    return interior.reshape(n, n, n, L, L, L, 5).transpose(0, 3, 1, 4, 2, 5, 6).reshape(n * L, n * L, n * L, 5)


def _fill_ghosts(P, g, L, periodic, oracle):
    S = L + 2 * g
    inn = slice(g, g + L)
    if periodic:
        for ax in range(3):
            Pm = np.moveaxis(P, ax, 0)
            for k in range(g):
                Pm[k, inn, inn] = Pm[L + k, inn, inn]
                Pm[g + L + k, inn, inn] = Pm[g + k, inn, inn]
        return P
    flat = P.ravel()
    for f in range(6):
        flat = oracle.fv_boundary_transmissive(L, g, 5, flat, f)
    return flat.reshape(S, S, S, 5)


@pytest.mark.parametrize("periodic", [True, False])
def test_fvgrid_patch_split_changes_nothing(oracle, periodic):
    s = oracle.make_solver(3)
    s.cfl = 0.9
    n, L, g = 2, 4, 1
    G, interior = _grid_interior(n, L, _smooth)
    grid = oracle.FVGrid(s, oracle.FV_GODUNOV, L, n, periodic=periodic)
    grid.set_solution(interior)
    steps = 3
    dts = grid.step(steps)
    # one patch of (nL)^3 subcells with the same boundary treatment
    Lb = n * L
    S = Lb + 2 * g
    P = np.zeros((S, S, S, 5))
    inn = slice(g, g + Lb)
    P[inn, inn, inn] = G
    dt = oracle.fv_stable_dt(s, Lb, g, P, (1.0, 1.0, 1.0), 0.9)
    ref_dts = []
    for _ in range(steps):
        ref_dts.append(dt)
        P = _fill_ghosts(P, g, Lb, periodic, oracle)
        new, _ = oracle.godunov_update(s, Lb, g, P, (1.0, 1.0, 1.0), dt)
        P = new.reshape(P.shape)
        dt = oracle.fv_stable_dt(s, Lb, g, P, (1.0, 1.0, 1.0), 0.9)
    assert np.allclose(dts, ref_dts, rtol=1e-15, atol=0)
    got = _to_global(grid.solution(), n, L)
    assert np.array_equal(got, P[inn, inn, inn]) or np.abs(got - P[inn, inn, inn]).max() < 1e-14
    assert abs(grid.time() - sum(ref_dts)) <= 1e-15 * grid.time()
    assert grid.dt() == dt or abs(grid.dt() - dt) <= 1e-15 * dt


def test_fvgrid_first_dt_closed_form(oracle):
    s = oracle.make_solver(3)
    s.cfl = 0.9
    n, L = 3, 5
    q = np.array([1.1, 0.4, -0.2, 0.1, 2.2])
    grid = oracle.FVGrid(s, oracle.FV_GODUNOV, L, n, periodic=True)
    grid.set_solution(np.tile(q, n ** 3 * L ** 3))
    rho, u = q[0], q[1:4] / q[0]
    c = np.sqrt(GAMMA * (GAMMA - 1) * (q[4] - 0.5 * rho * (u @ u)) / rho)
    assert abs(grid.dt() - 0.9 / 3 * (1.0 / (n * L)) / (np.abs(u).max() + c)) <= 1e-14 * grid.dt()
    grid.step(2)
    assert np.abs(grid.solution() - np.tile(q, n ** 3 * L ** 3)).max() <= 1e-14
