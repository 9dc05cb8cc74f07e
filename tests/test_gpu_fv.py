"""Finite volumes on the GPU against the oracle's restatement of the
reference's Godunov scheme and FiniteVolumesSolver time step.

Bars: the single-patch Godunov update (godunov.cpph:35) against the oracle
on every value of the patch -- interior, the boundary-face fluxes the
reference adds to the first ghost layer, untouched ghosts bit-equal -- and
dt_max_allowed (1e-12 of each variable's scale; dt 1e-13); the limiting grid
with limiter-type godunov against the oracle's driver (statuses identical
after every step, dt 1e-13, solution 1e-12); the device-resident FV grid
against the oracle's patch-and-ghost-layer driver after every step on
periodic and transmissive grids whose shapes exercise the kernel's tiling
(one tile, several tiles along x and y, several z chunks, a single cell):
dt of every step 1e-13, time stamp, solution 1e-12; at the full workload
size (54^3 cells of 10^3 subcells), conservation of every conserved total on
the periodic grid, positivity and the time stamp as the sum of the dts.
"""
import numpy as np
import pytest

from refutil import blast_cells, fv_cells_to_global, fv_grid_state, fv_smooth, rel_max

pytestmark = pytest.mark.gpu


def _var_err(a, b):
    a, b = np.asarray(a).reshape(-1, 5), np.asarray(b).reshape(-1, 5)
    return max(np.abs(a[:, v] - b[:, v]).max() / max(np.abs(b[:, v]).max(), 1e-300) for v in range(5))


@pytest.mark.parametrize("L,g", [(4, 1), (10, 1), (7, 1), (5, 2)])
def test_godunov_patch_vs_oracle(exa, oracle, L, g):
    print(f"case: single-patch Godunov vs oracle, L={L} g={g}", flush=True)
    rng = np.random.default_rng(40 + L + g)
    S = L + 2 * g
    c = (np.arange(S) - g + 0.5) / L
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    P = fv_smooth(Z, Y, X) * (1.0 + 0.01 * rng.random((S, S, S, 1)))
    new_in = P * (1.0 + 1e-3 * rng.random(P.shape))  # luh_new need not equal luh
    s = exa.euler_solver(3)
    dx, dt = (0.3, 0.25, 0.2), 2e-4
    got, dtm = exa.godunov_update(s, P, dx, dt, L, g, luh_new=new_in)
    ref_new, ref_dt = oracle.godunov_update(oracle.make_solver(3), L, g, P, dx, dt, cfl=0.9, luh_new=new_in)
    assert _var_err(got, ref_new) <= 1e-12
    changed = np.abs(ref_new - new_in.ravel()) > 0
    assert np.array_equal(got[~changed], new_in.ravel()[~changed])  # untouched ghosts stay as they are
    assert abs(dtm - ref_dt) <= 1e-13 * ref_dt


@pytest.mark.parametrize("n,periodic,radius,steps", [(6, False, 0.2, 6), (6, True, 0.3, 5)])
def test_limiting_grid_godunov_vs_oracle(exa, oracle, n, periodic, radius, steps):
    order = 3
    print(f"case: limiting grid (godunov) vs oracle, N={order} {n}^3 periodic={periodic} {steps} steps", flush=True)
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    luh = blast_cells(order, n, nodes=nodes, radius=radius)
    lim_ic = blast_cells(order, n, subcells=True, radius=radius)
    g = exa.LimitingGrid(exa.euler_solver(order), n, boundary=exa.BC_PERIODIC if periodic else exa.BC_TRANSMISSIVE,
                         fv_scheme=exa.FV_GODUNOV)
    g.upload(luh, lim_ic)
    o = oracle.LimitingGrid(oracle.make_solver(order), n, periodic=periodic, fv_scheme=oracle.FV_GODUNOV)
    o.set_solution(luh, lim_ic)
    st0, _ = g.status()
    assert np.array_equal(st0, o.status())
    irr_prev = 0
    for k in range(steps):
        _, dt, _ = g.time()
        assert abs(dt - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        g.step(1)
        ost, _, oirr = o.step(1)
        st, irr = g.status()
        assert np.array_equal(st, ost[0]), f"limiter status after step {k}"
        assert (irr - irr_prev) == int(oirr[0]), f"irregular change in step {k}"
        irr_prev = irr
        assert _var_err(g.solution(), o.solution()) <= 1e-12, f"solution after step {k}"
    st, _ = g.status()
    L = 2 * order + 1
    for c in np.nonzero(st >= 2)[0][:8]:
        a = g.limiter_solution(int(c)).reshape(L + 4, L + 4, L + 4, 5)[2:-2, 2:-2, 2:-2]
        b = o.limiter_solution(int(c)).reshape(L + 2, L + 2, L + 2, 5)[1:-1, 1:-1, 1:-1]
        assert rel_max(a, b) <= 1e-11
    assert (st >= 2).any()
    g.close()


# (n, L, periodic): one tile; several tiles along x (X = 70 > 62) and y;
# several z chunks (Z = 40); a single cell (every neighbour wraps to itself)
CASES = [((3, 2, 4), 4, True), ((3, 2, 4), 4, False), ((14, 3, 2), 5, True), ((2, 3, 8), 5, False),
         ((2, 2, 8), 5, True), ((1, 1, 1), 3, True)]


@pytest.mark.parametrize("n,L,periodic", CASES)
def test_fvgrid_vs_oracle(exa, oracle, n, L, periodic):
    steps = 4
    print(f"case: FV grid vs oracle, cells {n} L={L} periodic={periodic}, {steps} steps", flush=True)
    _, cells = fv_grid_state(n, L, fv_smooth)
    s = exa.euler_solver(3)
    g = exa.FVGrid(s, L, n, boundary=exa.BC_PERIODIC if periodic else exa.BC_TRANSMISSIVE)
    g.upload(cells)
    os_ = oracle.make_solver(3)
    o = oracle.FVGrid(os_, oracle.FV_GODUNOV, L, n, periodic=periodic)
    o.set_solution(cells)
    assert np.array_equal(g.solution(), cells.ravel())
    for k in range(steps):
        t, dt, nst = g.time()
        assert nst == k
        assert abs(dt - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        assert abs(t - o.time()) <= 1e-13 * max(o.time(), 1e-300)
        g.step(1)
        o.step(1)
        err = _var_err(g.solution(), o.solution())
        assert err <= 1e-12, f"solution after step {k}: {err:.3e}"
    t, dt, nst = g.time()
    assert nst == steps and abs(t - o.time()) <= 1e-13 * o.time() and abs(dt - o.dt()) <= 1e-13 * o.dt()
    g.close()


def test_fvgrid_init_euler_matches_host(exa):
    print("case: FV grid device initial condition vs host formula, 3x2x2 cells L=6", flush=True)
    n, L = (3, 2, 2), 6
    g = exa.FVGrid(exa.euler_solver(3), L, n)
    g.init_euler(rho0=1.0, amp=0.2, k=(1, 2, 1), vel=(0.5, -0.25, 0.1), p=1.3)

    def fn(Z, Y, X):
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * (X + 2 * Y + Z))
        one = np.ones_like(rho)
        E = 1.3 / 0.4 + 0.5 * rho * (0.25 + 0.0625 + 0.01)
        return np.stack([rho, 0.5 * rho, -0.25 * rho, 0.1 * rho, E * one], -1)

    _, cells = fv_grid_state(n, L, fn)
    assert np.abs(g.solution() - cells.ravel()).max() <= 1e-14 * 4
    g.close()


def test_fvgrid_full_size_conservation(exa):
    """The FV workload at its real size: 54^3 cells of 10^3 subcells (540^3),
    periodic entropy wave + 8 seeded modes, 2 steps."""
    n, L = 54, 10
    print(f"case: FV grid {n}^3 cells L={L} (full workload), 2 steps", flush=True)
    g = exa.FVGrid(exa.euler_solver(3), L, n)
    g.init_euler(rho0=1.0, amp=0.2, k=(1, 1, 1), vel=(1.0, 1.0, 1.0), p=1.0, nmodes=8, mode_amp=0.01, seed=2024)
    u0 = g.solution().reshape(-1, 5)
    tot0 = u0.sum(axis=0)
    t0, dt0, _ = g.time()
    g.step(1)
    t1, dt1, _ = g.time()
    g.step(1)
    t2, dt2, n2 = g.time()
    u2 = g.solution().reshape(-1, 5)
    assert t0 == 0.0 and t1 == dt0 and t2 == dt0 + dt1 and n2 == 2
    assert dt0 > 0 and dt1 > 0 and dt2 > 0 and abs(dt1 - dt0) < 0.05 * dt0
    assert np.isfinite(u2).all() and u2[:, 0].min() > 0
    tot2 = u2.sum(axis=0)
    for v in range(5):
        scale = np.abs(u0[:, v]).sum()
        assert abs(tot2[v] - tot0[v]) <= 1e-11 * scale, f"total of variable {v}"
    assert np.abs(u2 - u0).max() > 1e-6  # it evolved
    g.close()


@pytest.mark.parametrize("n,L,periodic", [((2, 3, 2), 5, True), ((3, 2, 2), 7, False), ((3, 3, 3), 5, True),
                                          ((4, 1, 2), 7, True)])
def test_fvgrid_musclhancock_vs_oracle(exa, oracle, n, L, periodic):
    """The stand-alone MUSCL-Hancock FV solver (FiniteVolumesSolver with
    musclhancock.cpph:49): per-cell patches with two ghost layers on the
    device against the oracle's patch-and-ghost-layer driver after every
    step (dt 1e-13, time stamp, solution 1e-12 of each variable's scale)."""
    steps = 4
    print(f"case: MUSCL-Hancock FV grid vs oracle, cells {n} L={L} periodic={periodic}, {steps} steps", flush=True)
    _, cells = fv_grid_state(n, L, fv_smooth)
    s = exa.euler_solver(3)
    g = exa.FVGrid(s, L, n, boundary=exa.BC_PERIODIC if periodic else exa.BC_TRANSMISSIVE,
                   scheme=exa.FV_MUSCLHANCOCK)
    g.upload(cells)
    o = oracle.FVGrid(oracle.make_solver(3), oracle.FV_MUSCLHANCOCK, L, n, periodic=periodic)
    o.set_solution(cells)
    assert np.array_equal(g.solution(), cells.ravel())
    for k in range(steps):
        t, dt, nst = g.time()
        assert nst == k
        assert abs(dt - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        assert abs(t - o.time()) <= 1e-13 * max(o.time(), 1e-300)
        g.step(1)
        o.step(1)
        err = _var_err(g.solution(), o.solution())
        assert err <= 1e-12, f"solution after step {k}: {err:.3e}"
    g.close()


def test_fvgrid_musclhancock_init_and_refusals(exa):
    """The device initial condition through the patches; patch sizes other
    than 5 and 7 are refused."""
    n, L = (2, 2, 3), 5
    g = exa.FVGrid(exa.euler_solver(3), L, n, scheme=exa.FV_MUSCLHANCOCK)
    g.init_euler(rho0=1.0, amp=0.2, k=(1, 2, 1), vel=(0.5, -0.25, 0.1), p=1.3)
    h = exa.FVGrid(exa.euler_solver(3), L, n)
    h.init_euler(rho0=1.0, amp=0.2, k=(1, 2, 1), vel=(0.5, -0.25, 0.1), p=1.3)
    assert np.array_equal(g.solution(), h.solution())
    assert g.time()[1] == h.time()[1]
    g.close()
    h.close()
    with pytest.raises(exa.ExaError, match="status 2"):
        exa.FVGrid(exa.euler_solver(3), 6, 2, scheme=exa.FV_MUSCLHANCOCK)
