"""The subcell limiter on the GPU against the oracle's restatement of the
reference's limiter path.

Bars: projections against the reference's limiter golden vectors at the
reference test's tolerance (1e-13, LimiterKernelTest.cpp:30); single-cell
min/max, DMP and MUSCL-Hancock against the oracle (1e-13 of the data scale,
DMP verdicts identical); the limiting grid against the oracle's driver on
blast problems (limiter status of every cell identical after every step, the
same steps with an irregular change, dt to 1e-13, solution to 1e-12 of each
variable's scale); at the full 81^3 size, size-independent invariants.
"""
import json
import os

import numpy as np
import pytest

from refutil import blast_cells, random_euler_cells, rel_max

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "limiter_3d.json")))


def test_projections_golden(exa):
    print("case: limiter projections vs limiter_testdata.cpp (B=4, L=7, g=0)", flush=True)
    lim = exa.project_on_fv(np.array(GOLD["testFromLuhConversion"]["luh_in"]), 4, 0)
    assert np.abs(lim - np.array(GOLD["testFromLuhConversion"]["lim_out"])).max() <= 1e-13
    luh = exa.project_on_dg(np.array(GOLD["testToLuhConversion"]["lim_in"]), 4, 0)
    assert np.abs(luh - np.array(GOLD["testToLuhConversion"]["luh_out"])).max() <= 1e-13


@pytest.mark.parametrize("order", [2, 3])
def test_single_cell_min_max_dmp_vs_oracle(exa, oracle, order):
    print(f"case: local min/max + DMP vs oracle, N={order}", flush=True)
    rng = np.random.default_rng(11 + order)
    B = order + 1
    for trial in range(6):
        luh = random_euler_cells(rng, 1, order)
        mn, mx = exa.cell_local_min_max(luh, B)
        omn, omx = oracle.cell_local_min_max(order, 5, luh)
        assert rel_max(mn, omn) < 1e-13 and rel_max(mx, omx) < 1e-13
        # boundary extrema around the local ones: some pass, some fail the DMP
        scale = 0.0 if trial % 2 else 0.05
        bmin = np.tile(omn * (1 + scale), 6) + rng.normal(0, 1e-3, 30)
        bmax = np.tile(omx * (1 - scale), 6) + rng.normal(0, 1e-3, 30)
        ok, bmn, bmx = exa.dmp(luh, B, bmin, bmax)
        ook, obmn, obmx = oracle.dmp(order, 5, luh, bmin, bmax)
        assert ok == ook
        assert rel_max(bmn, obmn) < 1e-13 and rel_max(bmx, obmx) < 1e-13
    # projections with ghost layers
    luh = random_euler_cells(rng, 1, order)
    lim = exa.project_on_fv(luh, B, 2)
    olim = oracle.project_on_fv(order, 5, 2, luh)
    assert rel_max(lim, olim) < 1e-13
    back = exa.project_on_dg(olim, B, 2)
    assert rel_max(back, oracle.project_on_dg(order, 5, 2, olim)) < 1e-13


def _blast_patch(rng, L, g):
    S = L + 2 * g
    c = (np.arange(S) - g + 0.5) / L
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    rho = 1.0 + 0.01 * rng.random(Z.shape)
    p = np.where(np.sqrt((X - 0.3) ** 2 + (Y - 0.6) ** 2 + (Z - 0.5) ** 2) < 0.4, 10.0, 0.1)
    P = np.stack([rho, 0.1 * rho, -0.2 * rho, 0.05 * rho, p / 0.4 + 0.5 * rho * 0.0525], -1)
    idx = np.arange(S)
    out = (idx < g) | (idx >= g + L)
    oz, oy, ox = np.meshgrid(out, out, out, indexing="ij")
    P[(oz.astype(int) + oy + ox) >= 2] = 0.0  # edges and corners are never filled
    return P


@pytest.mark.parametrize("order", [2, 3])
def test_muscl_patch_vs_oracle(exa, oracle, order):
    print(f"case: MUSCL-Hancock patch vs oracle, N={order}", flush=True)
    rng = np.random.default_rng(5)
    L, g = 2 * order + 1, 2
    s_exa, s_ora = exa.euler_solver(order), oracle.make_solver(order)
    for dt in (1e-4, 1e-3):
        P = _blast_patch(rng, L, g)
        new, dtmax = exa.muscl_update(s_exa, P, 0.1, dt, return_dt=True)
        new = new.reshape(P.shape)
        onew, odtmax = oracle.muscl_update(s_ora, L, g, P, 0.1, dt)
        onew = onew.reshape(P.shape)
        assert abs(dtmax - odtmax) <= 1e-14 * odtmax
        inner = (slice(g, g + L),) * 3
        assert np.isfinite(new[inner]).all()
        assert np.abs(new[inner] - P[inner]).max() > 1e-8  # it evolved
        for v in range(5):
            assert rel_max(new[inner][..., v], onew[inner][..., v]) < 1e-12


def test_device_blast_ic_matches_host_generator(exa, oracle):
    print("case: device blast initial condition vs host generator, N=3 5^3", flush=True)
    order, n = 3, 5
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    g = exa.LimitingGrid(exa.euler_solver(order), n)
    g.init_blast(seed=7, radius=0.2)
    dev = g.solution()
    host = blast_cells(order, n, seed=7, nodes=nodes, radius=0.2)
    assert np.abs(dev - host).max() <= 1e-14 * np.abs(host).max()
    g.close()


def _grid_pair(exa, oracle, order, n, periodic, radius, steps):
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    luh = blast_cells(order, n, nodes=nodes, radius=radius)
    lim_ic = blast_cells(order, n, subcells=True, radius=radius)
    g = exa.LimitingGrid(exa.euler_solver(order), n,
                         boundary=exa.BC_PERIODIC if periodic else exa.BC_TRANSMISSIVE)
    g.upload(luh, lim_ic)
    o = oracle.LimitingGrid(oracle.make_solver(order), n, periodic=periodic)
    o.set_solution(luh, lim_ic)
    return g, o


@pytest.mark.parametrize("n,periodic,radius,steps", [(6, True, 0.3, 6), (6, False, 0.2, 6), (8, False, 0.25, 5)])
def test_limiting_grid_vs_oracle(exa, oracle, n, periodic, radius, steps):
    order = 3
    print(f"case: limiting grid vs oracle, N={order} {n}^3 periodic={periodic} r={radius} {steps} steps", flush=True)
    g, o = _grid_pair(exa, oracle, order, n, periodic, radius, steps)
    st0, _ = g.status()
    assert np.array_equal(st0, o.status())
    irr_gpu_prev = 0
    for k in range(steps):
        t, dt, _ = g.time()
        assert abs(dt - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        g.step(1)
        ost, odt, oirr = o.step(1)
        st, irr = g.status()
        assert np.array_equal(st, ost[0]), f"limiter status after step {k}: {np.nonzero(st != ost[0])[0][:10]}"
        assert (irr - irr_gpu_prev) == int(oirr[0]), f"irregular change in step {k}"
        irr_gpu_prev = irr
        sol, osol = g.solution().reshape(-1, 5), o.solution().reshape(-1, 5)
        for v in range(5):
            err = np.abs(sol[:, v] - osol[:, v]).max() / max(np.abs(osol[:, v]).max(), 1e-300)
            assert err <= 1e-12 or np.abs(sol[:, v] - osol[:, v]).max() <= 1e-14, f"var {v} step {k}: {err:.3e}"
    assert irr_gpu_prev > 0  # the run exercised rollback and recomputation
    # FV patches of limited cells agree in their interior.  Ghost layers are
    # refilled from the neighbours before every update; the reference's update
    # also adds the boundary-face fluxes to the first ghost layer of its new
    # patch (musclhancock.cpph:267-329), values nothing reads before the next
    # refill and the device does not write.
    st, _ = g.status()
    S = 2 * order + 1 + 4
    inner = (slice(2, S - 2),) * 3
    for c in np.nonzero(st >= 2)[0][:8]:
        a = g.limiter_solution(int(c)).reshape(S, S, S, 5)[inner]
        b = o.limiter_solution(int(c)).reshape(S, S, S, 5)[inner]
        assert rel_max(a, b) <= 1e-11
    g.close()


def test_limiter_workload_full_size_invariants(exa):
    """The limiter workload at its real shape (N=3, 81^3, transmissive blast):
    finite and positive, neighbouring statuses within one, a minority of cells
    limited."""
    print("case: limiter workload 81^3 N=3, 5 steps", flush=True)
    n = 81
    g = exa.LimitingGrid(exa.euler_solver(3), n)
    g.init_blast(seed=7, radius=0.2)
    g.step(5)
    st, irr = g.status()
    sol = g.solution().reshape(-1, 5)
    assert np.isfinite(sol).all() and sol[:, 0].min() > 0
    S3 = st.reshape(n, n, n)
    for ax in range(3):
        assert np.abs(np.diff(S3, axis=ax)).max() <= 1
    frac = (st >= 2).mean()
    print(f"  FV-active fraction {frac:.4f}, troubled {(st == 3).mean():.4f}, irregular steps {irr}", flush=True)
    assert 0.0 < frac < 0.5
    g.close()


@pytest.mark.parametrize("helper_layers,cure", [(1, 0), (2, 1), (1, 2)])
def test_limiter_status_machine_injected_cell_on_device(exa, helper_layers, cure):
    """The hand case of tests/test_oracle_limiter.py on the device: one cell
    with negative density at one node on a uniform periodic grid; the
    statuses after set-up and after every step in closed form."""
    print(f"case: injected cell status machine h={helper_layers} cure={cure}", flush=True)
    from test_oracle_limiter import expected_injected_statuses, injected_cell_case
    order, n = 3, 7
    c0 = (3 * n + 2) * n + 4
    luh, lim_ic = injected_cell_case(order, n, c0)
    g = exa.LimitingGrid(exa.euler_solver(order), n, boundary=exa.BC_PERIODIC, helper_layers=helper_layers,
                         steps_till_cured=cure)
    g.upload(luh, lim_ic)
    st, irr = g.status()
    assert np.array_equal(st, expected_injected_statuses(n, c0, helper_layers, cure, 0))
    T = 2 * helper_layers + 1
    for step in range(1, T + cure + 2):
        g.step(1)
        st, irr = g.status()
        assert np.array_equal(st, expected_injected_statuses(n, c0, helper_layers, cure, step)), step
    assert irr == 0
    sol = g.solution().reshape(-1, 5)
    assert np.abs(sol[:, 0] - 1.0).max() < 1e-12 and np.abs(sol[:, 1:4]).max() < 1e-12
    g.close()


@pytest.mark.timeout(600)
def test_limiting_grid_vs_oracle_mid_size(exa, oracle):
    """Statuses bit-exact against the oracle on a 16^3 transmissive blast over
    20 steps (N=3), with the dt of every step and the solution at the end
    (the oracle's limiting driver takes about 40 s of one host core here)."""
    order, n, steps = 3, 16, 20
    print(f"case: limiting grid vs oracle, N={order} {n}^3 blast {steps} steps", flush=True)
    g, o = _grid_pair(exa, oracle, order, n, False, 0.1, steps)
    assert np.array_equal(g.status()[0], o.status())
    irr_prev = 0
    for k in range(steps):
        assert abs(g.time()[1] - o.dt()) <= 1e-13 * o.dt(), f"dt of step {k}"
        g.step(1)
        ost, _, oirr = o.step(1)
        st, irr = g.status()
        assert np.array_equal(st, ost[0]), f"limiter status after step {k}: {np.nonzero(st != ost[0])[0][:10]}"
        assert irr - irr_prev == int(oirr[0]), f"irregular change in step {k}"
        irr_prev = irr
        if k % 5 == 4:
            print(f"  step {k + 1}: statuses identical", flush=True)
    sol, osol = g.solution().reshape(-1, 5), o.solution().reshape(-1, 5)
    for v in range(5):
        assert rel_max(sol[:, v], osol[:, v]) <= 1e-11, v
    print(f"  FV-active {(st >= 2).mean():.3f}, irregular steps {irr_prev}", flush=True)
    assert irr_prev > 0 and (st >= 2).any()
    g.close()
