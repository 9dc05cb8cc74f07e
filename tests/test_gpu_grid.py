"""Multi-step parity of the device-resident grid (exa_grid) with the oracle's
restatement of the reference's non-fused time stepping.

Bar: <= 1e-12 relative per DOF after 10 steps; identical dt sequence (1e-13)
and identical per-cell Picard iteration counts.  At the full 81^3, N=5 size
the oracle cannot run the whole grid, so size-independent properties are
checked instead: conservation, the time step, and one step of sampled cells
recomputed from their neighbourhoods.
"""
import numpy as np
import pytest

from refutil import cell_weights, euler_state, node_coords, rel_max, rel_per_dof

pytestmark = pytest.mark.gpu

TOL = 1e-12


def run_pair(exa, oracle, order, n, steps, nmodes=0, seed=0, amp=0.2):
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(amp=amp, nmodes=nmodes, mode_amp=0.01, seed=seed)
    luh0 = g.solution()
    g.step(steps)
    g.synchronize()
    gpu = g.solution()
    t, dt, nst = g.time()
    dts = g.dt_history(steps + 1)
    it_gpu = g.picard_iterations()
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n)
    o.set_solution(luh0)
    odts, oit = o.step(steps, nthreads=0, record_iters=True)
    return dict(g=g, luh0=luh0, gpu=gpu, t=t, dt=dt, steps=nst, dts=dts, it=it_gpu, o=o, cpu=o.solution(),
                odts=odts, ot=o.time(), oit=oit)


def test_initial_condition_matches_closed_form(exa):
    """The device IC equals rho = 1 + 0.2 sin(2 pi (x+y+z)), u = 1, p = 1 at the GL nodes."""
    order, n = 3, 6
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler()
    X = node_coords(order, n)
    exact = euler_state(1.0 + 0.2 * np.sin(2 * np.pi * X.sum(-1)), (1.0, 1.0, 1.0), 1.0)
    assert rel_per_dof(g.solution(), exact.reshape(-1)) < 1e-14


def test_config1_entropy_wave_27cube_n3_10steps(exa, oracle):
    """Plain method, small: Euler N=3, 27^3 periodic, entropy wave, CFL 0.9, 10 steps."""
    print("case: grid 27^3 N=3 10 steps", flush=True)
    r = run_pair(exa, oracle, 3, 27, 10)
    assert r["steps"] == 10
    assert rel_per_dof(r["gpu"], r["cpu"]) < TOL
    assert np.abs(r["dts"][:10] - r["odts"]).max() <= 1e-13 * r["odts"].max()
    assert abs(r["t"] - r["ot"]) <= 1e-13 * r["ot"]
    # Picard counts of the predictor that ran last (for the step after the 10th)
    assert np.array_equal(r["it"], r["o"].picard_iterations())


@pytest.mark.parametrize("order,n,steps,seed", [(2, 7, 10, 1), (4, 5, 10, 2), (5, 5, 10, 3), (6, 4, 6, 4), (7, 3, 4, 5)])
def test_grid_with_fourier_modes_vs_oracle(exa, oracle, order, n, steps, seed):
    """Entropy wave plus 8 seeded Fourier modes (the precision workload's initial data)."""
    print(f"case: grid {n}^3 N={order} {steps} steps modes seed={seed}", flush=True)
    r = run_pair(exa, oracle, order, n, steps, nmodes=8, seed=seed)
    assert rel_per_dof(r["gpu"], r["cpu"]) < TOL
    assert np.abs(r["dts"][:steps] - r["odts"]).max() <= 1e-13 * r["odts"].max()


def test_picard_iterations_identical_every_step(exa, oracle):
    """Per-cell Picard iteration counts of every predictor pass are identical."""
    order, n, steps = 5, 4, 6
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(nmodes=8, mode_amp=0.01, seed=11)
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n)
    o.set_solution(g.solution())
    _, oit = o.step(steps, record_iters=True)
    g.step(0)  # initial time step size + first predictor
    for k in range(steps):
        assert np.array_equal(g.picard_iterations(), oit[k]), k
        g.step(1)
    assert np.array_equal(g.picard_iterations(), o.picard_iterations())


def test_conservation_on_device(exa):
    order, n, steps = 3, 8, 20
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(nmodes=8, mode_amp=0.01, seed=5)
    w3 = cell_weights(order)[None, ..., None]
    B = order + 1
    before = (w3 * g.solution().reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    g.step(steps)
    after = (w3 * g.solution().reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    assert np.abs(after - before).max() < 1e-12 * np.abs(before).max()


def neighbourhood_update(oracle, s, luh, n, c, dx, dt):
    """One reference step of cell c from its own and its 6 neighbours' states."""
    B = s.order + 1
    nb3 = B ** 3 * 5
    i, j, k = c % n, (c // n) % n, c // (n * n)

    def cid(a, b, cc):
        return ((cc % n) * n + (b % n)) * n + (a % n)

    nbrs = {"c": c, "x-": cid(i - 1, j, k), "x+": cid(i + 1, j, k), "y-": cid(i, j - 1, k),
            "y+": cid(i, j + 1, k), "z-": cid(i, j, k - 1), "z+": cid(i, j, k + 1)}
    pred = {key: oracle.space_time_predictor(s, luh[v * nb3:(v + 1) * nb3], dx, dt) for key, v in nbrs.items()}
    lduh = oracle.volume_integral(s, pred["c"]["lFhi"], dx)
    fsz = B * B * 5
    lF = pred["c"]["lFhbnd"].copy()
    for d, (lo, hi) in enumerate((("x-", "x+"), ("y-", "y+"), ("z-", "z+"))):
        # lower face 2d: neighbour is the left state; upper face 2d+1: cell is left
        QLn = pred[lo]["lQhbnd"].reshape(6, fsz)[2 * d + 1]
        FLn = pred[lo]["lFhbnd"].reshape(6, fsz)[2 * d + 1]
        Qc_lo = pred["c"]["lQhbnd"].reshape(6, fsz)[2 * d]
        _, FR = oracle.riemann(s, FLn, lF.reshape(6, fsz)[2 * d], QLn, Qc_lo, dt, d)
        lF.reshape(6, fsz)[2 * d] = FR
        QRn = pred[hi]["lQhbnd"].reshape(6, fsz)[2 * d]
        FRn = pred[hi]["lFhbnd"].reshape(6, fsz)[2 * d]
        Qc_hi = pred["c"]["lQhbnd"].reshape(6, fsz)[2 * d + 1]
        FL, _ = oracle.riemann(s, lF.reshape(6, fsz)[2 * d + 1], FRn, Qc_hi, QRn, dt, d)
        lF.reshape(6, fsz)[2 * d + 1] = FL
    lduh = oracle.surface_integral(s, lduh, lF, dx)
    return oracle.solution_update(s, luh[c * nb3:(c + 1) * nb3], lduh, dt), pred["c"]["iters"]


def test_precision_workload_full_size_81cube_n5(exa, oracle):
    """Full-size precision workload (81^3, N=5, entropy wave + 8 modes): one step of
    64 sampled cells recomputed by the oracle from their neighbourhoods (<=1e-12),
    the first dt recomputed by the oracle, conservation over two steps."""
    order, n = 5, 81
    B = order + 1
    nb3 = B ** 3 * 5
    print("case: full-size 81^3 N=5 sampled parity", flush=True)
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(nmodes=8, mode_amp=0.01, seed=2024)
    luh0 = g.solution()
    s = oracle.make_solver(order, 5, 0)
    dx = np.full(3, 1.0 / n)
    g.step(1)
    dts = g.dt_history(2)
    it = g.picard_iterations()  # predictor of step 2 (on the updated state)
    luh1 = g.solution()
    # dt^0: min over all cells, recomputed on the CPU
    og = oracle.Grid(s, n)
    og.set_solution(luh0)
    dt0 = og.stable_dt()
    assert abs(dts[0] - dt0) <= 1e-13 * dt0
    rng = np.random.default_rng(0)
    cells = rng.choice(n ** 3, 64, replace=False)
    for c in cells:
        ref, _ = neighbourhood_update(oracle, s, luh0, n, int(c), dx, dts[0])
        assert rel_per_dof(luh1[c * nb3:(c + 1) * nb3], ref) < TOL, int(c)
    # Picard counts of sampled cells for the next predictor (on luh1 with dt^1)
    for c in cells[:16]:
        o = oracle.space_time_predictor(s, luh1[c * nb3:(c + 1) * nb3], dx, dts[1])
        assert it[c] == o["iters"], int(c)
    w3 = cell_weights(order)[None, ..., None]
    tot0 = (w3 * luh0.reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    g.step(1)
    tot2 = (w3 * g.solution().reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    assert np.abs(tot2 - tot0).max() < 1e-11 * np.abs(tot0).max()
    g.close()


@pytest.mark.parametrize("order,n,steps", [(3, 6, 8), (5, 4, 5)])
def test_transmissive_grid_vs_oracle(exa, oracle, order, n, steps):
    """Transmissive boundaries (boundaryConditions with outside = inside state):
    device grid vs the oracle driver, which solves the boundary Riemann problems."""
    print(f"case: transmissive grid vs oracle, N={order} {n}^3 {steps} steps", flush=True)
    g = exa.Grid(exa.euler_solver(order), n, boundary=exa.BC_TRANSMISSIVE)
    g.init_euler(amp=0.2, nmodes=8, mode_amp=0.01, seed=5)
    luh0 = g.solution()
    g.step(steps)
    g.synchronize()
    gpu = g.solution()
    o = oracle.Grid(oracle.make_solver(order, 5, 0), n, transmissive=True)
    o.set_solution(luh0)
    odts, oit = o.step(steps, nthreads=0, record_iters=True)
    dts = g.dt_history(steps)
    assert np.abs(dts - odts).max() <= 1e-13 * odts.max()
    assert np.array_equal(g.picard_iterations(), o.picard_iterations())
    cpu = o.solution()
    assert rel_per_dof(gpu.reshape(-1, 5)[:, [0, 4]], cpu.reshape(-1, 5)[:, [0, 4]]) <= TOL
    assert rel_max(gpu, cpu) <= TOL
    g.close()


def test_grid_no_time_averaging_vs_oracle(exa, oracle):
    """A grid whose solver sets noTimeAveraging, against the oracle's own
    noTimeAveraging branch, 5 steps."""
    order, n, steps = 3, 5, 5
    print(f"case: grid {n}^3 N={order} noTimeAveraging {steps} steps", flush=True)
    s = exa.euler_solver(order)
    s.no_time_averaging = 1
    g = exa.Grid(s, n)
    g.init_euler(amp=0.2, nmodes=8, mode_amp=0.01, seed=9)
    luh0 = g.solution()
    g.step(steps)
    g.synchronize()
    o = oracle.Grid(oracle.make_solver(order, 5, 0, no_time_averaging=True), n)
    o.set_solution(luh0)
    odts, _ = o.step(steps, nthreads=0)
    assert rel_per_dof(g.solution(), o.solution()) < TOL
    assert np.abs(g.dt_history(steps) - odts).max() <= 1e-13 * odts.max()
    g.close()
