"""The linear ADER-DG kernels on the device against the oracle, called
through the C-ABI (exa_*_linear and the linear mode of exa_grid):

* spaceTimePredictorLinear, elasticity at orders 2..5, with and without
  point sources: faces, the five temporaries (lQi, lFi, gradQ, lQhi, lFhi)
  -- each within 1e-12 of its own scale;
* riemannSolverLinear / volumeIntegralLinear / surfaceIntegralLinear on
  seeded inputs (1e-12 of scale) and on the reference's Euler 3-D golden
  vectors at the bars the oracle meets there;
* the device grid in linear mode (Cauchy-Kowalevski predictor + linear
  volume, linear Riemann, linear surface + update, global dt) against the
  oracle's grid driver in linear mode: dt to 1e-13, every variable to 1e-12
  of its scale after several steps.
"""
import numpy as np
import pytest

from refutil import elastic_cells, golden

pytestmark = pytest.mark.gpu

GE = golden("generic_euler_3d.json")


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _elastic_state(rng, B, cells=1):
    """Random smooth-scaled elastic states: stresses/velocities O(1), material
    (lambda, mu, rho) in a physical range."""
    q = rng.standard_normal((cells, B ** 3, 12))
    q[..., 9] = 2.0 + rng.random((cells, B ** 3))      # lambda
    q[..., 10] = 1.0 + rng.random((cells, B ** 3))     # mu
    q[..., 11] = 1.5 + rng.random((cells, B ** 3))     # rho
    return q.reshape(-1)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
@pytest.mark.parametrize("point_sources", [False, True])
def test_predictor_linear_vs_oracle(exa, oracle, order, point_sources):
    print(f"case: linear predictor N={order} point_sources={point_sources}", flush=True)
    rng = np.random.default_rng(100 + order)
    B = order + 1
    luh = _elastic_state(rng, B)
    dx, dt = (0.11, 0.13, 0.17), 2.3e-3
    PSi = rng.standard_normal((B + 1) * B ** 3 * 9) if point_sources else None
    dev = exa.space_time_predictor_linear(exa.elastic_solver(order), luh, dx, dt, temps=True, PSi=PSi)
    ref = oracle.space_time_predictor_linear(oracle.elastic_solver(order), luh, dx, dt, PSi)
    for key in ("lQhbnd", "lFhbnd", "lQi", "lFi", "gradQ", "lQhi", "lFhi"):
        err = _rel(dev[key], ref[key])
        print(f"  {key}: {err:.2e}", flush=True)
        assert err <= 1e-12, key


def test_predictor_linear_without_temporaries_matches(exa):
    rng = np.random.default_rng(7)
    s = exa.elastic_solver(3)
    luh = _elastic_state(rng, 4)
    a = exa.space_time_predictor_linear(s, luh, (0.1, 0.1, 0.1), 1e-3, temps=True)
    b = exa.space_time_predictor_linear(s, luh, (0.1, 0.1, 0.1), 1e-3, temps=False)
    # two instantiations of the kernel: the compiler may contract differently
    assert _rel(a["lQhbnd"], b["lQhbnd"]) <= 1e-14 and _rel(a["lFhbnd"], b["lFhbnd"]) <= 1e-14


@pytest.mark.parametrize("order", [2, 3, 4, 5])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_riemann_linear_vs_oracle(exa, oracle, order, d):
    print(f"case: linear Riemann N={order} d={d}", flush=True)
    rng = np.random.default_rng(10 * order + d)
    B = order + 1
    QL = _elastic_state(rng, B)[: B * B * 12]
    QR = _elastic_state(rng, B)[: B * B * 12]
    FL, FR = exa.riemann_solver_linear(exa.elastic_solver(order), QL, QR, 1e-3, d)
    oFL, oFR = oracle.riemann_linear(oracle.elastic_solver(order), QL, QR, 1e-3, d)
    assert _rel(FL, oFL) <= 1e-12 and _rel(FR, oFR) <= 1e-12


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_volume_and_surface_linear_vs_oracle(exa, oracle, order):
    print(f"case: linear volume/surface N={order}", flush=True)
    rng = np.random.default_rng(50 + order)
    B = order + 1
    lFhi = rng.standard_normal(4 * B ** 3 * 9)
    assert _rel(exa.volume_integral_linear(exa.elastic_solver(order), lFhi),
                oracle.volume_integral_linear(oracle.elastic_solver(order), lFhi[: 3 * B ** 3 * 9])) <= 1e-12
    lduh = rng.standard_normal(B ** 3 * 9)
    F = rng.standard_normal(6 * B * B * 9)
    dx = (0.21, 0.17, 0.13)
    assert _rel(exa.surface_integral_linear(exa.elastic_solver(order), lduh, F, dx),
                oracle.surface_integral_linear(oracle.elastic_solver(order), lduh, F, dx)) <= 1e-12


def test_linear_euler_golden_vectors(exa):
    """GenericEulerKernelTest3D.cpp testVolumeIntegralLinear /
    testSurfaceIntegralLinear / testRiemannSolverLinear (order 3, 5 variables,
    1 parameter, zero coefficient matrix)."""
    print("case: linear Euler golden vectors", flush=True)
    s = exa.euler_solver(3, npar=1, linear=True)
    lFhi = np.zeros((4, 64, 5))
    lFhi[0, :, 1], lFhi[1, :, 2], lFhi[2, :, 3] = 1.0, 2.0, 3.0
    ref = np.array(GE["testVolumeIntegralLinear"]["lduh"])
    assert _rel(exa.volume_integral_linear(s, lFhi.ravel()), ref) <= 1e-14
    F = np.zeros((6, 16, 5))
    F[0, :, 1] = F[1, :, 1] = 1.0
    F[2, :, 2] = F[3, :, 2] = 1.0
    F[4, :, 3] = F[5, :, 3] = 1.0
    ref = np.array(GE["testSurfaceIntegralLinear"]["lduh"])
    assert _rel(exa.surface_integral_linear(s, np.arange(320) / 10.0, F.ravel(), 0.1), ref) <= 1e-13
    QL, QR = np.zeros((16, 6)), np.zeros((16, 6))
    QL[:, :5] = np.array(GE["testRiemannSolver"]["QL"]).reshape(16, 5)
    QR[:, :5] = np.array(GE["testRiemannSolver"]["QR"]).reshape(16, 5)
    FL, FR = exa.riemann_solver_linear(s, QL.ravel(), QR.ravel(), 1.40831757919882352703e-03, 1)
    assert _rel(FL, np.array(GE["testRiemannSolverLinear"]["FL"])) <= 1e-13
    assert _rel(FR, np.array(GE["testRiemannSolverLinear"]["FR"])) <= 1e-13


def test_linear_solver_support_flags(exa):
    assert exa.supported(exa.elastic_solver(4, linear=True))
    assert not exa.supported(exa.elastic_solver(7, linear=True))
    with pytest.raises(exa.ExaError):
        exa.Grid(exa.elastic_solver(7, linear=True), 3)


@pytest.mark.parametrize("order,n,steps", [(2, (4, 3, 2), 4), (4, (3, 2, 3), 3)])
def test_linear_grid_transmissive_vs_oracle(exa, oracle, order, n, steps):
    print(f"case: linear elastic grid transmissive vs oracle, N={order} {n} {steps} steps", flush=True)
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    # elastic_cells builds a cube; take a non-cubic grid's cells from a cube of the largest extent
    m = max(n)
    full = elastic_cells(order, m, seed=7, nodes=nodes, width=0.2).reshape(m, m, m, -1)
    luh = np.ascontiguousarray(full[:n[2], :n[1], :n[0]]).reshape(-1)
    g = exa.Grid(exa.elastic_solver(order, linear=True), n, boundary=exa.BC_TRANSMISSIVE)
    g.upload(luh)
    g.step(steps)
    g.synchronize()
    dts = g.dt_history(steps)
    s = oracle.elastic_solver(order)
    s.linear = 1
    o = oracle.Grid(s, n, transmissive=True)
    o.set_solution(luh)
    odts, _ = o.step(steps, nthreads=0)
    assert np.abs(dts - odts).max() <= 1e-13 * odts.max()
    a, b = g.solution().reshape(-1, 12), o.solution().reshape(-1, 12)
    err = max(np.abs(a[:, v] - b[:, v]).max() / max(np.abs(b[:, v]).max(), 1e-300) for v in range(12))
    print(f"  max rel error per variable {err:.3e}", flush=True)
    assert err <= 1e-12
    g.close()


@pytest.mark.parametrize("order,n,steps", [(2, 4, 5), (3, 3, 4), (4, 3, 3), (5, 2, 3)])
def test_linear_grid_vs_oracle(exa, oracle, order, n, steps):
    print(f"case: linear elastic grid vs oracle, N={order} {n}^3 {steps} steps", flush=True)
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    luh = elastic_cells(order, n, seed=5, nodes=nodes, width=0.15)
    g = exa.Grid(exa.elastic_solver(order, linear=True), n)
    g.upload(luh)
    g.step(steps)
    g.synchronize()
    dts = g.dt_history(steps)
    s = oracle.elastic_solver(order)
    s.linear = 1
    o = oracle.Grid(s, n)
    o.set_solution(luh)
    odts, _ = o.step(steps, nthreads=0)
    assert np.abs(dts - odts).max() <= 1e-13 * odts.max()
    a, b = g.solution().reshape(-1, 12), o.solution().reshape(-1, 12)
    err = max(np.abs(a[:, v] - b[:, v]).max() / max(np.abs(b[:, v]).max(), 1e-300) for v in range(12))
    print(f"  max rel error per variable {err:.3e}", flush=True)
    assert err <= 1e-12
    g.close()


@pytest.mark.parametrize("order", [1, 3, 4, 5, 7])
def test_point_source_projection_vs_oracle(exa, oracle, order):
    print(f"case: pointSource projection N={order}", flush=True)
    rng = np.random.default_rng(300 + order)
    B = order + 1
    n = (B + 1) * B ** 3
    center, dx = np.array([0.3, -0.2, 1.1]), np.array([0.05, 0.07, 0.04])
    PSi = rng.standard_normal(n * 9)
    # per-node positions: most inside the cell, some outside
    X0 = center - dx / 2 + rng.uniform(-0.2, 1.2, (n, 3)) * dx
    dev = exa.point_source_projection(exa.elastic_solver(order), PSi, X0, center, dx)
    ref = oracle.point_source_projection(oracle.elastic_solver(order), PSi, X0, center, dx)
    assert _rel(dev, ref) <= 1e-12
    assert np.array_equal(dev == 0.0, ref == 0.0)
    assert all(exa.lib().exa_quadrature_node(order, i) == oracle.table(oracle.T_NODES, order, i) for i in range(B))


def test_linear_elastic_2d_golden_vectors_on_device(exa, oracle):
    """ElasticityKernelTest2D.cpp testRiemannSolverLinear / testVolumeIntegralLinear /
    testSurfaceIntegralLinear (order 4, non-zero coefficient matrix) through the
    z-invariant 3-D embedding the oracle tests use, on the device, at the
    reference test's eps 1e-9 (ElasticityKernelTest.h:69)."""
    print("case: elastic 2-D golden vectors (embedded) on the device", flush=True)
    EL = golden("elasticity_2d.json")
    s = exa.elastic_solver(4)
    B = 5
    w = np.array([oracle.table(oracle.T_WEIGHTS, 4, i) for i in range(B)])
    g = EL["testRiemannSolverLinear"]
    q2L = np.concatenate([np.array(g["qL_IN"]).reshape(B, 9), np.array(g["paramL_IN"]).reshape(B, 3)], 1)
    q2R = np.concatenate([np.array(g["qR_IN"]).reshape(B, 9), np.array(g["paramR_IN"]).reshape(B, 3)], 1)
    QL, QR = np.broadcast_to(q2L, (B, B, 12)).ravel(), np.broadcast_to(q2R, (B, B, 12)).ravel()
    FL, FR = exa.riemann_solver_linear(s, QL, QR, 1.916666666666667e-4, 1)
    for got, key in ((FL, "FL_OUT"), (FR, "FR_OUT")):
        ref = np.array(g[key]).reshape(B, 9)
        for i in range(B):
            assert np.abs(got.reshape(B, B, 9)[i] - ref).max() <= 1e-9, key
    g = EL["testVolumeIntegralLinear"]
    l2 = np.array(g["lFhi_IN"]).reshape(2, B, B, 9)
    lFhi = np.zeros((4, B, B, B, 9))
    lFhi[0], lFhi[1] = l2[0][None], l2[1][None]
    lduh = exa.volume_integral_linear(s, lFhi.ravel()).reshape(B, B, B, 9)
    ref = np.array(g["lduh_OUT"]).reshape(B, B, 9)
    for i in range(B):
        assert np.abs(lduh[i] / w[i] - ref).max() <= 1e-12 * np.abs(ref).max()
    g = EL["testSurfaceIntegralLinear"]
    F2 = np.array(g["lFbnd_IN"]).reshape(4, B, 9)
    F = np.zeros((6, B, B, 9))
    F[0], F[1], F[2], F[3] = F2[0][None], F2[1][None], F2[2][None], F2[3][None]
    lduh_in = w[:, None, None, None] * np.array(g["lduh_IN"]).reshape(B, B, 9)[None]
    out = exa.surface_integral_linear(s, lduh_in.ravel(), F.ravel(), (38.4615384615385, 35.7142857142857, 1.0))
    ref = np.array(g["lduh_OUT"]).reshape(B, B, 9)
    for i in range(B):
        assert np.abs(out.reshape(B, B, B, 9)[i] / w[i] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_linear_elastic_2d_predictor_golden_on_device(exa, oracle):
    """ElasticityKernelTest2D.cpp testSpaceTimePredictorLinear (order 4, NCP
    only) through the z-invariant 3-D embedding, on the device: lQhi, lFhi and
    both face arrays at every z slice, at the reference test's eps 1e-9 and
    1e-9 of each array's scale."""
    print("case: elastic 2-D predictor golden (embedded) on the device", flush=True)
    from test_oracle_linear import _check_stp_linear_elastic_golden
    g = golden("elasticity_2d.json")["testSpaceTimePredictorLinear"]
    s = exa.elastic_solver(4)
    B = 5
    luh2 = np.concatenate([np.array(g["luh_IN"]).reshape(B, B, 9), np.array(g["param_IN"]).reshape(B, B, 3)], 2)
    luh = np.broadcast_to(luh2, (B, B, B, 12)).copy()
    dx = (38.4615384615385, 35.7142857142857, 1.0)
    dt = 0.813172798364530
    o = exa.space_time_predictor_linear(s, luh.ravel(), dx, dt, temps=True)
    _check_stp_linear_elastic_golden(g, o["lQhi"], o["lFhi"], o["lQhbnd"], o["lFhbnd"])
