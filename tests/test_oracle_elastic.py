"""The oracle's elastic PDE (velocity-stress, non-conservative product form)
pinned before it is trusted:

* the non-conservative product equals sum_d B_d grad_d Q with coefficient
  matrices written independently from the velocity-stress equations;
* the eigenvalue list equals the spectrum of those matrices (+-cp, +-cs, 0);
* the whole nonlinear ADER-DG path with useNCP (predictor gradients, NCP in
  the Picard loop and in the Rusanov solver) converges at high order to an
  exact plane P-wave in a homogeneous medium;
* the workload's material stays in its stated ranges.
"""
import numpy as np
import pytest

from refutil import elastic_cells, elastic_material, elastic_matrices, node_coords


def _state(rng, lam, mu, rho):
    q = rng.normal(size=9)
    return np.concatenate([q, [lam, mu, rho]])


def test_ncp_matches_matrix_form(oracle):
    s = oracle.elastic_solver(3)
    rng = np.random.default_rng(1)
    for _ in range(10):
        rho = rng.uniform(2, 3)
        cp = rng.uniform(4, 6)
        cs = cp / np.sqrt(3)
        mu = rho * cs * cs
        lam = rho * cp * cp - 2 * mu
        Q = _state(rng, lam, mu, rho)
        grad = rng.normal(size=(3, 9))
        Bx, By, Bz = elastic_matrices(lam, mu, rho)
        ref = Bx @ grad[0] + By @ grad[1] + Bz @ grad[2]
        got = oracle.ncp(s, Q, grad)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_eigenvalues_are_the_spectrum_of_the_system(oracle):
    s = oracle.elastic_solver(3)
    rho, cp = 2.4, 5.0
    cs = cp / np.sqrt(3)
    mu = rho * cs * cs
    lam = rho * cp * cp - 2 * mu
    Q = _state(np.random.default_rng(2), lam, mu, rho)
    mats = elastic_matrices(lam, mu, rho)
    for d in range(3):
        ev = np.sort(oracle.eigenvalues(s, Q, d))
        spec = np.sort(np.linalg.eigvals(mats[d]).real)
        assert np.allclose(ev, spec, atol=1e-12 * cp)
        assert abs(ev.max() - cp) < 1e-14 * cp


def _p_wave(oracle, order, n, t, lam, mu, rho, amp=0.1):
    """u = amp sin(2 pi (x - cp t)), sxx = -rho cp u, syy = szz = -(lam / cp) u."""
    nodes = np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])
    X = node_coords(order, n, nodes=nodes)
    cp = np.sqrt((lam + 2 * mu) / rho)
    u = amp * np.sin(2 * np.pi * (X[..., 0] - cp * t))
    z = np.zeros_like(u)
    one = np.ones_like(u)
    Q = np.stack([-rho * cp * u, -(lam / cp) * u, -(lam / cp) * u, z, z, z, u, z, z, lam * one, mu * one, rho * one],
                 -1)
    return Q.reshape(-1)


def test_plane_p_wave_converges(oracle):
    order = 3
    rho, cp = 2.5, 5.0
    cs = cp / np.sqrt(3)
    mu = rho * cs * cs
    lam = rho * cp * cp - 2 * mu
    s = oracle.elastic_solver(order)
    errs = []
    for n in (3, 6):
        g = oracle.Grid(s, n)
        g.set_solution(_p_wave(oracle, order, n, 0.0, lam, mu, rho))
        dts, _ = g.step(4 * n)
        t = g.time()
        sol = g.solution().reshape(-1, 12)
        ex = _p_wave(oracle, order, n, t, lam, mu, rho).reshape(-1, 12)
        assert np.array_equal(sol[:, 9:], ex[:, 9:])  # parameters are carried unchanged
        errs.append(np.abs(sol[:, :9] - ex[:, :9]).max())
    assert errs[0] < 2e-2 * 0.1 * rho * cp
    assert errs[0] / errs[1] > 2 ** 3.5, errs


def test_workload_material_ranges(oracle):
    X = node_coords(3, 6)
    lam, mu, rho = elastic_material(X)
    cs = np.sqrt(mu / rho)
    cp = np.sqrt((lam + 2 * mu) / rho)
    assert rho.min() >= 2.0 and rho.max() <= 3.0
    assert cp.min() >= 4.0 and cp.max() <= 6.0
    assert np.allclose(cs, cp / np.sqrt(3))
    q = elastic_cells(3, 4).reshape(-1, 12)
    assert np.abs(q[:, :6]).max() == 0.0 and q[:, 6:9].max() <= 1.0
