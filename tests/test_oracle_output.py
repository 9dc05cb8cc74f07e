"""The output path: the Cartesian VTK plotter's interpolation in the oracle
(ADERDG2CartesianVTK.cpp:288-336) and the legacy VTK writer.

* equidistantGridProjector1d (the reference's printed literals) against its
  closed form L_n(m / N), the Gauss-Legendre Lagrange basis at the
  equidistant points;
* the interpolation is exact on polynomials of degree <= N per direction:
  the vertex values equal the polynomial at offset + i h / N;
* the writer's file, read back: vertex coordinates, hexahedra spanning one
  sub-cell each with positive volume, and the values.
"""
import numpy as np
import pytest

from exahype_amd import vtk


def _nodes(oracle, order):
    return np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])


def _lagrange(x, nodes, n):
    v = 1.0
    for k, xk in enumerate(nodes):
        if k != n:
            v *= (x - xk) / (nodes[n] - xk)
    return v


@pytest.mark.parametrize("order", range(1, 10))
def test_equidistant_projector_closed_form(oracle, order):
    nodes = _nodes(oracle, order)
    err = max(abs(oracle.equi_proj(order, n, m) - _lagrange(m / order, nodes, n))
              for n in range(order + 1) for m in range(order + 1))
    assert err <= 5e-12


@pytest.mark.parametrize("order", [1, 2, 3, 5, 6])
def test_plot_vertices_exact_on_polynomials(oracle, order):
    rng = np.random.default_rng(order)
    B, D = order + 1, 6
    s = oracle.make_solver(order, nvar=5, npar=1)
    nodes = _nodes(oracle, order)
    coef = rng.standard_normal((D, B, B, B))

    def p(x, y, z):
        xs, ys, zs = x ** np.arange(B), y ** np.arange(B), z ** np.arange(B)
        return np.einsum("vabc,a,b,c->v", coef, xs, ys, zs)
    u = np.array([p(nodes[i0], nodes[i1], nodes[i2]) for i2 in range(B) for i1 in range(B) for i0 in range(B)])
    out = oracle.plot_vertices(s, u.ravel(), 1).reshape(B ** 3, D)
    exact = np.array([p(i0 / order, i1 / order, i2 / order) for i2 in range(B) for i1 in range(B) for i0 in range(B)])
    assert np.abs(out - exact).max() <= 1e-10 * np.abs(exact).max()


def test_vtk_writer_round_trip(tmp_path, oracle):
    order, n, lo, hi = 2, (2, 1, 3), (0.0, -1.0, 0.5), (1.0, 0.0, 2.0)
    B, D = order + 1, 5
    ncells = 6
    q = np.random.default_rng(1).standard_normal(ncells * B ** 3 * D)
    path = vtk.write_cartesian_vertices(str(tmp_path / "a.vtk"), q, order, n, lo, hi, time=0.25, binary=False)
    pts, cells, qq = vtk.read_ascii(path)
    assert pts.shape == (ncells * B ** 3, 3) and cells.shape == (ncells * order ** 3, 8)
    assert np.array_equal(qq, q.reshape(-1, D))
    h = (np.array(hi) - np.array(lo)) / np.array(n) / order
    for hexa in cells:
        c = pts[hexa]
        assert np.allclose(c[6] - c[0], h)          # opposite corners one sub-cell apart
        assert np.allclose(c[1] - c[0], [h[0], 0, 0]) and np.allclose(c[3] - c[0], [0, h[1], 0])
        assert np.allclose(c[4] - c[0], [0, 0, h[2]])
    # cell 1 (i = 1, j = 0, k = 0) starts at lo + (1, 0, 0) * cell size
    assert np.allclose(pts[B ** 3], np.array(lo) + np.array([0.5, 0.0, 0.0]))
    binp = vtk.write_cartesian_vertices(str(tmp_path / "b.vtk"), q, order, n, lo, hi, binary=True)
    assert open(binp, "rb").read().startswith(b"# vtk DataFile Version 3.0")


def test_legendre_vertices_writer(tmp_path):
    """The Legendre vertex plotter (ADERDG2LegendreVTK.cpp:278-385): the
    stored nodal values themselves at offset + x_i h, x_i the Gauss-Legendre
    nodes (numpy's rule, independent of the library's table), N^3 hexahedra
    per cell between neighbouring nodes."""
    from refutil import gauss_legendre
    order, n, lo, hi = 3, (2, 3, 1), (0.0, -1.0, 0.5), (1.0, 0.5, 2.0)
    B, D = order + 1, 5
    ncells = 6
    luh = np.random.default_rng(2).standard_normal(ncells * B ** 3 * D)
    path = vtk.write_legendre_vertices(str(tmp_path / "l.vtk"), luh, order, n, lo, hi, time=0.5, binary=False)
    pts, cells, q = vtk.read_ascii(path)
    assert np.array_equal(q, luh.reshape(-1, D))
    x, _ = gauss_legendre(B)
    h = (np.array(hi) - np.array(lo)) / np.array(n)
    for c in (0, 1, 5):
        ci = np.array([c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1])])
        for node in (0, 1, B, B * B + 2, B ** 3 - 1):
            idx = np.array([node % B, (node // B) % B, node // (B * B)])
            assert np.allclose(pts[c * B ** 3 + node], np.array(lo) + (ci + x[idx]) * h, rtol=0, atol=1e-14)
    assert cells.shape == (ncells * order ** 3, 8)
    for hexa in cells[:order ** 3]:
        c = pts[hexa]
        assert (c[6] > c[0]).all()                     # opposite corners
        assert np.allclose(c[1, 1:], c[0, 1:]) and c[1, 0] > c[0, 0]
        assert np.allclose(c[4, :2], c[0, :2]) and c[4, 2] > c[0, 2]
