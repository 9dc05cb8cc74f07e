"""The Cartesian VTK plotter's interpolation on the device against the oracle
(1e-12 of scale; the device applies the tensor-product projector direction by
direction), for host arrays and for a device-resident grid, and the written
file."""
import numpy as np
import pytest

from exahype_amd import vtk

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
def test_plot_vertices_vs_oracle(exa, oracle, order):
    print(f"case: plot vertices N={order}", flush=True)
    rng = np.random.default_rng(order)
    ncells, B = 7, order + 1
    luh = rng.standard_normal(ncells * B ** 3 * 6)
    dev = exa.plot_vertices(exa.euler_solver(order, npar=1), luh, ncells)
    ref = oracle.plot_vertices(oracle.make_solver(order, nvar=5, npar=1), luh, ncells)
    assert _rel(dev, ref) <= 1e-12


def test_grid_plot_vertices_and_file(exa, oracle, tmp_path):
    print("case: grid plot vertices N=3 5^3", flush=True)
    s = exa.euler_solver(3)
    g = exa.Grid(s, 5)
    g.init_euler()
    g.step(2)
    luh = g.solution()
    out, ms = g.plot_vertices(timed=True)
    assert ms > 0.0
    assert _rel(out, oracle.plot_vertices(oracle.make_solver(3), luh, 125)) <= 1e-12
    path = vtk.write_cartesian_vertices(str(tmp_path / "g.vtk"), out, 3, (5, 5, 5), binary=False)
    pts, cells, q = vtk.read_ascii(path)
    assert q.shape == (125 * 64, 5) and np.array_equal(q, out.reshape(-1, 5))
    g.close()
