"""Legacy VTK output of ADER-DG grids: the Cartesian and Legendre vertex
plotters.

Mirrors the reference's ``vtk::Cartesian::vertices`` plotters
(``exahype/plotters/VTK/ADERDG2CartesianVTK.cpp:288-336``, ``:398-420``):
every cell becomes a patch of N^3 hexahedra on its (N+1)^3 equidistant
vertices ``offset + i h / N``, carrying the interpolated values as vertex data
``Q`` (all V+P stored values, the identity post-processing) and the time
stamp.  The interpolation runs on the device (``exa_plot_vertices`` /
``exa_grid_plot_vertices``); this module only writes the file.

The Legendre vertex plotter (``vtk::Legendre::vertices``,
``exahype/plotters/VTK/ADERDG2LegendreVTK.cpp:278-385``) writes the stored
nodal values themselves at the Gauss-Legendre nodes ``offset + x_i h``, with
the same N^3 hexahedra per cell between neighbouring nodes.

Each patch owns its vertices (vertices on cell faces appear once per cell),
as the reference's patch writer does.
"""
import numpy as np

VTK_HEXAHEDRON = 12


def patch_geometry(order, n, lo, hi, nodes=None):
    """Points [ncells * (N+1)^3, 3] and hexahedra [ncells * N^3, 8] of a uniform
    n = (nx, ny, nz) grid on [lo, hi], cells in grid order (x fastest) and each
    cell's vertices (i2, i1, i0) with x fastest.  The vertices sit at the
    reference-cell positions i/N, or at nodes[i] when given."""
    B = order + 1
    n = np.asarray(n)
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    h = (hi - lo) / n
    c = np.arange(int(np.prod(n)))
    cell_lo = lo + h * np.stack([c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1])], -1)
    i = np.arange(B ** 3)
    idx = np.stack([i % B, (i // B) % B, i // (B * B)], -1)
    frac = idx / order if nodes is None else np.asarray(nodes, dtype=np.float64)[idx]
    points = (cell_lo[:, None, :] + frac[None, :, :] * h[None, None, :]).reshape(-1, 3)
    # hexahedra of one patch: sub-cell (a, b, c) -> its 8 corners in VTK order
    s = np.arange(order)
    a, b, cc = np.meshgrid(s, s, s, indexing="ij")
    a, b, cc = cc.ravel(), b.ravel(), a.ravel()  # x fastest
    corner = lambda x, y, z: (z * B + y) * B + x
    hexa = np.stack([corner(a, b, cc), corner(a + 1, b, cc), corner(a + 1, b + 1, cc), corner(a, b + 1, cc),
                     corner(a, b, cc + 1), corner(a + 1, b, cc + 1), corner(a + 1, b + 1, cc + 1),
                     corner(a, b + 1, cc + 1)], -1)
    cells = (hexa[None, :, :] + (c * B ** 3)[:, None, None]).reshape(-1, 8)
    return points, cells


def write_cartesian_vertices(path, vertices, order, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), time=0.0,
                             binary=True, title="exahype_amd ADER-DG vertices"):
    """Write a legacy VTK unstructured grid.  vertices: [ncells * (N+1)^3 * D]
    as exa_plot_vertices returns them."""
    if order < 1:
        raise ValueError("the Cartesian vertex plotter needs order >= 1")
    return _write(path, vertices, order, n, lo, hi, time, binary, title, None)


def write_legendre_vertices(path, luh, order, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), time=0.0,
                            binary=True, title="exahype_amd ADER-DG Legendre vertices", nodes=None):
    """Write the stored solution luh [ncells * (N+1)^3 * D] (as exa_grid
    downloads it) at the Gauss-Legendre nodes.  nodes defaults to the
    library's quadrature nodes (exa_quadrature_node)."""
    if order < 1:
        raise ValueError("the Legendre vertex plotter needs order >= 1")
    if nodes is None:
        from . import capi
        nodes = [capi.lib().exa_quadrature_node(order, i) for i in range(order + 1)]
    return _write(path, luh, order, n, lo, hi, time, binary, title, nodes)


def _write(path, vertices, order, n, lo, hi, time, binary, title, nodes):
    points, cells = patch_geometry(order, n, lo, hi, nodes)
    npts = len(points)
    q = np.asarray(vertices, dtype=np.float64).reshape(npts, -1)
    D = q.shape[1]
    head = (f"# vtk DataFile Version 3.0\n{title}\n{'BINARY' if binary else 'ASCII'}\n"
            f"DATASET UNSTRUCTURED_GRID\n")
    with open(path, "wb") as f:
        f.write(head.encode())

        def block(header, arr, fmt):
            f.write(header.encode())
            if binary:
                f.write(np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder(">")).tobytes())
                f.write(b"\n")
            else:
                np.savetxt(f, arr.reshape(len(arr), -1), fmt=fmt)

        block(f"POINTS {npts} double\n", points, "%.17g")
        conn = np.concatenate([np.full((len(cells), 1), 8, dtype=np.int32), cells.astype(np.int32)], 1)
        block(f"CELLS {len(cells)} {conn.size}\n", conn, "%d")
        block(f"CELL_TYPES {len(cells)}\n", np.full((len(cells), 1), VTK_HEXAHEDRON, dtype=np.int32), "%d")
        f.write(f"POINT_DATA {npts}\nFIELD FieldData 2\n".encode())
        block(f"Q {D} {npts} double\n", q, "%.17g")
        block(f"time 1 {npts} double\n", np.full((npts, 1), float(time)), "%.17g")
    return path


def read_ascii(path):
    """Minimal reader of files written with binary=False (tests): points,
    cells (without the count column) and the Q array."""
    tok = open(path).read().split()
    i = tok.index("POINTS")
    npts = int(tok[i + 1])
    pts = np.array(tok[i + 3:i + 3 + 3 * npts], dtype=np.float64).reshape(npts, 3)
    i = tok.index("CELLS")
    ncell, size = int(tok[i + 1]), int(tok[i + 2])
    cells = np.array(tok[i + 3:i + 3 + size], dtype=np.int64).reshape(ncell, 9)[:, 1:]
    i = tok.index("Q")
    D = int(tok[i + 1])
    q = np.array(tok[i + 4:i + 4 + D * npts], dtype=np.float64).reshape(npts, D)
    return pts, cells, q
