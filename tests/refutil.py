"""Shared helpers for the tests: golden vectors, closed-form DG quantities,
exact Euler entropy-wave solutions, comparison metrics."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GAMMA = 1.4


def golden(name="generic_euler_3d.json"):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def gauss_legendre(B):
    """Nodes and weights on [0,1] from numpy (independent of the reference tables)."""
    xi, wi = np.polynomial.legendre.leggauss(B)
    return (xi + 1.0) / 2.0, wi / 2.0


def lagrange(nodes):
    nodes = np.asarray(nodes, dtype=float)
    B = len(nodes)

    def phi(j, x):
        r = 1.0
        for k in range(B):
            if k != j:
                r *= (x - nodes[k]) / (nodes[j] - nodes[k])
        return r

    def dphi(j, x):
        s = 0.0
        for m in range(B):
            if m == j:
                continue
            r = 1.0 / (nodes[j] - nodes[m])
            for k in range(B):
                if k != j and k != m:
                    r *= (x - nodes[k]) / (nodes[j] - nodes[k])
            s += r
        return s

    return phi, dphi


def euler_state(rho, u, p):
    """Conserved variables (rho, rho u, E) for arrays rho and constant u, p."""
    rho = np.asarray(rho, dtype=float)
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    return np.stack([rho, rho * u[0], rho * u[1], rho * u[2], E], axis=-1)


def euler_flux(Q, d):
    """Physical flux in direction d of conserved states Q[...,5] (textbook form)."""
    rho = Q[..., 0]
    u = Q[..., 1:4] / rho[..., None]
    p = (GAMMA - 1.0) * (Q[..., 4] - 0.5 * rho * (u ** 2).sum(-1))
    F = Q * u[..., d][..., None]
    F[..., 1 + d] += p
    F[..., 4] += p * u[..., d]
    return F


def node_coords(order, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), nodes=None):
    """Physical coordinates of every (cell, z, y, x) node: array [ncells, B,B,B, 3] (x,y,z).
    Cell c = (k*ny + j)*nx + i; node x = centre + dx (xi - 1/2)."""
    B = order + 1
    if nodes is None:
        nodes, _ = gauss_legendre(B)
    n = (n, n, n) if isinstance(n, int) else n
    dx = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    ci = [i.reshape(-1), j.reshape(-1), k.reshape(-1)]
    out = np.zeros((n[0] * n[1] * n[2], B, B, B, 3))
    for d in range(3):
        centre = lo[d] + (ci[d] + 0.5) * dx[d]
        loc = dx[d] * (np.asarray(nodes) - 0.5)
        shape = [1, 1, 1]
        shape[2 - d] = B  # axis order (z,y,x)
        out[..., d] = centre[:, None, None, None] + loc.reshape(shape)[None]
    return out


def cell_weights(order):
    """w_z w_y w_x at each (z,y,x) node."""
    _, w = gauss_legendre(order + 1)
    return w[:, None, None] * w[None, :, None] * w[None, None, :]


def rel_max(a, b):
    """max |a-b| / max |b| (normalised by the array's scale)."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def rel_per_dof(a, b):
    """max over entries of |a-b| / |b| (for strictly non-zero data such as rho, E)."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b) / np.abs(b)))


def random_euler_cells(rng, ncells, order, npar=0):
    """Random admissible Euler states at every node: rho in [0.5,2], u in [-1,1], p in [0.5,2]."""
    B = order + 1
    n = ncells * B ** 3
    rho = rng.uniform(0.5, 2.0, n)
    u = rng.uniform(-1.0, 1.0, (n, 3))
    p = rng.uniform(0.5, 2.0, n)
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u ** 2).sum(1)
    Q = np.concatenate([rho[:, None], rho[:, None] * u, E[:, None], np.zeros((n, npar))], axis=1)
    return np.ascontiguousarray(Q.reshape(-1))


def smooth_euler_cells(rng, ncells, order, scale=0.05):
    """Smooth per-cell data: a random low-degree polynomial perturbation of a constant state."""
    B = order + 1
    x, _ = gauss_legendre(B)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    out = []
    for _ in range(ncells):
        c = rng.uniform(-1, 1, (5, 4))
        base = np.array([1.0, 0.3, -0.2, 0.1, 2.5])
        prim = [base[v] + scale * (c[v, 0] * X + c[v, 1] * Y + c[v, 2] * Z + c[v, 3] * X * Y) for v in range(5)]
        rho, u, v, w, p = prim
        E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
        out.append(np.stack([rho, rho * u, rho * v, rho * w, E], -1))
    return np.ascontiguousarray(np.array(out).reshape(-1))


def splitmix64_uniform(counter, seed):
    """Uniform [0,1) doubles from a counter-based splitmix64: state = seed +
    (counter+1) * 0x9E3779B97F4A7C15, one mixing round, top 53 bits."""
    with np.errstate(over="ignore"):
        c = np.asarray(counter, dtype=np.uint64)
        z = np.uint64(seed) + (c + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def blast_cells(order, n, seed=7, nodes=None, radius=0.2, p_in=10.0, p_out=0.1, noise=0.01, subcells=False,
                box=None):
    """Spherical blast on the unit cube: rho = 1 + noise*(2u-1) with u from
    splitmix64 keyed by the global node index (layout independent), u = 0,
    p = p_in inside |x - 1/2| < radius, p_out outside.  [ncells*B^3*5].
    subcells=True evaluates it at the 2N+1 FV subcell centres per direction
    instead (noise keyed by the global subcell index, stream seed + 1).
    box=(i0, m): only the m^3 cells i0..i0+m-1 (per direction) of the n^3
    grid, with the full grid's coordinates and noise."""
    if subcells:
        L = 2 * order + 1
        nodes = (2.0 * np.arange(L) + 1.0) / (2.0 * L)
        order, seed = L - 1, seed + 1
    B = order + 1
    n3 = (n, n, n) if isinstance(n, int) else n
    if box is None:
        X = node_coords(order, n3, nodes=nodes)  # [c, z, y, x, 3]
        off, m3 = (0, 0, 0), n3
    else:
        i0, m = box
        off, m3 = (i0, i0, i0), (m, m, m)
        X = node_coords(order, m3, lo=tuple(i0 / n3[d] for d in range(3)),
                        hi=tuple((i0 + m) / n3[d] for d in range(3)), nodes=nodes)
    nc = X.shape[0]
    c = np.arange(nc)
    ci = c % m3[0] + off[0]
    cj = (c // m3[0]) % m3[1] + off[1]
    ck = c // (m3[0] * m3[1]) + off[2]
    z, y, x = np.meshgrid(np.arange(B), np.arange(B), np.arange(B), indexing="ij")
    gx = ci[:, None, None, None] * B + x[None]
    gy = cj[:, None, None, None] * B + y[None]
    gz = ck[:, None, None, None] * B + z[None]
    gid = (gz * (n3[1] * B) + gy) * (n3[0] * B) + gx
    u = splitmix64_uniform(gid.astype(np.uint64), seed)
    rho = 1.0 + noise * (2.0 * u - 1.0)
    r = np.sqrt(((X - 0.5) ** 2).sum(-1))
    p = np.where(r < radius, p_in, p_out)
    Q = np.stack([rho, 0 * rho, 0 * rho, 0 * rho, p / (GAMMA - 1.0)], -1)
    return np.ascontiguousarray(Q.reshape(-1))


def elastic_material(X, seed=3):
    """Seeded piecewise-smooth material at points X[..., 3] (x, y, z): two
    layers split by z0(x, y) = 0.5 + 0.1 sin(2 pi x + a) sin(2 pi y + b); per
    layer rho in [2.05, 2.95], cp in [4.1, 5.9], each modulated smoothly by
    +-0.05 / +-0.1; cs = cp / sqrt(3).  Returns (lambda, mu, rho)."""
    u = splitmix64_uniform(np.arange(6, dtype=np.uint64), seed)
    a, b = 2 * np.pi * u[0], 2 * np.pi * u[1]
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    z0 = 0.5 + 0.1 * np.sin(2 * np.pi * x + a) * np.sin(2 * np.pi * y + b)
    upper = z > z0
    rho = np.where(upper, 2.05 + 0.9 * u[2], 2.05 + 0.9 * u[3]) + 0.05 * np.sin(2 * np.pi * ((x + y) + z))
    cp = np.where(upper, 4.1 + 1.8 * u[4], 4.1 + 1.8 * u[5]) + 0.1 * np.sin(2 * np.pi * ((x - y) + z))
    cs = cp / np.sqrt(3.0)
    mu = rho * cs * cs
    lam = rho * cp * cp - 2.0 * mu
    return lam, mu, rho


def elastic_cells(order, n, seed=3, nodes=None, width=0.05, amp=1.0):
    """Elastic workload state at the GL nodes: zero stresses, a Gaussian
    velocity pulse u = v = w = amp exp(-|x - c|^2 / (2 width^2)) at the cube
    centre, parameters (lambda, mu, rho) of elastic_material.  [ncells*B^3*12]."""
    X = node_coords(order, n, nodes=nodes)
    lam, mu, rho = elastic_material(X, seed)
    r2 = (((X[..., 0] - 0.5) ** 2 + (X[..., 1] - 0.5) ** 2) + (X[..., 2] - 0.5) ** 2)
    g = amp * np.exp(-r2 / (2.0 * width * width))
    z = np.zeros_like(g)
    Q = np.stack([z, z, z, z, z, z, g, g, g, lam, mu, rho], -1)
    return np.ascontiguousarray(Q.reshape(-1))


def elastic_matrices(lam, mu, rho):
    """Coefficient matrices B_x, B_y, B_z of dQ/dt + sum_d B_d dQ/dx_d = 0 for the
    9 elastic variables (sxx, syy, szz, sxy, syz, sxz, u, v, w), written from the
    velocity-stress equations directly."""
    Bx, By, Bz = np.zeros((9, 9)), np.zeros((9, 9)), np.zeros((9, 9))
    ir = 1.0 / rho
    # stresses from velocity gradients
    Bx[0, 6], By[0, 7], Bz[0, 8] = -(lam + 2 * mu), -lam, -lam
    Bx[1, 6], By[1, 7], Bz[1, 8] = -lam, -(lam + 2 * mu), -lam
    Bx[2, 6], By[2, 7], Bz[2, 8] = -lam, -lam, -(lam + 2 * mu)
    Bx[3, 7], By[3, 6] = -mu, -mu          # sxy: dv/dx + du/dy
    By[4, 8], Bz[4, 7] = -mu, -mu          # syz: dw/dy + dv/dz
    Bx[5, 8], Bz[5, 6] = -mu, -mu          # sxz: dw/dx + du/dz
    # velocities from stress divergence
    Bx[6, 0], By[6, 3], Bz[6, 5] = -ir, -ir, -ir
    Bx[7, 3], By[7, 1], Bz[7, 4] = -ir, -ir, -ir
    Bx[8, 5], By[8, 4], Bz[8, 2] = -ir, -ir, -ir
    return Bx, By, Bz


def fv_state(rho, u, v, w, p):
    """Conserved Euler variables from primitive arrays (any shape), stacked last."""
    return np.stack([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def fv_smooth(Z, Y, X):
    """A smooth periodic Euler state on [0,1]^3 with all velocity components."""
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (X + 2 * Y)) * np.cos(2 * np.pi * Z)
    u = 0.3 + 0.1 * np.cos(2 * np.pi * Y)
    v = -0.2 + 0.1 * np.sin(2 * np.pi * Z)
    w = 0.1 * np.cos(2 * np.pi * X)
    p = 1.0 + 0.1 * np.sin(2 * np.pi * (X - Z))
    return fv_state(rho, u, v, w, p)


def fv_grid_state(n, L, fn, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """fn(Z, Y, X) at the subcell centres of an n = (nx, ny, nz) grid of L^3
    patches: returns (global array [Z][Y][X][5], cell-major array
    [cells][L][L][L][5] with cell c = (k ny + j) nx + i)."""
    nx, ny, nz = n
    cs = [lo[d] + (np.arange(n[d] * L) + 0.5) * (hi[d] - lo[d]) / (n[d] * L) for d in range(3)]
    Z, Y, X = np.meshgrid(cs[2], cs[1], cs[0], indexing="ij")
    G = fn(Z, Y, X)
    cells = G.reshape(nz, L, ny, L, nx, L, 5).transpose(0, 2, 4, 1, 3, 5, 6).reshape(nx * ny * nz, L, L, L, 5)
    return G, cells


def fv_cells_to_global(cells, n, L):
    nx, ny, nz = n
    return np.asarray(cells).reshape(nz, ny, nx, L, L, L, 5).transpose(0, 3, 1, 4, 2, 5, 6).reshape(
        nz * L, ny * L, nx * L, 5)
