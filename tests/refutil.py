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
