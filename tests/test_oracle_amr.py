"""The AMR part of the oracle (face and volume unknowns prolongation and
restriction, amrRoutines.cpph), pinned before it is trusted:

* the fine-grid projector table (the reference's printed literals) against
  its closed form: coarse Lagrange basis function n (on the Gauss-Legendre
  nodes of [0,1]) at fine node m of subinterval j, L_n((x_m + j) / 3);
* prolongation is exact on polynomials of degree <= N per direction, over
  one and two levels and at every subinterval position (the fine nodal values
  equal the polynomial at the fine nodes);
* restriction summed over all subintervals recovers the coarse nodal values
  of such a polynomial (each subinterval's Gauss quadrature is exact for the
  degree-2N products), so restriction after prolongation is the identity;
* the reference's own tests (GenericEulerKernelTest3D.cpp
  testFaceUnknownsProjection, testVolumeUnknownsProjection): constants are
  reproduced by prolongation and by complete restriction;
* restriction accumulates into the coarse arrays (added, not overwritten).
A transposed projector, a wrong weight, a dropped 1/9 or 1/27 or a wrong
digit order each fail one of these.
"""
import itertools

import numpy as np
import pytest


def _nodes(oracle, order):
    return np.array([oracle.table(oracle.T_NODES, order, i) for i in range(order + 1)])


def _lagrange(x, nodes, n):
    v = 1.0
    for k, xk in enumerate(nodes):
        if k != n:
            v *= (x - xk) / (nodes[n] - xk)
    return v


def _solver(oracle, order, nvar=2, npar=1):
    return oracle.make_solver(order, nvar=nvar, npar=npar)


def _poly(rng, order, nvar, ndim):
    """Random polynomials of degree <= order per direction, one per variable."""
    coef = rng.standard_normal((nvar,) + (order + 1,) * ndim)

    def f(*xs):
        out = np.zeros(nvar)
        for idx in itertools.product(range(order + 1), repeat=ndim):
            mono = 1.0
            for d, e in enumerate(idx):
                mono *= xs[d] ** e
            out += coef[(slice(None),) + idx] * mono
        return out
    return f


def _volume(f, nodes, lo, h, nvar):
    """Nodal values (z, y, x, var) on the cell [lo, lo + h]^3 (lo per direction x, y, z)."""
    B = len(nodes)
    u = np.zeros((B, B, B, nvar))
    for k, j, i in itertools.product(range(B), repeat=3):
        u[k, j, i] = f(lo[0] + h * nodes[i], lo[1] + h * nodes[j], lo[2] + h * nodes[k])
    return u.ravel()


def _face(f, nodes, lo, h, nvar):
    """Nodal values (b, a, var) on a face; subface index (a-like, b-like) = (index 0, index 1)."""
    B = len(nodes)
    u = np.zeros((B, B, nvar))
    for b, a in itertools.product(range(B), repeat=2):
        u[b, a] = f(lo[0] + h * nodes[a], lo[1] + h * nodes[b])
    return u.ravel()


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_fine_grid_projector_closed_form(oracle, order):
    nodes = _nodes(oracle, order)
    err = 0.0
    for j in range(3):
        for n in range(order + 1):
            for m in range(order + 1):
                exact = _lagrange((nodes[m] + j) / 3.0, nodes, n)
                err = max(err, abs(oracle.fine_proj(order, j, n, m) - exact))
    assert err <= 5e-12


@pytest.mark.parametrize("order,level_delta", [(2, 1), (3, 1), (5, 1), (3, 2), (4, 2)])
def test_volume_prolongation_exact_on_polynomials(oracle, order, level_delta):
    rng = np.random.default_rng(order * 10 + level_delta)
    s = _solver(oracle, order)
    D = s.nvar + s.npar
    nodes = _nodes(oracle, order)
    f = _poly(rng, order, D, 3)
    uc = _volume(f, nodes, (0.0, 0.0, 0.0), 1.0, D)
    n = 3 ** level_delta
    h = 1.0 / n
    for sub in [(0, 0, 0), (n - 1, 1, 0), (1, n - 1, n // 2), (n - 1, n - 1, n - 1)]:
        uf = oracle.volume_unknowns_prolongation(s, uc, 2, 2 + level_delta, sub)
        exact = _volume(f, nodes, tuple(h * c for c in sub), h, D)
        assert np.abs(uf - exact).max() <= 1e-10 * np.abs(exact).max(), sub


@pytest.mark.parametrize("order,level_delta", [(2, 1), (3, 1), (4, 1), (3, 2)])
def test_volume_restriction_recovers_coarse_polynomial(oracle, order, level_delta):
    rng = np.random.default_rng(100 + order)
    s = _solver(oracle, order)
    D = s.nvar + s.npar
    nodes = _nodes(oracle, order)
    f = _poly(rng, order, D, 3)
    exact = _volume(f, nodes, (0.0, 0.0, 0.0), 1.0, D)
    n = 3 ** level_delta
    h = 1.0 / n
    start = rng.standard_normal(exact.size)
    uc = start.copy()
    for sub in itertools.product(range(n), repeat=3):
        uf = _volume(f, nodes, tuple(h * c for c in sub), h, D)
        uc = oracle.volume_unknowns_restriction(s, uc, uf, 0, level_delta, sub)
    # accumulated onto the coarse array's previous content
    assert np.abs((uc - start) - exact).max() <= 1e-10 * np.abs(exact).max()


def test_volume_restriction_after_prolongation_is_identity(oracle):
    rng = np.random.default_rng(5)
    s = _solver(oracle, 4, nvar=3, npar=0)
    uc = rng.standard_normal(125 * 3)
    acc = np.zeros_like(uc)
    for sub in itertools.product(range(3), repeat=3):
        uf = oracle.volume_unknowns_prolongation(s, uc, 1, 2, sub)
        acc = oracle.volume_unknowns_restriction(s, acc, uf, 1, 2, sub)
    assert np.abs(acc - uc).max() <= 1e-11 * np.abs(uc).max()


@pytest.mark.parametrize("order,level_delta", [(3, 1), (5, 1), (3, 2)])
def test_face_prolongation_and_restriction_on_polynomials(oracle, order, level_delta):
    rng = np.random.default_rng(7 + order)
    s = _solver(oracle, order)
    D, V = s.nvar + s.npar, s.nvar
    nodes = _nodes(oracle, order)
    fq, ff = _poly(rng, order, D, 2), _poly(rng, order, V, 2)
    qc, fc = _face(fq, nodes, (0.0, 0.0), 1.0, D), _face(ff, nodes, (0.0, 0.0), 1.0, V)
    n = 3 ** level_delta
    h = 1.0 / n
    acc_q, acc_f = np.zeros_like(qc), np.zeros_like(fc)
    for sub in itertools.product(range(n), repeat=2):
        qf, f_f = oracle.face_unknowns_prolongation(s, qc, fc, 0, level_delta, sub)
        lo = (h * sub[0], h * sub[1])
        eq, ef = _face(fq, nodes, lo, h, D), _face(ff, nodes, lo, h, V)
        assert np.abs(qf - eq).max() <= 1e-10 * np.abs(eq).max()
        assert np.abs(f_f - ef).max() <= 1e-10 * np.abs(ef).max()
        acc_q, acc_f = oracle.face_unknowns_restriction(s, acc_q, acc_f, eq, ef, 0, level_delta, sub)
    assert np.abs(acc_q - qc).max() <= 1e-10 * np.abs(qc).max()
    assert np.abs(acc_f - fc).max() <= 1e-10 * np.abs(fc).max()


def test_reference_constant_round_trips(oracle):
    """GenericEulerKernelTest3D.cpp testFaceUnknownsProjection /
    testVolumeUnknownsProjection: order 3, one variable and one parameter,
    constant 1 data; prolongation reproduces it, complete restriction too."""
    s = _solver(oracle, 3, nvar=1, npar=1)
    for level_coarse in range(3):
        uc = np.ones(64 * 2)
        out = np.zeros(64 * 2)
        for sub in itertools.product(range(3), repeat=3):
            uf = oracle.volume_unknowns_prolongation(s, uc, level_coarse, level_coarse + 1, sub)
            assert np.allclose(uf, 1.0, rtol=0, atol=1e-9)
            out = oracle.volume_unknowns_restriction(s, out, np.ones(64 * 2), level_coarse, level_coarse + 1, sub)
        assert np.allclose(out, 1.0, rtol=0, atol=1e-12)
        qc, fc = np.ones(16 * 2), np.ones(16)
        aq, af = np.zeros(16 * 2), np.zeros(16)
        for sub in itertools.product(range(3), repeat=2):
            qf, ff = oracle.face_unknowns_prolongation(s, qc, fc, level_coarse, level_coarse + 1, sub)
            assert np.allclose(qf, 1.0, rtol=0, atol=1e-12) and np.allclose(ff, 1.0, rtol=0, atol=1e-12)
            aq, af = oracle.face_unknowns_restriction(s, aq, af, qf, ff, level_coarse, level_coarse + 1, sub)
        assert np.allclose(aq, 1.0, rtol=0, atol=1e-12) and np.allclose(af, 1.0, rtol=0, atol=1e-12)
