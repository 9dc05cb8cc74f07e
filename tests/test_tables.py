"""The DG tables used by both the oracle and the HIP kernels.

The tables are the reference's printed literals (ExaHyPE/kernels/DGMatrices.cpp,
GaussLegendreQuadrature.cpp).  They are pinned here against the closed forms
they approximate, computed independently from numpy's Gauss-Legendre rule:
  nodes/weights = Gauss-Legendre on [0,1];  FLCoeff_j = phi_j(0), FRCoeff_j = phi_j(1)
  Kxi[l][n] = w_n phi_l'(x_n);  dudx[l][n] = phi_n'(x_l);  iK1 = (FR FR^T - Kxi)^-1
A transposed operand or a wrong node ordering fails by O(1).
"""
import os
import re

import numpy as np
import pytest

from refutil import gauss_legendre, lagrange

ORDERS = range(0, 10)


def table(O, which, order, B, two_d=False):
    if two_d:
        return np.array([[O.table(which, order, i, j) for j in range(B)] for i in range(B)])
    return np.array([O.table(which, order, i) for i in range(B)])


@pytest.mark.parametrize("order", ORDERS)
def test_tables_match_closed_forms(oracle, order):
    O = oracle
    B = order + 1
    x, w = gauss_legendre(B)
    phi, dphi = lagrange(x)
    FL = np.array([phi(j, 0.0) for j in range(B)])
    FR = np.array([phi(j, 1.0) for j in range(B)])
    Kxi = np.array([[w[n] * dphi(l, x[n]) for n in range(B)] for l in range(B)])
    dudx = np.array([[dphi(n, x[l]) for n in range(B)] for l in range(B)])
    iK1 = np.linalg.inv(np.outer(FR, FR) - Kxi)
    assert np.abs(table(O, O.T_NODES, order, B) - x).max() < 1e-15
    assert np.abs(table(O, O.T_WEIGHTS, order, B) - w).max() < 1e-15
    assert np.abs(table(O, O.T_FLCOEFF, order, B) - FL).max() < 1e-13
    assert np.abs(table(O, O.T_FRCOEFF, order, B) - FR).max() < 1e-13
    # printed with 13 significant digits in the reference
    assert np.abs(table(O, O.T_KXI, order, B, True) - Kxi).max() < 1e-12
    assert np.abs(table(O, O.T_DUDX, order, B, True) - dudx).max() < 1e-11
    assert np.abs(table(O, O.T_IK1, order, B, True) - iK1).max() < 1e-12
    if order > 0:  # the transposed operator is far off: the check pins orientation
        assert np.abs(table(O, O.T_KXI, order, B, True) - Kxi.T).max() > 1e-3


@pytest.mark.parametrize("order", range(1, 10))
def test_lobatto_tables(oracle, order):
    """Gauss-Lobatto nodes are stored in descending order (GaussLobattoQuadrature.cpp)."""
    O = oracle
    B = order + 1
    nodes = table(O, O.T_LOB_NODES, order, B)
    weights = table(O, O.T_LOB_WEIGHTS, order, B)
    assert nodes[0] == 1.0 and nodes[-1] == 0.0
    assert np.all(np.diff(nodes) < 0)
    # Lobatto quadrature with B points integrates x^k exactly up to k = 2B-3
    for k in range(2 * B - 2):
        assert abs(np.dot(weights, nodes ** k) - 1.0 / (k + 1)) < 1e-14


def test_tables_are_the_reference_literals():
    """When the reference sources are present, the generated table file holds
    exactly the doubles the reference's literals denote."""
    ref = "/root/reference/ExaHyPE/kernels/DGMatrices.cpp"
    if not os.path.exists(ref):
        pytest.skip("reference sources not present on this machine")
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = open(os.path.join(here, "include", "exahype_amd", "device", "dg_tables.inc")).read()

    def block(name):
        m = re.search(r"EXA_TABLE double " + name + r"\[[^=]*=\s*\{(.*?)\};", inc, re.S)
        return np.array([float(v) for v in re.findall(r"[-+0-9.eE]+", m.group(1))])

    kxi = block("exa_kxi").reshape(10, 10, 10)
    pat = re.compile(r"^\s*Kxi\[(\d+)\]\[(\d+)\]\[(\d+)\]\s*=\s*([-+0-9.eE]+);")
    n = 0
    for line in open(ref):
        m = pat.match(line)
        if m:
            o, i, j = int(m.group(1)), int(m.group(2)), int(m.group(3))
            assert kxi[o, i, j] == float(m.group(4))
            n += 1
    assert n == 385
