"""The FP32-bulk form (predictor, volume integral and Riemann solver in float;
surface integral, solution update and time step in double) against the
plain FP64 form on the same initial data.

Bar (the precision workload's requirement): relative L2 error <= 1e-6 after
20 steps; no more Picard iterations than FP64; conservation to the rounding
of the float volume integral (its node contributions sum to zero per cell only
up to float rounding, ~1e-7 of the flux per step; face fluxes are shared by
both cells and the update is double, so they cancel exactly).
"""
import numpy as np
import pytest

from refutil import cell_weights

pytestmark = pytest.mark.gpu


def run(exa, order, n, steps, precision, seed=2024):
    g = exa.Grid(exa.euler_solver(order, precision=precision), n)
    g.init_euler(nmodes=8, mode_amp=0.01, seed=seed)
    u0 = g.solution()
    g.set_profiling(True)
    g.step(steps)
    sol = g.solution()
    its, cells = g.picard_totals()
    t, dt, nst = g.time()
    g.close()
    return u0, sol, its / cells, t


@pytest.mark.parametrize("order,n", [(5, 9), (3, 12), (7, 5)])
def test_fp32_bulk_error_vs_fp64(exa, order, n):
    print(f"case: fp32-bulk vs fp64, N={order} {n}^3, 20 steps", flush=True)
    u0, s64, it64, t64 = run(exa, order, n, 20, exa.PRECISION_FP64)
    u0b, s32, it32, t32 = run(exa, order, n, 20, exa.PRECISION_FP32_BULK)
    assert np.array_equal(u0, u0b)
    assert abs(t32 - t64) <= 1e-6 * t64
    l2 = np.linalg.norm(s32 - s64) / np.linalg.norm(s64)
    linf = np.abs(s32 - s64).max() / np.abs(s64).max()
    print(f"  L2 rel {l2:.3e}  Linf rel {linf:.3e}  picard fp64 {it64:.2f} fp32 {it32:.2f}", flush=True)
    assert l2 <= 1e-6
    assert linf <= 1e-5
    assert it32 <= it64 + 1e-12
    # conservation of the double solution, to float rounding of the volume terms
    B = order + 1
    w3 = cell_weights(order)[None, ..., None]
    tot0 = (w3 * u0.reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    tot = (w3 * s32.reshape(-1, B, B, B, 5)).sum(axis=(0, 1, 2, 3))
    drift = np.abs(tot - tot0).max() / np.abs(tot0).max()
    print(f"  conservation drift {drift:.3e}", flush=True)
    assert drift < 2e-6


@pytest.mark.parametrize("order,n", [(2, 8), (3, 6), (4, 6), (5, 5), (6, 4), (7, 4)])
def test_matrix_core_predictor_vs_vector_alu(exa, order, n):
    """The FP32-bulk predictor on the matrix cores (EXA_PREDICTOR_MFMA) and on
    the vector ALUs (AUTO) evaluate the same float arithmetic in another
    rounding order: after 6 steps they agree to float rounding (relative L2
    <= 2e-7, Picard means within 0.1), and each is within the precision bar of
    the FP64 form."""
    print(f"case: matrix-core vs vector-ALU predictor, N={order} {n}^3, 6 steps", flush=True)
    sols, its = {}, {}
    for variant in (exa.PREDICTOR_MFMA, exa.PREDICTOR_AUTO):
        g = exa.Grid(exa.euler_solver(order, precision=exa.PRECISION_FP32_BULK), n)
        g.set_predictor_variant(variant)
        g.init_euler(nmodes=8, mode_amp=0.01, seed=11)
        g.set_profiling(True)
        g.step(6)
        sols[variant] = g.solution()
        it, cells = g.picard_totals()
        its[variant] = it / cells
        g.close()
    g = exa.Grid(exa.euler_solver(order), n)
    g.init_euler(nmodes=8, mode_amp=0.01, seed=11)
    g.step(6)
    s64 = g.solution()
    g.close()
    a, v = sols[exa.PREDICTOR_MFMA], sols[exa.PREDICTOR_AUTO]
    d = np.linalg.norm(a - v) / np.linalg.norm(v)
    e_a = np.linalg.norm(a - s64) / np.linalg.norm(s64)
    e_v = np.linalg.norm(v - s64) / np.linalg.norm(s64)
    print(f"  mfma vs valu L2 {d:.3e}; vs fp64: mfma {e_a:.3e} valu {e_v:.3e}; "
          f"picard mfma {its[exa.PREDICTOR_MFMA]:.3f} valu {its[exa.PREDICTOR_AUTO]:.3f}", flush=True)
    assert d <= 2e-7
    assert e_a <= 1e-6 and e_v <= 1e-6
    assert abs(its[exa.PREDICTOR_MFMA] - its[exa.PREDICTOR_AUTO]) <= 0.1


def test_predictor_variant_setter_rejects_unsupported(exa):
    """Variants the solver does not compile are refused (the FP64 form has no
    matrix-core kernel); out-of-range values are invalid."""
    g = exa.Grid(exa.euler_solver(3), 2)
    lib = exa.lib()
    assert lib.exa_grid_set_predictor_variant(g._h, exa.PREDICTOR_MFMA) == exa.EXA_E_UNSUPPORTED
    assert lib.exa_grid_set_predictor_variant(g._h, 7) == exa.EXA_E_INVALID
    assert lib.exa_grid_set_predictor_variant(g._h, exa.PREDICTOR_VALU_LDS144) == exa.EXA_OK
    assert lib.exa_grid_set_predictor_variant(g._h, exa.PREDICTOR_VALU_LDS64) == exa.EXA_OK
    g.close()
    g = exa.Grid(exa.euler_solver(5, precision=exa.PRECISION_FP32_BULK), 2)
    assert lib.exa_grid_set_predictor_variant(g._h, exa.PREDICTOR_MFMA) == exa.EXA_OK
    g.close()
