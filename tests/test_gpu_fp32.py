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
