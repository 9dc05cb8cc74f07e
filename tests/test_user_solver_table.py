"""The user-solver route on the host side: the table EXA_USER_SOLVER exports
from the example shared object (a host function; no GPU call).  The hook
flags must see the PDE's __device__ hooks from host code, and the launch
functions must all be present."""
import ctypes

import pytest

from exahype_amd import capi


@pytest.mark.parametrize("order", [3, 5])
def test_example_table(order):
    fn = getattr(capi.user_example(), f"exa_user_euler_order{order}")
    fn.restype = ctypes.POINTER(capi.UserSolverTable)
    t = fn().contents
    assert t.abi_version == 1
    assert (t.order, t.num_variables, t.num_parameters, t.terms) == (order, 5, 0, capi.TERM_FLUX)
    assert t.cfl == 0.9
    assert t.has_boundary_values == 1
    assert t.has_adjust == 1
    assert all((t.predictor, t.corrector, t.stable_dt, t.adjust))
