"""MI355X-native ADER-DG time-stepping core (HIP kernels for gfx950 behind a C-ABI).

The host side mirrors the reference's kernel interface
(kernels::aderdg::generic::c, ExaHyPE/kernels/aderdg/generic/Kernels.h:85-213);
see exahype_amd.capi.  Build the library with ``python -m exahype_amd.build``.
"""
from . import capi  # noqa: F401
from .capi import Grid, Solver, euler_solver, ExaError  # noqa: F401

__all__ = ["capi", "Grid", "Solver", "euler_solver", "ExaError"]
