"""ctypes binding of the C-ABI in include/exahype_amd.h (libexahype_amd.so).

Function names follow the reference kernels they replace
(kernels::aderdg::generic::c::*, ExaHyPE/kernels/aderdg/generic/Kernels.h:85-213).
There is no CPU fallback: if the HIP library is missing or a call fails,
an exception is raised.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexahype_amd.so")

EXA_OK = 0
EXA_E_HIP, EXA_E_UNSUPPORTED, EXA_E_INVALID, EXA_E_NOMEM, EXA_E_NONFINITE = 1, 2, 3, 4, 5
PDE_EULER = 0
PDE_ELASTIC = 1
TERM_FLUX, TERM_SOURCE, TERM_NCP = 1, 2, 4
STAGE_PREDICTOR, STAGE_RIEMANN, STAGE_UPDATE, STAGE_DT = range(4)
STAGE_NAMES = ("predictor", "riemann", "update", "dt")
BC_PERIODIC = 0
BC_TRANSMISSIVE = 1
BC_USER = 2
PDE_USER = 100
FV_GODUNOV, FV_MUSCLHANCOCK = 0, 1
PRECISION_FP64, PRECISION_FP32_BULK = 0, 1
PREDICTOR_AUTO, PREDICTOR_VALU_LDS64, PREDICTOR_VALU_LDS144, PREDICTOR_MFMA = 0, 1, 2, 3

# Every symbol declared in include/exahype_amd.h.
EXPORTED = (
    "exa_status_string", "exa_init", "exa_solver_supported",
    "exa_space_time_predictor_nonlinear", "exa_volume_integral_nonlinear",
    "exa_surface_integral_nonlinear", "exa_riemann_solver_nonlinear", "exa_solution_update",
    "exa_stable_time_step_size", "exa_batch_predictor_volume",
    "exa_grid_create", "exa_grid_destroy", "exa_grid_num_cells", "exa_grid_upload_solution",
    "exa_grid_download_solution", "exa_grid_init_euler", "exa_grid_step", "exa_grid_synchronize",
    "exa_grid_time", "exa_grid_dt_history", "exa_grid_picard_iterations",
    "exa_grid_set_profiling", "exa_grid_stage_times", "exa_grid_download_cells", "exa_grid_picard_totals",
    "exa_grid_init_elastic", "exa_project_on_fv_limiter_space", "exa_project_on_dg_space", "exa_find_cell_local_min_and_max",
    "exa_discrete_maximum_principle_and_min_and_max_search", "exa_musclhancock_solution_update",
    "exa_lgrid_create", "exa_lgrid_destroy", "exa_lgrid_num_cells", "exa_lgrid_upload", "exa_lgrid_init_blast",
    "exa_lgrid_step", "exa_lgrid_synchronize", "exa_lgrid_time", "exa_lgrid_download", "exa_lgrid_status",
    "exa_lgrid_limiter_solution", "exa_lgrid_min_max", "exa_lgrid_set_profiling", "exa_lgrid_stage_times",
    "exa_lgrid_picard_totals", "exa_lgrid_set_fused_time_stepping", "exa_lgrid_predictor_reruns",
    "exa_grid_set_predictor_variant", "exa_grid_set_fused_corrector", "exa_grid_create_user",
    "exa_grid_set_fused_time_stepping", "exa_grid_predictor_reruns",
    "exa_grid_set_adjust_solution", "exa_grid_adjust_solution", "exa_patchset_create", "exa_patchset_run", "exa_patchset_stats", "exa_patchset_predictor_variant",
    "exa_patchset_solution", "exa_patchset_time", "exa_patchset_destroy", "exa_patchset_group_variant",
    "exa_patchset_set_fused_corrector", "exa_patchset_set_profiling", "exa_patchset_group_times",
    "exa_godunov_solution_update", "exa_fvgrid_create", "exa_fvgrid_destroy", "exa_fvgrid_num_cells",
    "exa_fvgrid_upload", "exa_fvgrid_download", "exa_fvgrid_init_euler", "exa_fvgrid_step",
    "exa_fvgrid_synchronize", "exa_fvgrid_time", "exa_fvgrid_set_profiling", "exa_fvgrid_stage_times",
    "exa_space_time_predictor_linear", "exa_volume_integral_linear", "exa_surface_integral_linear",
    "exa_riemann_solver_linear",
    "exa_face_unknowns_prolongation", "exa_face_unknowns_restriction", "exa_volume_unknowns_prolongation",
    "exa_volume_unknowns_restriction", "exa_device_alloc", "exa_device_free", "exa_copy_to_device",
    "exa_copy_to_host", "exa_volume_unknowns_prolongation_batch", "exa_volume_unknowns_restriction_batch",
    "exa_volume_unknowns_restriction_batch_noalloc",
    "exa_face_unknowns_prolongation_batch", "exa_face_unknowns_restriction_batch", "exa_point_source_projection",
    "exa_quadrature_node", "exa_plot_vertices", "exa_grid_plot_vertices",
)


class ExaError(RuntimeError):
    pass


class Solver(ctypes.Structure):
    """exa_solver: the reference SolverType's constants."""
    _fields_ = [("pde", ctypes.c_int), ("order", ctypes.c_int), ("num_variables", ctypes.c_int),
                ("num_parameters", ctypes.c_int), ("terms", ctypes.c_int),
                ("no_time_averaging", ctypes.c_int), ("cfl", ctypes.c_double), ("precision", ctypes.c_int),
                ("linear", ctypes.c_int)]

    @property
    def B(self):
        return self.order + 1

    @property
    def D(self):
        return self.num_variables + self.num_parameters


class UserSolverTable(ctypes.Structure):
    """exa_user_solver: a user PDE compiled into the kernel templates
    (include/exahype_amd/user_solver.hpp); the launch functions are opaque here."""
    _fields_ = [("abi_version", ctypes.c_int), ("order", ctypes.c_int), ("num_variables", ctypes.c_int),
                ("num_parameters", ctypes.c_int), ("terms", ctypes.c_int), ("cfl", ctypes.c_double),
                ("has_boundary_values", ctypes.c_int), ("has_adjust", ctypes.c_int),
                ("predictor", ctypes.c_void_p), ("corrector", ctypes.c_void_p), ("stable_dt", ctypes.c_void_p),
                ("adjust", ctypes.c_void_p)]


class EulerIC(ctypes.Structure):
    _fields_ = [("rho0", ctypes.c_double), ("amp", ctypes.c_double), ("k", ctypes.c_double * 3),
                ("vel", ctypes.c_double * 3), ("p", ctypes.c_double), ("nmodes", ctypes.c_int),
                ("mode_amp", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


class PatchSpec(ctypes.Structure):
    """exa_patch_spec: one periodic unit-cube patch: Euler with an entropy
    wave, or elasticity with a seeded material and a velocity pulse."""
    _fields_ = [("n", ctypes.c_int), ("order", ctypes.c_int), ("k", ctypes.c_double * 3), ("amp", ctypes.c_double),
                ("pde", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("width", ctypes.c_double)]


def elastic_patch(n, order, seed, width=0.05, amp=1.0):
    """An elastic patch spec for PatchSet (Euler patches are (n, order, k, amp))."""
    return {"pde": PDE_ELASTIC, "n": n, "order": order, "seed": seed, "width": width, "amp": amp}


def _patch_fields(spec):
    """(pde, n, order, k, amp, seed, width) of an Euler tuple or an elastic dict."""
    if isinstance(spec, dict):
        return (spec["pde"], spec["n"], spec["order"], (0.0, 0.0, 0.0), spec.get("amp", 1.0), spec.get("seed", 0),
                spec.get("width", 0.05))
    n, order, k, amp = spec
    return (PDE_EULER, n, order, k, amp, 0, 0.0)


def euler_solver(order, npar=0, terms=TERM_FLUX, cfl=0.9, precision=PRECISION_FP64, linear=False):
    return Solver(PDE_EULER, order, 5, npar, terms, 0, cfl, precision, int(linear))


def elastic_solver(order, cfl=0.9, precision=PRECISION_FP64, linear=False):
    """Linear elasticity: 9 variables + (lambda, mu, rho), non-conservative product only.
    linear=True selects the linear kernels (Cauchy-Kowalevski predictor, linear
    Riemann solver and surface integral) in grids."""
    return Solver(PDE_ELASTIC, order, 9, 3, TERM_NCP, 0, cfl, precision, int(linear))


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib():
    """Load libexahype_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ExaError(f"{LIB_PATH} is missing: build it with `python -m exahype_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    S = ctypes.POINTER(Solver)
    V = ctypes.c_void_p
    L.exa_status_string.argtypes = [ctypes.c_int]
    L.exa_status_string.restype = ctypes.c_char_p
    L.exa_init.argtypes = [ctypes.c_int]
    L.exa_solver_supported.argtypes = [S]
    L.exa_space_time_predictor_nonlinear.argtypes = [S, _D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_double, _I]
    L.exa_volume_integral_nonlinear.argtypes = [S, _D, _D, _D]
    L.exa_surface_integral_nonlinear.argtypes = [S, _D, _D, _D]
    L.exa_riemann_solver_nonlinear.argtypes = [S, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int]
    L.exa_solution_update.argtypes = [S, _D, _D, ctypes.c_double]
    L.exa_space_time_predictor_linear.argtypes = [S, _D, _D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_double, _D]
    L.exa_volume_integral_linear.argtypes = [S, _D, _D, _D]
    L.exa_surface_integral_linear.argtypes = [S, _D, _D, _D]
    L.exa_riemann_solver_linear.argtypes = [S, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int]
    I, Lg, V = ctypes.c_int, ctypes.c_long, ctypes.c_void_p
    F = ctypes.POINTER(ctypes.c_float)
    L.exa_point_source_projection.argtypes = [S, _D, _D, _D, _D]
    L.exa_plot_vertices.argtypes = [S, Lg, _D, _D]
    L.exa_grid_plot_vertices.argtypes = [V, _D, F]
    L.exa_quadrature_node.argtypes = [I, I]
    L.exa_quadrature_node.restype = ctypes.c_double
    L.exa_face_unknowns_prolongation.argtypes = [S, _D, _D, _D, _D, I, I, _I]
    L.exa_face_unknowns_restriction.argtypes = [S, _D, _D, _D, _D, I, I, _I]
    L.exa_volume_unknowns_prolongation.argtypes = [S, _D, _D, I, I, _I]
    L.exa_volume_unknowns_restriction.argtypes = [S, _D, _D, I, I, _I]
    L.exa_device_alloc.argtypes = [ctypes.c_size_t]
    L.exa_device_alloc.restype = V
    L.exa_device_free.argtypes = [V]
    L.exa_copy_to_device.argtypes = [V, V, ctypes.c_size_t]
    L.exa_copy_to_host.argtypes = [V, V, ctypes.c_size_t]
    L.exa_volume_unknowns_prolongation_batch.argtypes = [S, V, Lg, V, Lg, V, Lg, I, F]
    L.exa_volume_unknowns_restriction_batch.argtypes = [S, V, Lg, V, Lg, V, Lg, V, Lg, I, F]
    L.exa_volume_unknowns_restriction_batch_noalloc.argtypes = [S, V, Lg, V, Lg, V, Lg, V, Lg, I]
    L.exa_face_unknowns_prolongation_batch.argtypes = [S, V, V, Lg, V, V, Lg, V, Lg, I, F]
    L.exa_face_unknowns_restriction_batch.argtypes = [S, V, V, Lg, V, V, Lg, V, Lg, V, Lg, I, F]
    L.exa_stable_time_step_size.argtypes = [S, _D, _D, _D]
    L.exa_batch_predictor_volume.argtypes = [S, ctypes.c_long, _D, _D, _D, _D, _D, ctypes.c_double, _I]
    L.exa_grid_create.argtypes = [S, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int,
                                  ctypes.POINTER(V)]
    L.exa_grid_destroy.argtypes = [V]
    L.exa_grid_num_cells.argtypes = [V]
    L.exa_grid_num_cells.restype = ctypes.c_long
    L.exa_grid_upload_solution.argtypes = [V, _D]
    L.exa_grid_download_solution.argtypes = [V, _D]
    L.exa_grid_init_euler.argtypes = [V, ctypes.POINTER(EulerIC)]
    L.exa_grid_init_elastic.argtypes = [V, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double]
    L.exa_grid_step.argtypes = [V, ctypes.c_int]
    L.exa_grid_synchronize.argtypes = [V]
    L.exa_grid_time.argtypes = [V, _D, _D, ctypes.POINTER(ctypes.c_long)]
    L.exa_grid_dt_history.argtypes = [V, _D, ctypes.c_long]
    L.exa_grid_picard_iterations.argtypes = [V, _I]
    L.exa_grid_set_profiling.argtypes = [V, ctypes.c_int]
    L.exa_grid_set_predictor_variant.argtypes = [V, ctypes.c_int]
    L.exa_grid_set_fused_corrector.argtypes = [V, ctypes.c_int]
    L.exa_grid_create_user.argtypes = [ctypes.POINTER(UserSolverTable), I, I, I, I, _D, _D, I, ctypes.POINTER(V)]
    L.exa_grid_set_adjust_solution.argtypes = [V, I]
    L.exa_grid_set_fused_time_stepping.argtypes = [V, I, ctypes.c_double]
    L.exa_grid_predictor_reruns.argtypes = [V, ctypes.POINTER(Lg)]
    L.exa_grid_adjust_solution.argtypes = [V]
    L.exa_grid_stage_times.argtypes = [V, _D, ctypes.POINTER(ctypes.c_long)]
    L.exa_grid_download_cells.argtypes = [V, ctypes.c_long, ctypes.c_long, _D]
    L.exa_grid_picard_totals.argtypes = [V, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]
    i, d_ = ctypes.c_int, ctypes.c_double
    L.exa_project_on_fv_limiter_space.argtypes = [_D, i, i, i, _D]
    L.exa_project_on_dg_space.argtypes = [_D, i, i, i, _D]
    L.exa_find_cell_local_min_and_max.argtypes = [_D, i, i, _D, _D]
    L.exa_discrete_maximum_principle_and_min_and_max_search.argtypes = [_D, i, i, d_, d_, _D, _D, _I]
    L.exa_musclhancock_solution_update.argtypes = [S, _D, _D, _D, d_, i, i, _D]
    L.exa_lgrid_create.argtypes = [S, i, i, i, _D, _D, i, i, i, d_, d_, i, ctypes.POINTER(V)]
    L.exa_lgrid_destroy.argtypes = [V]
    L.exa_lgrid_num_cells.argtypes = [V]
    L.exa_lgrid_num_cells.restype = ctypes.c_long
    L.exa_lgrid_upload.argtypes = [V, _D, _D]
    L.exa_lgrid_init_blast.argtypes = [V, ctypes.c_ulonglong, d_, d_, d_, d_]
    L.exa_lgrid_step.argtypes = [V, i]
    L.exa_lgrid_synchronize.argtypes = [V]
    L.exa_lgrid_time.argtypes = [V, _D, _D, ctypes.POINTER(ctypes.c_long)]
    L.exa_lgrid_download.argtypes = [V, _D]
    L.exa_lgrid_status.argtypes = [V, _I, ctypes.POINTER(ctypes.c_long)]
    L.exa_lgrid_limiter_solution.argtypes = [V, ctypes.c_long, _D]
    L.exa_lgrid_min_max.argtypes = [V, _D, _D]
    L.exa_lgrid_set_profiling.argtypes = [V, i]
    L.exa_lgrid_stage_times.argtypes = [V, _D, ctypes.POINTER(ctypes.c_long)]
    L.exa_lgrid_picard_totals.argtypes = [V, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]
    L.exa_lgrid_set_fused_time_stepping.argtypes = [V, i, d_]
    L.exa_lgrid_predictor_reruns.argtypes = [V, ctypes.POINTER(ctypes.c_long)]
    L.exa_patchset_create.argtypes = [ctypes.POINTER(PatchSpec), i, i, d_, ctypes.c_long, ctypes.POINTER(V)]
    L.exa_patchset_run.argtypes = [V, i, i, i, i]
    L.exa_patchset_stats.argtypes = [V, _D, _D, _I]
    L.exa_patchset_predictor_variant.argtypes = [V, i]
    L.exa_patchset_group_variant.argtypes = [V, i, i]
    L.exa_patchset_set_fused_corrector.argtypes = [V, i]
    L.exa_patchset_set_profiling.argtypes = [V, i]
    L.exa_patchset_group_times.argtypes = [V, i, i, _D, _D, ctypes.POINTER(ctypes.c_long),
                                           ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]
    L.exa_patchset_solution.argtypes = [V, i, _D]
    L.exa_patchset_time.argtypes = [V, i, _D, _D]
    L.exa_patchset_destroy.argtypes = [V]
    L.exa_godunov_solution_update.argtypes = [S, _D, _D, _D, d_, i, i, _D]
    L.exa_fvgrid_create.argtypes = [S, i, i, i, i, i, _D, _D, i, ctypes.POINTER(V)]
    L.exa_fvgrid_destroy.argtypes = [V]
    L.exa_fvgrid_num_cells.argtypes = [V]
    L.exa_fvgrid_num_cells.restype = ctypes.c_long
    L.exa_fvgrid_upload.argtypes = [V, _D]
    L.exa_fvgrid_download.argtypes = [V, _D]
    L.exa_fvgrid_init_euler.argtypes = [V, ctypes.POINTER(EulerIC)]
    L.exa_fvgrid_step.argtypes = [V, i]
    L.exa_fvgrid_synchronize.argtypes = [V]
    L.exa_fvgrid_time.argtypes = [V, _D, _D, ctypes.POINTER(ctypes.c_long)]
    L.exa_fvgrid_set_profiling.argtypes = [V, i]
    L.exa_fvgrid_stage_times.argtypes = [V, _D, ctypes.POINTER(ctypes.c_long)]
    _lib = L
    return L


def check(status, what=""):
    if status != EXA_OK:
        msg = lib().exa_status_string(status).decode()
        raise ExaError(f"{what}: {msg} (status {status})")


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dx3(dx):
    return _f64(np.broadcast_to(np.asarray(dx, dtype=np.float64), (3,)))


def supported(s):
    return bool(lib().exa_solver_supported(ctypes.byref(s)))


# ---------------------------------------------------------------------------
# kernels::aderdg::generic::c single-cell entry points
# ---------------------------------------------------------------------------
def space_time_predictor_nonlinear(s, luh, dx, dt, temps=False):
    """spaceTimePredictorNonlinear: returns dict(lQhbnd, lFhbnd, iters[, lQi, lFi, lQhi, lFhi])."""
    B, D, V = s.B, s.D, s.num_variables
    out = dict(lQhbnd=np.zeros(6 * B * B * D), lFhbnd=np.zeros(6 * B * B * V))
    if temps:
        out.update(lQi=np.zeros(B ** 4 * D), lFi=np.zeros(4 * B ** 4 * V), lQhi=np.zeros(B ** 3 * D),
                   lFhi=np.zeros(4 * B ** 3 * V))
    it = ctypes.c_int(0)
    check(lib().exa_space_time_predictor_nonlinear(
        ctypes.byref(s), _dp(out["lQhbnd"]), _dp(out["lFhbnd"]), _dp(out.get("lQi")), _dp(out.get("lFi")),
        _dp(out.get("lQhi")), _dp(out.get("lFhi")), _dp(_f64(luh)), _dp(_dx3(dx)), float(dt),
        ctypes.byref(it)), "spaceTimePredictorNonlinear")
    out["iters"] = it.value
    return out


def volume_integral_nonlinear(s, lFhi, dx):
    """volumeIntegralNonlinear of lFhi (4 B^3 V), or of the space-time fluxes
    lFi (4 B^4 V) when s.no_time_averaging is set."""
    lduh = np.zeros(s.B ** 3 * s.num_variables)
    check(lib().exa_volume_integral_nonlinear(ctypes.byref(s), _dp(lduh), _dp(_f64(lFhi)), _dp(_dx3(dx))),
          "volumeIntegralNonlinear")
    return lduh


def surface_integral_nonlinear(s, lduh, lFhbnd, dx):
    lduh = np.array(lduh, dtype=np.float64)
    check(lib().exa_surface_integral_nonlinear(ctypes.byref(s), _dp(lduh), _dp(_f64(lFhbnd)), _dp(_dx3(dx))),
          "surfaceIntegralNonlinear")
    return lduh


def riemann_solver_nonlinear(s, FL, FR, QL, QR, dt, direction):
    FL = np.array(FL, dtype=np.float64)
    FR = np.array(FR, dtype=np.float64)
    check(lib().exa_riemann_solver_nonlinear(ctypes.byref(s), _dp(FL), _dp(FR), _dp(_f64(QL)), _dp(_f64(QR)),
                                             float(dt), int(direction)), "riemannSolverNonlinear")
    return FL, FR


def solution_update(s, luh, lduh, dt):
    luh = np.array(luh, dtype=np.float64)
    check(lib().exa_solution_update(ctypes.byref(s), _dp(luh), _dp(_f64(lduh)), float(dt)), "solutionUpdate")
    return luh


def stable_time_step_size(s, luh, dx):
    out = np.zeros(1)
    check(lib().exa_stable_time_step_size(ctypes.byref(s), _dp(_f64(luh)), _dp(_dx3(dx)), _dp(out)),
          "stableTimeStepSize")
    return float(out[0])


def space_time_predictor_linear(s, luh, dx, dt, temps=False, PSi=None):
    """spaceTimePredictorLinear: returns dict(lQhbnd, lFhbnd[, lQi, lFi, gradQ, lQhi, lFhi]).
    PSi: point sources (B+1) B^3 V, or None."""
    B, D, V = s.B, s.D, s.num_variables
    out = dict(lQhbnd=np.zeros(6 * B * B * D), lFhbnd=np.zeros(6 * B * B * V))
    if temps:
        out.update(lQi=np.zeros((B + 1) * B ** 3 * D), lFi=np.zeros(4 * B ** 4 * V), gradQ=np.zeros(3 * B ** 4 * V),
                   lQhi=np.zeros(B ** 3 * D), lFhi=np.zeros(4 * B ** 3 * V))
    check(lib().exa_space_time_predictor_linear(
        ctypes.byref(s), _dp(out["lQhbnd"]), _dp(out["lFhbnd"]), _dp(out.get("lQi")), _dp(out.get("lFi")),
        _dp(out.get("gradQ")), _dp(out.get("lQhi")), _dp(out.get("lFhi")), _dp(_f64(luh)), _dp(_dx3(dx)), float(dt),
        _dp(None if PSi is None else _f64(PSi))), "spaceTimePredictorLinear")
    return out


def volume_integral_linear(s, lFhi, dx=(1.0, 1.0, 1.0)):
    """volumeIntegralLinear of lFhi (4 B^3 V; the source block is not read)."""
    lduh = np.zeros(s.B ** 3 * s.num_variables)
    check(lib().exa_volume_integral_linear(ctypes.byref(s), _dp(lduh), _dp(_f64(lFhi)), _dp(_dx3(dx))),
          "volumeIntegralLinear")
    return lduh


def surface_integral_linear(s, lduh, lFhbnd, dx):
    lduh = np.array(lduh, dtype=np.float64)
    check(lib().exa_surface_integral_linear(ctypes.byref(s), _dp(lduh), _dp(_f64(lFhbnd)), _dp(_dx3(dx))),
          "surfaceIntegralLinear")
    return lduh


def riemann_solver_linear(s, QL, QR, dt, direction):
    """riemannSolverLinear: returns the fluctuations (FL, FR)."""
    n = s.B * s.B * s.num_variables
    FL, FR = np.zeros(n), np.zeros(n)
    check(lib().exa_riemann_solver_linear(ctypes.byref(s), _dp(FL), _dp(FR), _dp(_f64(QL)), _dp(_f64(QR)),
                                          float(dt), int(direction)), "riemannSolverLinear")
    return FL, FR


def plot_vertices(s, luh, ncells):
    """The Cartesian VTK plotter's values at each cell's (N+1)^3 equidistant vertices."""
    out = np.zeros(ncells * s.B ** 3 * s.D)
    check(lib().exa_plot_vertices(ctypes.byref(s), ncells, _dp(_f64(luh)), _dp(out)), "plotVertexData")
    return out


def point_source_projection(s, PSi, x0, center, dx):
    """pointSource's projection: PSi (B+1) B^3 V forces, x0 [(B+1) B^3, 3] source positions."""
    PSi = np.array(PSi, dtype=np.float64).ravel()
    check(lib().exa_point_source_projection(ctypes.byref(s), _dp(PSi), _dp(_f64(np.ravel(x0))), _dp(_dx3(center)),
                                            _dp(_dx3(dx))), "pointSource")
    return PSi


# ---------------------------------------------------------------------------
# AMR transfer operators (amrRoutines.cpph)
# ---------------------------------------------------------------------------
def _sub(v, n):
    return (ctypes.c_int * n)(*[int(x) for x in v])


def face_unknowns_prolongation(s, QC, FC, coarse_level, fine_level, subface):
    """faceUnknownsProlongation: returns (lQhbndFine, lFhbndFine)."""
    B, D, V = s.B, s.D, s.num_variables
    QF, FF = np.zeros(B * B * D), np.zeros(B * B * V)
    check(lib().exa_face_unknowns_prolongation(ctypes.byref(s), _dp(QF), _dp(FF), _dp(_f64(QC)), _dp(_f64(FC)),
                                               coarse_level, fine_level, _sub(subface, 2)),
          "faceUnknownsProlongation")
    return QF, FF


def face_unknowns_restriction(s, QC, FC, QF, FF, coarse_level, fine_level, subface):
    """faceUnknownsRestriction: returns the coarse arrays with the restricted fine values added."""
    QC, FC = np.array(QC, dtype=np.float64), np.array(FC, dtype=np.float64)
    check(lib().exa_face_unknowns_restriction(ctypes.byref(s), _dp(QC), _dp(FC), _dp(_f64(QF)), _dp(_f64(FF)),
                                              coarse_level, fine_level, _sub(subface, 2)),
          "faceUnknownsRestriction")
    return QC, FC


def volume_unknowns_prolongation(s, uC, coarse_level, fine_level, subcell, uF=None):
    uF = np.zeros(s.B ** 3 * s.D) if uF is None else np.array(uF, dtype=np.float64)
    check(lib().exa_volume_unknowns_prolongation(ctypes.byref(s), _dp(uF), _dp(_f64(uC)), coarse_level, fine_level,
                                                 _sub(subcell, 3)), "volumeUnknownsProlongation")
    return uF


def volume_unknowns_restriction(s, uC, uF, coarse_level, fine_level, subcell):
    uC = np.array(uC, dtype=np.float64)
    check(lib().exa_volume_unknowns_restriction(ctypes.byref(s), _dp(uC), _dp(_f64(uF)), coarse_level, fine_level,
                                                _sub(subcell, 3)), "volumeUnknownsRestriction")
    return uC


class DeviceArray:
    """A device buffer (exa_device_alloc) holding a numpy array's bytes."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.dtype, self.shape, self.nbytes = a.dtype, a.shape, a.nbytes
        self.ptr = lib().exa_device_alloc(max(a.nbytes, 8))
        if not self.ptr:
            raise ExaError("exa_device_alloc failed")
        check(lib().exa_copy_to_device(self.ptr, a.ctypes.data, a.nbytes), "exa_copy_to_device")

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().exa_copy_to_host(out.ctypes.data, self.ptr, self.nbytes), "exa_copy_to_host")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib().exa_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ms(timed):
    return ctypes.byref(timed) if timed is not None else None


def volume_prolongation_batch(s, fine, coarse, items, level_delta):
    """fine, coarse: DeviceArray of blocks; items: DeviceArray int32 [n, 5].  Returns the kernel time (ms)."""
    t = ctypes.c_float(0.0)
    check(lib().exa_volume_unknowns_prolongation_batch(
        ctypes.byref(s), fine.ptr, fine.shape[0], coarse.ptr, coarse.shape[0], items.ptr, items.shape[0],
        level_delta, _ms(t)), "volumeUnknownsProlongation (batch)")
    return t.value


def volume_restriction_batch(s, coarse, fine, ptr, items, level_delta):
    t = ctypes.c_float(0.0)
    check(lib().exa_volume_unknowns_restriction_batch(
        ctypes.byref(s), coarse.ptr, coarse.shape[0], fine.ptr, fine.shape[0], ptr.ptr, ptr.shape[0] - 1, items.ptr,
        items.shape[0], level_delta, _ms(t)), "volumeUnknownsRestriction (batch)")
    return t.value


def face_prolongation_batch(s, QF, FF, QC, FC, items, level_delta):
    t = ctypes.c_float(0.0)
    check(lib().exa_face_unknowns_prolongation_batch(
        ctypes.byref(s), QF.ptr, FF.ptr, QF.shape[0], QC.ptr, FC.ptr, QC.shape[0], items.ptr, items.shape[0],
        level_delta, _ms(t)), "faceUnknownsProlongation (batch)")
    return t.value


def face_restriction_batch(s, QC, FC, QF, FF, ptr, items, level_delta):
    t = ctypes.c_float(0.0)
    check(lib().exa_face_unknowns_restriction_batch(
        ctypes.byref(s), QC.ptr, FC.ptr, QC.shape[0], QF.ptr, FF.ptr, QF.shape[0], ptr.ptr, ptr.shape[0] - 1,
        items.ptr, items.shape[0], level_delta, _ms(t)), "faceUnknownsRestriction (batch)")
    return t.value


def batch_predictor_volume(s, luh, dx, dt):
    """performPredictionAndVolumeIntegral over many cells: luh [ncells, B^3*D], dx [ncells,3]."""
    luh = _f64(luh)
    B, D, V = s.B, s.D, s.num_variables
    n = luh.size // (B ** 3 * D)
    dxa = _f64(np.broadcast_to(np.asarray(dx, dtype=np.float64), (n, 3)))
    out = dict(lQhbnd=np.zeros(n * 6 * B * B * D), lFhbnd=np.zeros(n * 6 * B * B * V),
               lduh=np.zeros(n * B ** 3 * V), iters=np.zeros(n, dtype=np.int32))
    check(lib().exa_batch_predictor_volume(ctypes.byref(s), n, _dp(out["lQhbnd"]), _dp(out["lFhbnd"]),
                                           _dp(out["lduh"]), _dp(luh), _dp(dxa), float(dt),
                                           out["iters"].ctypes.data_as(_I)), "batch predictor")
    return out


USER_EXAMPLE_PATH = os.path.join(HERE, "libexa_user_example.so")
_user_example = None


def user_example():
    """The example user solver (examples/user_euler.hip): an Euler PDE with a
    reflecting-wall boundaryValues and an entropy-wave adjustPointSolution,
    compiled through include/exahype_amd/user_solver.hpp into its own shared
    object.  Tables: exa_user_euler_order3, exa_user_euler_order5."""
    global _user_example
    if _user_example is None:
        if not os.path.exists(USER_EXAMPLE_PATH):
            raise ExaError(f"{USER_EXAMPLE_PATH} is missing: build it with `python -m exahype_amd.build`")
        _user_example = ctypes.CDLL(USER_EXAMPLE_PATH)
    return _user_example


# ---------------------------------------------------------------------------
# Device-resident grid
# ---------------------------------------------------------------------------
class Grid:
    """A uniform periodic grid level resident in HBM (exa_grid)."""

    def __init__(self, s, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), boundary=BC_PERIODIC):
        if isinstance(n, int):
            n = (n, n, n)
        self.s = s
        self.n = tuple(n)
        h = ctypes.c_void_p()
        check(lib().exa_grid_create(ctypes.byref(s), n[0], n[1], n[2], _dp(_dx3(lo)), _dp(_dx3(hi)), boundary,
                                    ctypes.byref(h)), "exa_grid_create")
        self._h = h
        self.ncells = lib().exa_grid_num_cells(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().exa_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_user_solver(cls, table_fn, n, precision=PRECISION_FP64, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                         boundary=BC_PERIODIC):
        """A grid whose PDE is a user solver: table_fn is the ctypes function
        EXA_USER_SOLVER exported (returns const exa_user_solver*)."""
        table_fn.restype = ctypes.POINTER(UserSolverTable)
        table_fn.argtypes = []
        us = table_fn()
        if not us:
            raise ExaError("user solver table is NULL")
        t = us.contents
        if isinstance(n, int):
            n = (n, n, n)
        self = cls.__new__(cls)
        self.s = Solver(PDE_USER, t.order, t.num_variables, t.num_parameters, t.terms, 0, t.cfl, precision, 0)
        self.n = tuple(n)
        self.user = t
        h = ctypes.c_void_p()
        check(lib().exa_grid_create_user(us, precision, n[0], n[1], n[2], _dp(_dx3(lo)), _dp(_dx3(hi)), boundary,
                                         ctypes.byref(h)), "exa_grid_create_user")
        self._h = h
        self.ncells = lib().exa_grid_num_cells(h)
        return self

    def set_adjust_solution(self, on):
        """adjustPointSolution after every update (user solvers with the hook)."""
        check(lib().exa_grid_set_adjust_solution(self._h, int(on)), "adjust solution")

    def adjust_solution(self):
        """adjustPointSolution at the grid's time with dt = 0 on every cell (the
        initial condition at t = 0)."""
        check(lib().exa_grid_adjust_solution(self._h), "exa_grid_adjust_solution")

    def init_euler(self, rho0=1.0, amp=0.2, k=(1, 1, 1), vel=(1.0, 1.0, 1.0), p=1.0, nmodes=0,
                   mode_amp=0.0, seed=0):
        ic = EulerIC(rho0, amp, (ctypes.c_double * 3)(*k), (ctypes.c_double * 3)(*vel), p, nmodes, mode_amp,
                     seed)
        check(lib().exa_grid_init_euler(self._h, ctypes.byref(ic)), "exa_grid_init_euler")

    def init_elastic(self, seed=3, width=0.05, amp=1.0):
        check(lib().exa_grid_init_elastic(self._h, seed, width, amp), "exa_grid_init_elastic")

    def upload(self, luh):
        check(lib().exa_grid_upload_solution(self._h, _dp(_f64(luh))), "upload")

    def solution(self):
        out = np.zeros(self.ncells * self.s.B ** 3 * self.s.D)
        check(lib().exa_grid_download_solution(self._h, _dp(out)), "download")
        return out

    def plot_vertices(self, timed=False):
        """Values at every cell's (N+1)^3 equidistant vertices (the Cartesian VTK
        plotter's interpolation, on the device); with timed, also the kernel ms."""
        out = np.zeros(self.ncells * self.s.B ** 3 * self.s.D)
        t = ctypes.c_float(0.0)
        check(lib().exa_grid_plot_vertices(self._h, _dp(out), ctypes.byref(t)), "exa_grid_plot_vertices")
        return (out, t.value) if timed else out

    def cells(self, first, count):
        out = np.zeros(count * self.s.B ** 3 * self.s.D)
        check(lib().exa_grid_download_cells(self._h, int(first), int(count), _dp(out)), "download cells")
        return out

    def picard_totals(self):
        it, nc = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(lib().exa_grid_picard_totals(self._h, ctypes.byref(it), ctypes.byref(nc)), "picard totals")
        return it.value, nc.value

    def step(self, nsteps):
        check(lib().exa_grid_step(self._h, int(nsteps)), "exa_grid_step")

    def synchronize(self):
        check(lib().exa_grid_synchronize(self._h), "synchronize")

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
        check(lib().exa_grid_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), "time")
        return t.value, dt.value, n.value

    def dt_history(self, n):
        out = np.zeros(n)
        check(lib().exa_grid_dt_history(self._h, _dp(out), n), "dt history")
        return out

    def picard_iterations(self):
        out = np.zeros(self.ncells, dtype=np.int32)
        check(lib().exa_grid_picard_iterations(self._h, out.ctypes.data_as(_I)), "picard iterations")
        return out

    def set_profiling(self, on=True):
        check(lib().exa_grid_set_profiling(self._h, int(on)), "profiling")

    def set_fused_corrector(self, on):
        """One-kernel corrector (Riemann + surface + update + dt) or the two passes."""
        check(lib().exa_grid_set_fused_corrector(self._h, int(on)), "fused corrector")

    def set_fused_time_stepping(self, on=True, weight=0.9):
        """fuse-algorithmic-steps with fuse-algorithmic-steps-factor weight."""
        check(lib().exa_grid_set_fused_time_stepping(self._h, int(on), float(weight)), "fused time stepping")

    def predictor_reruns(self):
        n = ctypes.c_long()
        check(lib().exa_grid_predictor_reruns(self._h, ctypes.byref(n)), "predictor reruns")
        return n.value

    def set_predictor_variant(self, variant):
        """EXA_PREDICTOR_AUTO (0), _VALU_LDS64 (1), _VALU_LDS144 (2) or _MFMA (3)."""
        check(lib().exa_grid_set_predictor_variant(self._h, int(variant)), "predictor variant")

    def stage_times(self):
        ms = np.zeros(4)
        n = (ctypes.c_long * 4)()
        check(lib().exa_grid_stage_times(self._h, _dp(ms), n), "stage times")
        return {STAGE_NAMES[i]: (float(ms[i]), int(n[i])) for i in range(4)}


# ---------------------------------------------------------------------------
# kernels::limiter::generic::c and MUSCL-Hancock single-cell entry points
# ---------------------------------------------------------------------------
def project_on_fv(luh, basis_size, ghost_layer_width, lim=None, num_variables=5):
    L = 2 * basis_size - 1
    S = L + 2 * ghost_layer_width
    out = np.zeros(S ** 3 * num_variables) if lim is None else _f64(lim).copy()
    check(lib().exa_project_on_fv_limiter_space(_dp(_f64(luh).ravel()), num_variables, basis_size,
                                                ghost_layer_width, _dp(out)), "project_on_fv")
    return out


def project_on_dg(lim, basis_size, ghost_layer_width, num_variables=5):
    out = np.zeros(basis_size ** 3 * num_variables)
    check(lib().exa_project_on_dg_space(_dp(_f64(lim).ravel()), num_variables, basis_size, ghost_layer_width,
                                        _dp(out)), "project_on_dg")
    return out


def cell_local_min_max(luh, basis_size, num_variables=5):
    mn, mx = np.zeros(num_variables), np.zeros(num_variables)
    check(lib().exa_find_cell_local_min_and_max(_dp(_f64(luh).ravel()), num_variables, basis_size, _dp(mn),
                                                _dp(mx)), "local min/max")
    return mn, mx


def dmp(luh, basis_size, bmin, bmax, relax=1e-4, scaling=1e-3, num_variables=5):
    bmin = _f64(bmin).ravel().copy()
    bmax = _f64(bmax).ravel().copy()
    ok = np.zeros(1, dtype=np.int32)
    check(lib().exa_discrete_maximum_principle_and_min_and_max_search(
        _dp(_f64(luh).ravel()), num_variables, basis_size, relax, scaling, _dp(bmin), _dp(bmax),
        ok.ctypes.data_as(_I)), "dmp")
    return bool(ok[0]), bmin, bmax


def muscl_update(s, patch, dx, dt, return_dt=False):
    """One MUSCL-Hancock update of a patch with filled ghost layers (interior
    written); with return_dt also the admissible dt the reference returns."""
    old = _f64(patch).ravel()
    new = old.copy()
    L = 2 * s.order + 1
    dtmax = ctypes.c_double()
    check(lib().exa_musclhancock_solution_update(ctypes.byref(s), _dp(new), _dp(old), _dp(_dx3(dx)), dt, L, 2,
                                                 ctypes.byref(dtmax)), "musclhancock")
    return (new, dtmax.value) if return_dt else new


def godunov_update(s, patch, dx, dt, patch_size, ghost_layer_width=1, luh_new=None):
    """godunov::c::solutionUpdate of one patch with filled ghost layers;
    returns (new patch, dt_max_allowed).  luh_new defaults to a copy of the
    patch (the reference's caller copies the solution first)."""
    old = _f64(patch).ravel()
    new = old.copy() if luh_new is None else _f64(luh_new).ravel().copy()
    dtmax = ctypes.c_double()
    check(lib().exa_godunov_solution_update(ctypes.byref(s), _dp(new), _dp(old), _dp(_dx3(dx)), dt, patch_size,
                                            ghost_layer_width, ctypes.byref(dtmax)), "godunov")
    return new, dtmax.value


class FVGrid:
    """Device-resident finite-volumes grid (exa_fvgrid): nx*ny*nz cells of
    patch_size^3 subcells, Godunov, global time stepping."""

    def __init__(self, s, patch_size, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), boundary=BC_PERIODIC,
                 scheme=FV_GODUNOV):
        if isinstance(n, int):
            n = (n, n, n)
        self.s, self.L, self.n = s, patch_size, tuple(n)
        h = ctypes.c_void_p()
        check(lib().exa_fvgrid_create(ctypes.byref(s), scheme, patch_size, n[0], n[1], n[2], _dp(_dx3(lo)),
                                      _dp(_dx3(hi)), boundary, ctypes.byref(h)), "exa_fvgrid_create")
        self._h = h
        self.ncells = lib().exa_fvgrid_num_cells(h)
        self.size = self.ncells * patch_size ** 3 * 5

    def close(self):
        if getattr(self, "_h", None):
            lib().exa_fvgrid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, interior):
        a = _f64(interior).ravel()
        assert a.size == self.size
        check(lib().exa_fvgrid_upload(self._h, _dp(a)), "upload")

    def init_euler(self, rho0=1.0, amp=0.2, k=(1, 1, 1), vel=(1.0, 1.0, 1.0), p=1.0, nmodes=0, mode_amp=0.0,
                   seed=0):
        ic = EulerIC(rho0, amp, (ctypes.c_double * 3)(*k), (ctypes.c_double * 3)(*vel), p, nmodes, mode_amp,
                     seed)
        check(lib().exa_fvgrid_init_euler(self._h, ctypes.byref(ic)), "exa_fvgrid_init_euler")

    def solution(self):
        out = np.zeros(self.size)
        check(lib().exa_fvgrid_download(self._h, _dp(out)), "download")
        return out

    def step(self, n=1):
        check(lib().exa_fvgrid_step(self._h, n), "step")

    def synchronize(self):
        check(lib().exa_fvgrid_synchronize(self._h), "synchronize")

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
        check(lib().exa_fvgrid_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), "time")
        return t.value, dt.value, n.value

    def set_profiling(self, on=True):
        check(lib().exa_fvgrid_set_profiling(self._h, int(on)), "set_profiling")

    def stage_times(self):
        ms = np.zeros(2)
        n = (ctypes.c_long * 2)()
        check(lib().exa_fvgrid_stage_times(self._h, _dp(ms), n), "stage_times")
        return {"update": (float(ms[0]), int(n[0])), "finalize": (float(ms[1]), int(n[1]))}


class LimitingGrid:
    """A uniform grid with the a-posteriori subcell limiter (exa_lgrid)."""

    def __init__(self, s, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), boundary=BC_TRANSMISSIVE, helper_layers=1,
                 steps_till_cured=0, relax=1e-4, scaling=1e-3, fv_scheme=None):
        if isinstance(n, int):
            n = (n, n, n)
        self.s = s
        self.n = tuple(n)
        self.fv_scheme = FV_MUSCLHANCOCK if fv_scheme is None else fv_scheme
        h = ctypes.c_void_p()
        check(lib().exa_lgrid_create(ctypes.byref(s), n[0], n[1], n[2], _dp(_dx3(lo)), _dp(_dx3(hi)), boundary,
                                     helper_layers, steps_till_cured, relax, scaling, self.fv_scheme,
                                     ctypes.byref(h)), "exa_lgrid_create")
        self._h = h
        self.ncells = lib().exa_lgrid_num_cells(h)
        self.L = 2 * s.order + 1
        self.S = self.L + 4  # device patches keep two ghost layers for either scheme

    def close(self):
        if getattr(self, "_h", None):
            lib().exa_lgrid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, luh, lim_ic=None):
        check(lib().exa_lgrid_upload(self._h, _dp(_f64(luh).ravel()),
                                     None if lim_ic is None else _dp(_f64(lim_ic).ravel())), "upload")

    def init_blast(self, seed=7, radius=0.2, p_in=10.0, p_out=0.1, noise=0.01):
        check(lib().exa_lgrid_init_blast(self._h, seed, radius, p_in, p_out, noise), "init_blast")

    def step(self, n=1):
        check(lib().exa_lgrid_step(self._h, n), "step")

    def set_fused_time_stepping(self, on=True, weight=0.9):
        """fuse-algorithmic-steps of the limiting solver; set after upload or
        init and before the first step for the fused start."""
        check(lib().exa_lgrid_set_fused_time_stepping(self._h, int(on), float(weight)), "fused time stepping")

    def predictor_reruns(self):
        n = ctypes.c_long()
        check(lib().exa_lgrid_predictor_reruns(self._h, ctypes.byref(n)), "predictor reruns")
        return n.value

    def synchronize(self):
        check(lib().exa_lgrid_synchronize(self._h), "synchronize")

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
        check(lib().exa_lgrid_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), "time")
        return t.value, dt.value, n.value

    def solution(self):
        out = np.zeros(self.ncells * self.s.B ** 3 * 5)
        check(lib().exa_lgrid_download(self._h, _dp(out)), "download")
        return out

    def status(self):
        out = np.zeros(self.ncells, dtype=np.int32)
        irr = ctypes.c_long()
        check(lib().exa_lgrid_status(self._h, out.ctypes.data_as(_I), ctypes.byref(irr)), "status")
        return out, irr.value

    def limiter_solution(self, c):
        out = np.zeros(self.S ** 3 * 5)
        check(lib().exa_lgrid_limiter_solution(self._h, c, _dp(out)), "limiter_solution")
        return out

    def min_max(self):
        mn = np.zeros(self.ncells * 30)
        mx = np.zeros(self.ncells * 30)
        check(lib().exa_lgrid_min_max(self._h, _dp(mn), _dp(mx)), "min_max")
        return mn, mx

    def set_profiling(self, on):
        check(lib().exa_lgrid_set_profiling(self._h, int(on)), "set_profiling")

    def stage_times(self):
        """ms and launches per stage: predictor, riemann, update (the status
        kernel: DG update, DMP, status), dt, limiter, fv (FV patch updates)."""
        names = ("predictor", "riemann", "update", "dt", "limiter", "fv")
        ms = np.zeros(len(names))
        n = (ctypes.c_long * len(names))()
        check(lib().exa_lgrid_stage_times(self._h, _dp(ms), n), "stage_times")
        return {names[i]: (float(ms[i]), int(n[i])) for i in range(len(names))}

    def picard_totals(self):
        it, nc = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(lib().exa_lgrid_picard_totals(self._h, ctypes.byref(it), ctypes.byref(nc)), "picard totals")
        return it.value, nc.value


class PatchSet:
    """Many independent periodic Euler patches of uneven size and order,
    advanced together with per-patch dt (exa_patchset)."""

    def __init__(self, specs, precision=PRECISION_FP64, hbm_budget_gb=200.0, keep_max_cells=0):
        arr = (PatchSpec * len(specs))()
        for j, spec in enumerate(specs):
            pde, n, order, k, amp, seed, width = _patch_fields(spec)
            arr[j].n, arr[j].order, arr[j].amp = int(n), int(order), float(amp)
            arr[j].pde, arr[j].seed, arr[j].width = int(pde), int(seed), float(width)
            for d in range(3):
                arr[j].k[d] = float(k[d])
        self.specs = list(specs)
        h = ctypes.c_void_p()
        check(lib().exa_patchset_create(arr, len(specs), precision, hbm_budget_gb, int(keep_max_cells),
                                        ctypes.byref(h)), "exa_patchset_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().exa_patchset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps, use_graph=True, autotune=True, warmup=0):
        check(lib().exa_patchset_run(self._h, warmup, nsteps, int(use_graph), int(autotune)), "exa_patchset_run")

    def stats(self):
        sec, dof, nw = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().exa_patchset_stats(self._h, ctypes.byref(sec), ctypes.byref(dof), ctypes.byref(nw)), "stats")
        return sec.value, dof.value, nw.value

    def predictor_variant(self, order, pde=PDE_EULER):
        return lib().exa_patchset_group_variant(self._h, pde, order)

    def set_fused_corrector(self, on):
        check(lib().exa_patchset_set_fused_corrector(self._h, int(on)), "fused corrector")

    def set_profiling(self, on):
        check(lib().exa_patchset_set_profiling(self._h, int(on)), "profiling")

    def group_times(self, pde, order):
        """(predictor ms, corrector ms, predictor launches, Picard iterations,
        predicted cells) of one group over the last profiled run."""
        p, c, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
        it, cells = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(lib().exa_patchset_group_times(self._h, pde, order, ctypes.byref(p), ctypes.byref(c), ctypes.byref(n),
                                             ctypes.byref(it), ctypes.byref(cells)), "group times")
        return p.value, c.value, n.value, it.value, cells.value

    def solution(self, j):
        pde, n, order = _patch_fields(self.specs[j])[:3]
        D = 12 if pde == PDE_ELASTIC else 5
        out = np.zeros(n ** 3 * (order + 1) ** 3 * D)
        check(lib().exa_patchset_solution(self._h, j, _dp(out)), "solution")
        return out

    def time(self, j):
        t, dt = ctypes.c_double(), ctypes.c_double()
        check(lib().exa_patchset_time(self._h, j, ctypes.byref(t), ctypes.byref(dt)), "time")
        return t.value, dt.value


def draw_patches(seed, npatches=96, sides=(9, 15, 27, 45, 81), orders=(2, 3, 4, 5, 6, 7)):
    """The seeded patch set: side uniform in `sides`, order uniform in
    `orders`, wave vector uniform in {1,2,3}^3, amplitude uniform in
    [0.05, 0.3] (numpy PCG64 stream of `seed`, drawn in this order per patch)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(npatches):
        n = int(rng.choice(sides))
        order = int(rng.choice(orders))
        k = tuple(int(x) for x in rng.integers(1, 4, size=3))
        amp = float(rng.uniform(0.05, 0.3))
        out.append((n, order, k, amp))
    return out
