"""The C-ABI library loads and exports every entry point include/exahype_amd.h
declares; argument validation that needs no GPU."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exahype_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_python_binding_set(exa):
    assert declared_functions() == sorted(exa.EXPORTED)


def test_library_exports_every_declared_symbol(exa):
    out = subprocess.run(["nm", "-D", "--defined-only", exa.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (exa_[a-z0-9_]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_status_strings(exa):
    L = exa.lib()
    assert L.exa_status_string(0) == b"ok"
    assert L.exa_status_string(2).startswith(b"no compiled")


def test_supported_instantiations(exa):
    assert exa.supported(exa.euler_solver(3))
    assert exa.supported(exa.euler_solver(5))
    assert exa.supported(exa.euler_solver(3, npar=1, terms=exa.TERM_FLUX | exa.TERM_SOURCE | exa.TERM_NCP))
    assert not exa.supported(exa.euler_solver(9))
    s = exa.euler_solver(3)
    s.no_time_averaging = 1  # the same kernels (the time average is taken first)
    assert exa.supported(s)


def test_invalid_arguments_rejected_before_any_device_work(exa):
    L = exa.lib()
    s = exa.euler_solver(3)
    h = ctypes.c_void_p()
    lo = (ctypes.c_double * 3)(0, 0, 0)
    hi = (ctypes.c_double * 3)(1, 1, 1)
    assert L.exa_grid_create(ctypes.byref(s), 0, 4, 4, lo, hi, 0, ctypes.byref(h)) == 3
    assert L.exa_grid_create(ctypes.byref(s), 4, 4, 4, lo, hi, 7, ctypes.byref(h)) == 2
    bad = exa.euler_solver(9)
    assert L.exa_grid_create(ctypes.byref(bad), 4, 4, 4, lo, hi, 0, ctypes.byref(h)) == 2
    assert L.exa_space_time_predictor_nonlinear(ctypes.byref(s), None, None, None, None, None, None, None,
                                                None, 0.1, None) == 3
    with pytest.raises(exa.ExaError):
        exa.check(3, "probe")


def test_fp32_bulk_is_grid_only(exa):
    """The single-cell drop-ins are FP64 (the reference's arithmetic); the FP32-bulk
    form is rejected before any device work."""
    import numpy as np
    s = exa.euler_solver(3, precision=exa.PRECISION_FP32_BULK)
    assert exa.supported(s)
    with pytest.raises(exa.ExaError):
        exa.space_time_predictor_nonlinear(s, np.ones(64 * 5), 0.1, 1e-3)


def test_linear_solver_rejected_by_nonlinear_only_entry_points(exa):
    """A solver with the linear flag has no nonlinear predictor, volume or
    surface integral, Riemann solver or limiting grid: those entry points
    return EXA_E_UNSUPPORTED (2) before any device work, instead of running the
    Picard / Rusanov kernels for it.  solutionUpdate and stableTimeStepSize are
    shared by both kinds of solver."""
    import numpy as np
    L = exa.lib()
    s = exa.elastic_solver(4, linear=True)
    B, D, V = 5, 12, 9
    luh, dx = np.ones(B ** 3 * D), np.full(3, 0.1)
    qb, fb = np.zeros(6 * B * B * D), np.zeros(6 * B * B * V)
    P = ctypes.POINTER(ctypes.c_double)
    dp = lambda a: a.ctypes.data_as(P)
    assert L.exa_space_time_predictor_nonlinear(ctypes.byref(s), dp(qb), dp(fb), None, None, None, None, dp(luh),
                                                dp(dx), 1e-3, None) == 2
    lduh = np.zeros(B ** 3 * V)
    assert L.exa_volume_integral_nonlinear(ctypes.byref(s), dp(lduh), dp(np.zeros(4 * B ** 3 * V)), dp(dx)) == 2
    assert L.exa_surface_integral_nonlinear(ctypes.byref(s), dp(lduh), dp(fb), dp(dx)) == 2
    fl, fr, ql, qr = np.zeros(B * B * V), np.zeros(B * B * V), np.zeros(B * B * D), np.zeros(B * B * D)
    assert L.exa_riemann_solver_nonlinear(ctypes.byref(s), dp(fl), dp(fr), dp(ql), dp(qr), 1e-3, 0) == 2
    h = ctypes.c_void_p()
    lo = (ctypes.c_double * 3)(0, 0, 0)
    hi = (ctypes.c_double * 3)(1, 1, 1)
    assert L.exa_lgrid_create(ctypes.byref(s), 4, 4, 4, lo, hi, 1, 1, 0, 1e-4, 1e-3, 0, ctypes.byref(h)) == 2


GLUE = r'''
#include <exahype_amd/kernels.hpp>
#include <vector>
struct Vec3 { double v[3]; double operator[](int i) const { return v[i]; } };
struct MyEulerSolver {  // the constants a generated Abstract<Solver> carries
  static constexpr int NumberOfVariables = 5;
  static constexpr int NumberOfParameters = 0;
  static constexpr int Order = 3;
  static constexpr double CFL = 0.9;
  static constexpr int ExaDevicePde = EXA_PDE_EULER;
};
// The calls AbstractGenericADERDGSolverImplementation.template makes.
void glue(MyEulerSolver& solver, double* luh, double* lduh, double* lQhbnd, double* lFhbnd,
          double** stu, double** stfu, double* tu, double* tfu, double* FL, double* FR) {
  const Vec3 dx{{0.1, 0.1, 0.1}};
  kernels::aderdg::generic::c::spaceTimePredictorNonlinear<false, true, false, false, MyEulerSolver>(
      solver, lQhbnd, lFhbnd, stu, stfu, tu, tfu, luh, dx, 1e-3);
  kernels::aderdg::generic::c::volumeIntegralNonlinear<false, true, false, 5, 4>(lduh, stfu[0], tfu, dx);
  kernels::aderdg::generic::c::riemannSolverNonlinear<false, MyEulerSolver>(solver, FL, FR, lQhbnd, lQhbnd, 1e-3, 0);
  kernels::aderdg::generic::c::surfaceIntegralNonlinear<5, 4>(lduh, lFhbnd, dx);
  kernels::aderdg::generic::c::solutionUpdate<MyEulerSolver>(solver, luh, lduh, 1e-3);
  (void)kernels::aderdg::generic::c::stableTimeStepSize<MyEulerSolver>(solver, luh, dx);
}
// The limiting solver's calls (Limiter.h:48-141) and the FV glue's MUSCL call.
struct MyADERDGSolver {  // the accessors of exahype::solvers::ADERDGSolver used by the limiter
  int getNumberOfVariables() const { return 5; }
  int getNodesPerCoordinateAxis() const { return 4; }
  int getDMPObservables() const { return 5; }
};
struct MyFVSolver {  // the constants a generated Abstract<FVSolver> carries
  static constexpr int NumberOfVariables = 5;
  static constexpr int NumberOfParameters = 0;
  static constexpr int PatchSize = 7;
  static constexpr int GhostLayerWidth = 2;
  static constexpr double CFL = 0.9;
  static constexpr int ExaDevicePde = EXA_PDE_EULER;
};
double limiter_glue(const MyADERDGSolver* solver, MyFVSolver& fv, double* luh, double* lim, double* limNew,
                    double* mn, double* mx) {
  const Vec3 dx{{0.1, 0.1, 0.1}};
  kernels::limiter::generic::c::projectOnFVLimiterSpace(luh, 5, 4, 2, lim);
  kernels::limiter::generic::c::projectOnDGSpace(lim, 5, 4, 2, luh);
  kernels::limiter::generic::c::findCellLocalMinAndMax(luh, solver, mn, mx);
  const bool ok = kernels::limiter::generic::c::discreteMaximumPrincipleAndMinAndMaxSearch(luh, solver, 1e-4, 1e-3,
                                                                                          mn, mx);
  double maxAdmissibleDt = kernels::finitevolumes::musclhancock::c::solutionUpdate<false, false, true, MyFVSolver>(
      fv, limNew, lim, dx, 1e-3);
  return ok ? maxAdmissibleDt : 0.0;
}
// The EulerFV demonstrator's Godunov solver (patch-size 10, ghost width 1).
struct MyGodunovSolver {
  static constexpr int NumberOfVariables = 5;
  static constexpr int NumberOfParameters = 0;
  static constexpr int PatchSize = 10;
  static constexpr int GhostLayerWidth = 1;
  static constexpr double CFL = 0.9;
  static constexpr int ExaDevicePde = EXA_PDE_EULER;
};
double godunov_glue(MyGodunovSolver& fv, double* luhNew, const double* luh) {
  const Vec3 dx{{0.1, 0.1, 0.1}};
  return kernels::finitevolumes::godunov::c::solutionUpdate<false, false, true, MyGodunovSolver>(fv, luhNew, luh,
                                                                                               dx, 1e-3);
}
// The glue's isLinear branch for a linear elasticity solver.
struct MyElasticSolver {
  static constexpr int NumberOfVariables = 9;
  static constexpr int NumberOfParameters = 3;
  static constexpr int Order = 3;
  static constexpr double CFL = 0.9;
  static constexpr int ExaDevicePde = EXA_PDE_ELASTIC;
};
void linear_glue(MyElasticSolver& solver, double* luh, double* lduh, double* lQhbnd, double* lFhbnd,
                 double** stu, double** stfu, double* tu, double* tfu, double* FL, double* FR, double* ps) {
  const Vec3 dx{{0.1, 0.1, 0.1}};
  kernels::aderdg::generic::c::spaceTimePredictorLinear<true, false, false, true, MyElasticSolver>(
      solver, lQhbnd, lFhbnd, stu, stfu, tu, tfu, luh, dx, 1e-3, ps);
  kernels::aderdg::generic::c::volumeIntegralLinear<true, false, 9, 4>(lduh, tfu, dx);
  kernels::aderdg::generic::c::riemannSolverLinear<true, MyElasticSolver>(solver, FL, FR, lQhbnd, lQhbnd, 1e-3, 0);
  kernels::aderdg::generic::c::surfaceIntegralLinear<9, 4>(lduh, lFhbnd, dx);
  kernels::aderdg::generic::c::solutionUpdate<MyElasticSolver>(solver, luh, lduh, 1e-3);
}
// A solver with a point source (UserSolverInterface.h:75) and the pointSource kernel.
struct MyPointSourceSolver {
  void pointSource(const double* const x, const double t, const double dt, double* forceVector, double* x0) {
    for (int m = 0; m < 9; ++m) forceVector[m] = x[0] + t + dt;
    x0[0] = 0.5; x0[1] = 0.5; x0[2] = 0.5;
  }
};
void point_source_glue(MyPointSourceSolver& solver, double* ps) {
  const Vec3 c{{0.5, 0.5, 0.5}}, dx{{0.1, 0.1, 0.1}};
  kernels::aderdg::generic::c::pointSource(solver, 0.0, 1e-3, c, dx, 9, 3, 4, ps);
}
// The AMR calls of ADERDGSolver's prolongation/restriction (Kernels.h:163-199).
struct Idx2 { int v[2]; int operator[](int i) const { return v[i]; } };
struct Idx3 { int v[3]; int operator[](int i) const { return v[i]; } };
void amr_glue(double* lQf, double* lFf, const double* lQc, const double* lFc, double* luhF, double* luhC) {
  const Idx2 sf{{1, 2}};
  const Idx3 sc{{0, 1, 2}};
  kernels::aderdg::generic::c::faceUnknownsProlongation<5, 1, 4>(lQf, lFf, lQc, lFc, 1, 2, sf);
  kernels::aderdg::generic::c::faceUnknownsRestriction<5, 1, 4>(const_cast<double*>(lQc), const_cast<double*>(lFc),
                                                                  lQf, lFf, 1, 2, sf);
  kernels::aderdg::generic::c::volumeUnknownsProlongation<5, 1, 4>(luhF, luhC, 0, 1, sc);
  kernels::aderdg::generic::c::volumeUnknownsRestriction<5, 1, 4>(luhC, luhF, 0, 1, sc);
}
int main(int argc, char**) {
  if (argc > 99) {
    glue(*(MyEulerSolver*)nullptr, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0);
    linear_glue(*(MyElasticSolver*)nullptr, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0);
    amr_glue(0, 0, 0, 0, 0, 0);
    MyPointSourceSolver pss;
    point_source_glue(pss, 0);
    MyFVSolver fv;
    MyGodunovSolver gv;
    return (int)(limiter_glue(nullptr, fv, 0, 0, 0, 0, 0) + godunov_glue(gv, 0, 0));
  }
  return 0;
}
'''


def test_cpp_dropin_header_compiles_and_links(exa, tmp_path):
    """Generated-glue-style calls compile against include/exahype_amd/kernels.hpp and link
    against the library (not executed: no GPU needed)."""
    src = tmp_path / "glue.cpp"
    src.write_text(GLUE)
    libdir = os.path.dirname(exa.LIB_PATH)
    res = subprocess.run(["g++", "-std=c++14", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                          str(tmp_path / "glue"), "-L", libdir, "-lexahype_amd", "-Wl,-rpath," + libdir],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_patch_sets_rejected_before_any_device_work(exa):
    with pytest.raises(exa.ExaError, match="status 2"):
        exa.PatchSet([(3, 8, (1, 1, 1), 0.1)])  # order outside 2..7
    with pytest.raises(exa.ExaError, match="status 4"):
        exa.PatchSet([(81, 7, (1, 1, 1), 0.1)], hbm_budget_gb=1.0)  # one patch exceeds the budget
    with pytest.raises(exa.ExaError, match="status 3"):
        exa.PatchSet([(3, 3, (1, 1, 1), 0.1)], precision=5)


def test_seeded_patch_set_draw(exa):
    a = exa.draw_patches(2024)
    assert a == exa.draw_patches(2024) and a != exa.draw_patches(2025)
    assert len(a) == 96
    for n, order, k, amp in a:
        assert n in (9, 15, 27, 45, 81) and 2 <= order <= 7
        assert all(x in (1, 2, 3) for x in k) and 0.05 <= amp <= 0.3
    assert {o for _, o, _, _ in a} == set(range(2, 8))


HOOKS = r'''
#include <exahype_amd/kernels.hpp>
#include <cmath>
#include <cstdio>
struct Vec3 { double v[3]; double operator[](int i) const { return v[i]; } };
// The reference Euler demonstrator's reflecting wall and an entropy-wave
// adjustPointSolution, as host callbacks of a user solver.
struct MyEulerSolver {
  static constexpr int NumberOfVariables = 5;
  static constexpr int NumberOfParameters = 0;
  static constexpr int Order = 4;
  static void flux(const double* Q, int d, double* F) {
    const double irho = 1.0 / Q[0];
    const double p = 0.4 * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    F[0] = Q[d + 1];
    for (int m = 1; m < 4; ++m) F[m] = irho * Q[d + 1] * Q[m] + (m == d + 1 ? p : 0.0);
    F[4] = irho * Q[d + 1] * (Q[4] + p);
  }
  void boundaryValues(const double* const x, const double t, const double dt, const int faceIndex,
                      const int normalNonZero, const double* const fluxIn, const double* const stateIn,
                      double* fluxOut, double* stateOut) {
    (void)x; (void)t; (void)dt; (void)faceIndex; (void)fluxIn;
    for (int m = 0; m < 5; ++m) stateOut[m] = stateIn[m];
    stateOut[1 + normalNonZero] = -stateOut[1 + normalNonZero];
    flux(stateOut, normalNonZero, fluxOut);
  }
  void adjustPointSolution(const double* const x, const double t, const double dt, double* Q) {
    const double rho = 1.0 + 0.2 * std::sin(6.283185307179586476925286766559 * (x[0] + x[1] + x[2]));
    const double Q0[5] = {rho, rho, rho, rho, 1.0 / 0.4 + 0.5 * rho * 3.0};
    for (int m = 0; m < 5; ++m) Q[m] = t == 0.0 ? Q0[m] : Q[m] + 0.5 * dt * (Q0[m] - Q[m]);
  }
};
int main() {
  MyEulerSolver solver;
  constexpr int B = 5;
  const Vec3 centre{{0.3, 0.5, 0.7}}, size{{0.1, 0.2, 0.4}};
  double QIn[B * B * 5], FIn[B * B * 5], QOut[B * B * 5], FOut[B * B * 5];
  for (int i = 0; i < B * B * 5; ++i) {
    QIn[i] = (i % 5 == 0 || i % 5 == 4) ? 1.5 + 0.01 * i : 0.1 * std::sin(i);
    FIn[i] = 0.0;
  }
  for (int face = 0; face < 6; ++face) {
    kernels::aderdg::generic::c::boundaryConditions(solver, FOut, QOut, FIn, QIn, centre, size, 0.1, 0.01, face,
                                                    face / 2);
    for (int i = 0; i < B * B * 5; ++i) std::printf("%.17g %.17g\n", QOut[i], FOut[i]);
  }
  double luh[B * B * B * 5];
  for (int i = 0; i < B * B * B * 5; ++i) luh[i] = 1.0 + 0.001 * i;
  kernels::aderdg::generic::c::solutionAdjustment(solver, luh, centre, size, 0.25, 0.02);
  for (int i = 0; i < B * B * B * 5; ++i) std::printf("%.17g\n", luh[i]);
  return 0;
}
'''


def test_host_hook_templates_match_oracle(exa, oracle, tmp_path):
    """boundaryConditions / solutionAdjustment of include/exahype_amd/kernels.hpp
    (host loops over the user's callbacks) against the oracle's restatement
    with the same callbacks; runs on the CPU (only exa_quadrature_node is
    called in the library)."""
    import numpy as np
    src = tmp_path / "hooks.cpp"
    src.write_text(HOOKS)
    libdir = os.path.dirname(exa.LIB_PATH)
    exe = str(tmp_path / "hooks")
    res = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                          exe, "-L", libdir, "-lexahype_amd", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    vals = [list(map(float, line.split())) for line in out.stdout.strip().splitlines()]
    B, n = 5, 5 * 5 * 5
    s = oracle.make_solver(4, 5, 0)
    QIn = np.array([(1.5 + 0.01 * i) if i % 5 in (0, 4) else 0.1 * np.sin(i) for i in range(n)])
    for face in range(6):
        rows = np.array(vals[face * n:(face + 1) * n])
        FOut, QOut = oracle.boundary_conditions(s, None, np.zeros(n), QIn, (0.3, 0.5, 0.7), (0.1, 0.2, 0.4), 0.1,
                                                0.01, face)
        assert np.abs(rows[:, 0] - QOut).max() == 0.0
        assert np.abs(rows[:, 1] - FOut).max() < 1e-15 * np.abs(FOut).max() + 1e-300
    luh = np.array([r[0] for r in vals[6 * n:]])
    exp = oracle.solution_adjustment(s, None, 1.0 + 0.001 * np.arange(B ** 3 * 5), (0.3, 0.5, 0.7),
                                     (0.1, 0.2, 0.4), 0.25, 0.02)
    assert np.abs(luh - exp).max() < 1e-15
