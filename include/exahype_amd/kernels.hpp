// exahype_amd/kernels.hpp -- the kernels::aderdg::generic::c templates of the
// reference (ExaHyPE/kernels/aderdg/generic/Kernels.h:85-213), the limiter
// kernels (ExaHyPE/kernels/limiter/generic/Limiter.h:48-141) and the
// MUSCL-Hancock and Godunov updates (ExaHyPE/kernels/finitevolumes/
// musclhancock/c/musclhancock.h:38-41, godunov/c/godunov.h) re-declared on
// top of the C-ABI, so that generated
// solver glue (Toolkit/src/eu/exahype/solvers/templates/
// AbstractGenericADERDGSolverImplementation.template:46-161,
// AbstractGenericFiniteVolumesSolverInCImplementation.template:48-49) and the
// limiting solver's calls compile unchanged against the HIP implementation.
//
// Differences a caller must know:
//  * PDE callbacks cannot cross into device code, so SolverType names one of
//    the device PDEs compiled into the library:
//        static constexpr int ExaDevicePde = EXA_PDE_EULER;
//    (NumberOfVariables, NumberOfParameters, Order, CFL as in the reference).
//  * dx may be any type with operator[] (tarch::la::Vector<3,double> works).
//  * The limiter functions take any solver type with getNumberOfVariables(),
//    getNodesPerCoordinateAxis() and getDMPObservables() (ADERDGSolver has
//    them); the library implements the default observables (the first
//    getDMPObservables() == getNumberOfVariables() variables, Euler).
//  * Calls are synchronous and take host arrays; errors throw exahype_amd::Error.
//  * For whole grid levels use exa_grid_* instead: these single-cell calls pay
//    a host<->device copy each.
#pragma once

#include <stdexcept>
#include <string>

#include "../exahype_amd.h"

namespace exahype_amd {

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check(int status, const char* what) {
  if (status != EXA_OK) throw Error(std::string(what) + ": " + exa_status_string(status));
}

template <typename SolverType>
inline exa_solver descriptor(int terms, bool noTimeAveraging) {
  exa_solver s;
  s.pde = SolverType::ExaDevicePde;
  s.order = SolverType::Order;
  s.num_variables = SolverType::NumberOfVariables;
  s.num_parameters = SolverType::NumberOfParameters;
  s.terms = terms;
  s.no_time_averaging = noTimeAveraging ? 1 : 0;
  s.cfl = SolverType::CFL;
  s.precision = EXA_PRECISION_FP64;
  s.linear = 0;
  return s;
}

// Kernels that do not depend on the PDE (integrals) only need sizes.
inline exa_solver size_descriptor(int numberOfVariables, int basisSize, int terms, bool noTimeAveraging) {
  exa_solver s;
  s.pde = EXA_PDE_EULER;
  s.order = basisSize - 1;
  s.num_variables = numberOfVariables;
  s.num_parameters = 0;
  s.terms = terms;
  s.no_time_averaging = noTimeAveraging ? 1 : 0;
  s.cfl = 0.9;
  s.precision = EXA_PRECISION_FP64;
  s.linear = 0;
  return s;
}

inline int terms(bool useSource, bool useFlux, bool useNCP) {
  return (useFlux ? EXA_TERM_FLUX : 0) | (useSource ? EXA_TERM_SOURCE : 0) | (useNCP ? EXA_TERM_NCP : 0);
}

template <typename DxVector>
inline void dx3(const DxVector& dx, double out[3]) {
  out[0] = dx[0];
  out[1] = dx[1];
  out[2] = dx[2];
}

}  // namespace exahype_amd

namespace kernels {
namespace aderdg {
namespace generic {
namespace c {

template <bool useSource, bool useFlux, bool useNCP, bool noTimeAveraging, typename SolverType, typename DxVector>
void spaceTimePredictorNonlinear(SolverType& /*solver*/, double* lQhbnd, double* lFhbnd,
                                 double** tempSpaceTimeUnknowns, double** tempSpaceTimeFluxUnknowns,
                                 double* tempUnknowns, double* tempFluxUnknowns, const double* const luh,
                                 const DxVector& dx, const double dt) {
  const exa_solver s =
      exahype_amd::descriptor<SolverType>(exahype_amd::terms(useSource, useFlux, useNCP), noTimeAveraging);
  double d[3];
  exahype_amd::dx3(dx, d);
  exahype_amd::check(
      exa_space_time_predictor_nonlinear(&s, lQhbnd, lFhbnd,
                                         tempSpaceTimeUnknowns ? tempSpaceTimeUnknowns[0] : nullptr,
                                         tempSpaceTimeFluxUnknowns ? tempSpaceTimeFluxUnknowns[0] : nullptr,
                                         tempUnknowns, tempFluxUnknowns, luh, d, dt, nullptr),
      "spaceTimePredictorNonlinear");
}

template <bool useSourceOrNCP, bool useFlux, bool noTimeAveraging, int numberOfVariables, int basisSize,
          typename DxVector>
void volumeIntegralNonlinear(double* lduh, const double* const lFi, const double* const lFhi,
                             const DxVector& dx) {
  const exa_solver s = exahype_amd::size_descriptor(
      numberOfVariables, basisSize, (useFlux ? EXA_TERM_FLUX : 0) | (useSourceOrNCP ? EXA_TERM_SOURCE : 0),
      noTimeAveraging);
  double d[3];
  exahype_amd::dx3(dx, d);
  // the noTimeAveraging form integrates the space-time fluxes lFi
  exahype_amd::check(exa_volume_integral_nonlinear(&s, lduh, noTimeAveraging ? lFi : lFhi, d),
                     "volumeIntegralNonlinear");
}

template <int numberOfVariables, int basisSize, typename DxVector>
void surfaceIntegralNonlinear(double* lduh, const double* const lFbnd, const DxVector& dx) {
  const exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_FLUX, false);
  double d[3];
  exahype_amd::dx3(dx, d);
  exahype_amd::check(exa_surface_integral_nonlinear(&s, lduh, lFbnd, d), "surfaceIntegralNonlinear");
}

template <bool useNCP, typename SolverType>
void riemannSolverNonlinear(SolverType& /*solver*/, double* FL, double* FR, const double* const QL,
                            const double* const QR, const double dt, const int normalNonZero) {
  const exa_solver s = exahype_amd::descriptor<SolverType>(EXA_TERM_FLUX | (useNCP ? EXA_TERM_NCP : 0), false);
  exahype_amd::check(exa_riemann_solver_nonlinear(&s, FL, FR, QL, QR, dt, normalNonZero),
                     "riemannSolverNonlinear");
}

template <typename SolverType>
void solutionUpdate(SolverType& /*solver*/, double* luh, const double* const lduh, const double dt) {
  const exa_solver s = exahype_amd::descriptor<SolverType>(EXA_TERM_FLUX, false);
  exahype_amd::check(exa_solution_update(&s, luh, lduh, dt), "solutionUpdate");
}

template <typename SolverType, typename DxVector>
double stableTimeStepSize(SolverType& /*solver*/, const double* const luh, const DxVector& dx) {
  const exa_solver s = exahype_amd::descriptor<SolverType>(EXA_TERM_FLUX, false);
  double d[3], dt = 0.0;
  exahype_amd::dx3(dx, d);
  exahype_amd::check(exa_stable_time_step_size(&s, luh, d, &dt), "stableTimeStepSize");
  return dt;
}

// ---- linear kernels (the glue's isLinear branch) ----

template <bool usePointSource, bool useSource, bool useFlux, bool useNCP, typename SolverType, typename DxVector>
void spaceTimePredictorLinear(SolverType& /*solver*/, double* lQbnd, double* lFbnd, double** tempSpaceTimeUnknowns,
                              double** tempSpaceTimeFluxUnknowns, double* tempUnknowns, double* tempFluxUnknowns,
                              const double* const luh, const DxVector& dx, const double dt,
                              double* tempPointForceSources) {
  exa_solver s = exahype_amd::descriptor<SolverType>(exahype_amd::terms(useSource, useFlux, useNCP), false);
  s.linear = 1;
  double d[3];
  exahype_amd::dx3(dx, d);
  const bool temps = tempSpaceTimeUnknowns && tempSpaceTimeFluxUnknowns && tempUnknowns && tempFluxUnknowns;
  exahype_amd::check(
      exa_space_time_predictor_linear(&s, lQbnd, lFbnd, temps ? tempSpaceTimeUnknowns[0] : nullptr,
                                      temps ? tempSpaceTimeFluxUnknowns[0] : nullptr,
                                      temps ? tempSpaceTimeFluxUnknowns[1] : nullptr, temps ? tempUnknowns : nullptr,
                                      temps ? tempFluxUnknowns : nullptr, luh, d, dt,
                                      usePointSource ? tempPointForceSources : nullptr),
      "spaceTimePredictorLinear");
}

template <bool useSourceOrNCP, bool useFlux, int numberOfVariables, int basisSize, typename DxVector>
void volumeIntegralLinear(double* lduh, const double* const lFhi, const DxVector& dx) {
  exa_solver s = exahype_amd::size_descriptor(
      numberOfVariables, basisSize, (useFlux ? EXA_TERM_FLUX : 0) | (useSourceOrNCP ? EXA_TERM_NCP : 0), false);
  s.linear = 1;
  double d[3];
  exahype_amd::dx3(dx, d);
  exahype_amd::check(exa_volume_integral_linear(&s, lduh, lFhi, d), "volumeIntegralLinear");
}

template <int numberOfVariables, int basisSize, typename DxVector>
void surfaceIntegralLinear(double* lduh, const double* const lFbnd, const DxVector& dx) {
  exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_NCP, false);
  s.linear = 1;
  double d[3];
  exahype_amd::dx3(dx, d);
  exahype_amd::check(exa_surface_integral_linear(&s, lduh, lFbnd, d), "surfaceIntegralLinear");
}

template <bool useNCP, typename SolverType>
void riemannSolverLinear(SolverType& /*solver*/, double* FL, double* FR, const double* const QL,
                         const double* const QR, const double dt, const int normalNonZero) {
  exa_solver s = exahype_amd::descriptor<SolverType>(useNCP ? EXA_TERM_NCP : EXA_TERM_FLUX, false);
  s.linear = 1;
  exahype_amd::check(exa_riemann_solver_linear(&s, FL, FR, QL, QR, dt, normalNonZero), "riemannSolverLinear");
}

// pointSource (Kernels.h:203, pointSource.cpph:17): the user's pointSource is
// called on the host at every space-time node, as the reference calls it; the
// projection onto the nodal basis runs on the device.
template <typename SolverType, typename DxVector>
void pointSource(SolverType& solver, const double t, const double dt, const DxVector& center, const DxVector& dx,
                 const int numberOfVariables, const int numberOfParameters, const int basisSize,
                 double* pointForceSources) {
  const int order = basisSize - 1;
  const int nodes = (basisSize + 1) * basisSize * basisSize * basisSize;
  double* x0 = new double[3 * nodes];
  double x[3], x0n[3];
  for (int n = 0; n < basisSize + 1; ++n) {
    const double tn = n == 0 ? t : dt * exa_quadrature_node(order, n - 1) + t;
    for (int i = 0; i < basisSize; ++i)
      for (int j = 0; j < basisSize; ++j)
        for (int k = 0; k < basisSize; ++k) {
          x[2] = center[2] + dx[2] * (exa_quadrature_node(order, i) - 0.5);
          x[1] = center[1] + dx[1] * (exa_quadrature_node(order, j) - 0.5);
          x[0] = center[0] + dx[0] * (exa_quadrature_node(order, k) - 0.5);
          const int node = ((n * basisSize + i) * basisSize + j) * basisSize + k;
          solver.pointSource(x, tn, dt, pointForceSources + node * numberOfVariables, x0n);
          x0[3 * node] = x0n[0];
          x0[3 * node + 1] = x0n[1];
          x0[3 * node + 2] = x0n[2];
        }
  }
  exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_SOURCE, false);
  s.num_parameters = numberOfParameters;
  double c[3], d[3];
  exahype_amd::dx3(center, c);
  exahype_amd::dx3(dx, d);
  const int st = exa_point_source_projection(&s, pointForceSources, x0, c, d);
  delete[] x0;
  exahype_amd::check(st, "pointSource");
}

// boundaryConditions (Kernels.h:145, 3d/boundaryConditions.cpph:24-85): the
// user's boundaryValues at every node of a boundary face, on the host, as the
// reference calls it.  Face blocks are [i][j][var] with i along the second
// and j along the first free direction.  Grids with EXA_BC_USER call a
// compiled boundary_values on the device instead (exahype_amd/user_solver.hpp).
template <typename SolverType, typename DxVector>
void boundaryConditions(SolverType& solver, double* FOut, double* QOut, const double* const FIn,
                        const double* const QIn, const DxVector& cellCentre, const DxVector& cellSize,
                        const double t, const double dt, const int faceIndex, const int normalNonZero) {
  constexpr int V = SolverType::NumberOfVariables;
  constexpr int D = V + SolverType::NumberOfParameters;
  constexpr int order = SolverType::Order;
  constexpr int B = order + 1;
  if (normalNonZero < 0 || normalNonZero > 2) throw exahype_amd::Error("boundaryConditions: normalNonZero must be 0, 1 or 2");
  const int f = faceIndex - 2 * normalNonZero;
  const int d1 = normalNonZero == 0 ? 1 : 0;
  const int d2 = normalNonZero == 2 ? 1 : 2;
  double x[3];
  x[normalNonZero] = cellCentre[normalNonZero] + (-0.5 + f) * cellSize[normalNonZero];
  for (int i = 0; i < B; ++i) {
    x[d2] = cellCentre[d2] + cellSize[d2] * (exa_quadrature_node(order, i) - 0.5);
    for (int j = 0; j < B; ++j) {
      x[d1] = cellCentre[d1] + cellSize[d1] * (exa_quadrature_node(order, j) - 0.5);
      solver.boundaryValues(x, t, dt, faceIndex, normalNonZero, FIn + (i * B + j) * V, QIn + (i * B + j) * D,
                            FOut + (i * B + j) * V, QOut + (i * B + j) * D);
    }
  }
}

// solutionAdjustment (Kernels.h:128, 3d/solutionAdjustment.cpph:17-47): the
// user's adjustPointSolution at every node of luh [z][y][x][var], on the
// host.  Grids call a compiled adjust_point_solution on the device instead
// (exa_grid_set_adjust_solution, exa_grid_adjust_solution).
template <typename SolverType, typename DxVector>
void solutionAdjustment(SolverType& solver, double* luh, const DxVector& center, const DxVector& dx, const double t,
                        const double dt) {
  constexpr int D = SolverType::NumberOfVariables + SolverType::NumberOfParameters;
  constexpr int order = SolverType::Order;
  constexpr int B = order + 1;
  double x[3];
  for (int i = 0; i < B; ++i) {
    const double qi = exa_quadrature_node(order, i);
    for (int j = 0; j < B; ++j) {
      const double qj = exa_quadrature_node(order, j);
      for (int k = 0; k < B; ++k) {
        const double qk = exa_quadrature_node(order, k);
        x[2] = center[2] + dx[2] * (qi - 0.5);
        x[1] = center[1] + dx[1] * (qj - 0.5);
        x[0] = center[0] + dx[0] * (qk - 0.5);
        solver.adjustPointSolution(x, t, dt, luh + ((i * B + j) * B + k) * D);
      }
    }
  }
}

// ---- AMR transfer operators (Kernels.h:163-199, amrRoutines.cpph) ----

template <int numberOfVariables, int numberOfParameters, int basisSize, typename FaceIndex>
void faceUnknownsProlongation(double* lQhbndFine, double* lFhbndFine, const double* lQhbndCoarse,
                              const double* lFhbndCoarse, const int coarseGridLevel, const int fineGridLevel,
                              const FaceIndex& subfaceIndex) {
  exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_FLUX, false);
  s.num_parameters = numberOfParameters;
  const int sub[2] = {subfaceIndex[0], subfaceIndex[1]};
  exahype_amd::check(exa_face_unknowns_prolongation(&s, lQhbndFine, lFhbndFine, lQhbndCoarse, lFhbndCoarse,
                                                    coarseGridLevel, fineGridLevel, sub),
                     "faceUnknownsProlongation");
}

template <int numberOfVariables, int numberOfParameters, int basisSize, typename FaceIndex>
void faceUnknownsRestriction(double* lQhbndCoarse, double* lFhbndCoarse, const double* lQhbndFine,
                             const double* lFhbndFine, const int coarseGridLevel, const int fineGridLevel,
                             const FaceIndex& subfaceIndex) {
  exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_FLUX, false);
  s.num_parameters = numberOfParameters;
  const int sub[2] = {subfaceIndex[0], subfaceIndex[1]};
  exahype_amd::check(exa_face_unknowns_restriction(&s, lQhbndCoarse, lFhbndCoarse, lQhbndFine, lFhbndFine,
                                                   coarseGridLevel, fineGridLevel, sub),
                     "faceUnknownsRestriction");
}

template <int numberOfVariables, int numberOfParameters, int basisSize, typename CellIndex>
void volumeUnknownsProlongation(double* luhFine, const double* luhCoarse, const int coarseGridLevel,
                                const int fineGridLevel, const CellIndex& subcellIndex) {
  exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_FLUX, false);
  s.num_parameters = numberOfParameters;
  const int sub[3] = {subcellIndex[0], subcellIndex[1], subcellIndex[2]};
  exahype_amd::check(exa_volume_unknowns_prolongation(&s, luhFine, luhCoarse, coarseGridLevel, fineGridLevel, sub),
                     "volumeUnknownsProlongation");
}

template <int numberOfVariables, int numberOfParameters, int basisSize, typename CellIndex>
void volumeUnknownsRestriction(double* luhCoarse, const double* luhFine, const int coarseGridLevel,
                               const int fineGridLevel, const CellIndex& subcellIndex) {
  exa_solver s = exahype_amd::size_descriptor(numberOfVariables, basisSize, EXA_TERM_FLUX, false);
  s.num_parameters = numberOfParameters;
  const int sub[3] = {subcellIndex[0], subcellIndex[1], subcellIndex[2]};
  exahype_amd::check(exa_volume_unknowns_restriction(&s, luhCoarse, luhFine, coarseGridLevel, fineGridLevel, sub),
                     "volumeUnknownsRestriction");
}

}  // namespace c
}  // namespace generic
}  // namespace aderdg
}  // namespace kernels

namespace kernels {
namespace limiter {
namespace generic {
namespace c {

inline void projectOnFVLimiterSpace(const double* const luh, const int numberOfVariables, const int basisSize,
                                    const int ghostLayerWidth, double* const lim) {
  exahype_amd::check(exa_project_on_fv_limiter_space(luh, numberOfVariables, basisSize, ghostLayerWidth, lim),
                     "projectOnFVLimiterSpace");
}

inline void projectOnDGSpace(const double* const lim, const int numberOfVariables, const int basisSize,
                             const int ghostLayerWidth, double* const luh) {
  exahype_amd::check(exa_project_on_dg_space(lim, numberOfVariables, basisSize, ghostLayerWidth, luh),
                     "projectOnDGSpace");
}

template <typename ADERDGSolverType>
inline void findCellLocalMinAndMax(const double* const luh, const ADERDGSolverType* solver,
                                   double* const localMinPerVariables, double* const localMaxPerVariable) {
  if (solver->getDMPObservables() != solver->getNumberOfVariables())
    throw exahype_amd::Error("findCellLocalMinAndMax: only the default observables are implemented");
  exahype_amd::check(exa_find_cell_local_min_and_max(luh, solver->getNumberOfVariables(),
                                                     solver->getNodesPerCoordinateAxis(), localMinPerVariables,
                                                     localMaxPerVariable),
                     "findCellLocalMinAndMax");
}

template <typename ADERDGSolverType>
inline bool discreteMaximumPrincipleAndMinAndMaxSearch(const double* const luh, const ADERDGSolverType* solver,
                                                       const double relaxationParameter,
                                                       const double differenceScaling,
                                                       double* boundaryMinPerVariables,
                                                       double* boundaryMaxPerVariables) {
  if (solver->getDMPObservables() != solver->getNumberOfVariables())
    throw exahype_amd::Error("discreteMaximumPrincipleAndMinAndMaxSearch: only the default observables are implemented");
  int ok = 0;
  exahype_amd::check(exa_discrete_maximum_principle_and_min_and_max_search(
                         luh, solver->getNumberOfVariables(), solver->getNodesPerCoordinateAxis(),
                         relaxationParameter, differenceScaling, boundaryMinPerVariables, boundaryMaxPerVariables,
                         &ok),
                     "discreteMaximumPrincipleAndMinAndMaxSearch");
  return ok != 0;
}

}  // namespace c
}  // namespace generic
}  // namespace limiter

namespace finitevolumes {
namespace musclhancock {
namespace c {

// Returns the admissible time step size (min over faces of 0.99/3 h / s_max).
template <bool useSource, bool useNCP, bool useFlux, class SolverType, typename DxVector>
double solutionUpdate(SolverType& /*solver*/, double* luh_new, const double* luh, const DxVector& dx, double dt) {
  exa_solver s;
  s.pde = SolverType::ExaDevicePde;
  s.order = (SolverType::PatchSize - 1) / 2;
  s.num_variables = SolverType::NumberOfVariables;
  s.num_parameters = SolverType::NumberOfParameters;
  s.terms = exahype_amd::terms(useSource, useFlux, useNCP);
  s.no_time_averaging = 0;
  s.cfl = SolverType::CFL;
  s.precision = EXA_PRECISION_FP64;
  s.linear = 0;
  double d[3], dt_max = 0.0;
  exahype_amd::dx3(dx, d);
  exahype_amd::check(exa_musclhancock_solution_update(&s, luh_new, luh, d, dt, SolverType::PatchSize,
                                                      SolverType::GhostLayerWidth, &dt_max),
                     "musclhancock::solutionUpdate");
  return dt_max;
}

}  // namespace c
}  // namespace musclhancock

namespace godunov {
namespace c {

// godunov::c::solutionUpdate (ExaHyPE/kernels/finitevolumes/godunov/c/godunov.h,
// c/3d/godunov.cpph:35): the first-order update of one patch with filled ghost
// layers; luh_new holds a copy of luh on entry.  Returns the admissible time
// step size (min over faces of CFL/3 h / s_max).
template <bool useSource, bool useNCP, bool useFlux, class SolverType, typename DxVector>
double solutionUpdate(SolverType& /*solver*/, double* luh_new, const double* luh, const DxVector& dx, double dt) {
  exa_solver s;
  s.pde = SolverType::ExaDevicePde;
  s.order = 0;
  s.num_variables = SolverType::NumberOfVariables;
  s.num_parameters = SolverType::NumberOfParameters;
  s.terms = exahype_amd::terms(useSource, useFlux, useNCP);
  s.no_time_averaging = 0;
  s.cfl = SolverType::CFL;
  s.precision = EXA_PRECISION_FP64;
  s.linear = 0;
  double d[3], dt_max = 0.0;
  exahype_amd::dx3(dx, d);
  exahype_amd::check(exa_godunov_solution_update(&s, luh_new, luh, d, dt, SolverType::PatchSize,
                                                 SolverType::GhostLayerWidth, &dt_max),
                     "godunov::solutionUpdate");
  return dt_max;
}

}  // namespace c
}  // namespace godunov
}  // namespace finitevolumes
}  // namespace kernels
