// user_solver.hpp -- compiling a user's PDE into the device kernels.
//
// The reference's user solver derives from the generated Abstract<Solver>
// and implements flux, eigenvalues, nonConservativeProduct, algebraicSource,
// boundaryValues and adjustPointSolution (UserSolverInterface.h:64-171); the
// generic kernels are templates over that SolverType.  Here the same role is
// played by a struct of static __device__ functions (the interface is listed
// in device/exa_device.hpp, "PDE functor interface"), and this header
// instantiates the library's kernel templates for it in the user's own
// translation unit, compiled with hipcc --offload-arch=gfx950 -std=c++20
// -I<repo>/include -I<repo>/include/exahype_amd/device.  The result is a
// table of launch functions (exa_user_solver, exahype_amd.h) that
// exa_grid_create_user drives; the library itself is not rebuilt.
//
//   struct MyPDE { static constexpr int V = 5, P = 0; ... };
//   EXA_USER_SOLVER(my_solver_order5, MyPDE, 5, exa::kFlux)
//   // host: exa_grid_create_user(my_solver_order5(), EXA_PRECISION_FP64, ...)
#pragma once

#include "../exahype_amd.h"
#include "device/aderdg.hpp"
#include "device/corrector.hpp"
#include "device/exa_device.hpp"
#include "device/predictor.hpp"

namespace exa {

template <class PDE>
concept HasCfl = requires { PDE::CFL; };

template <class PDE, int ORDER, int TERMS>
struct UserSolver {
  static constexpr int B = ORDER + 1;
  static constexpr int V = PDE::V, D = PDE::V + PDE::P;
  static constexpr bool USE_NCP = (TERMS & kNCP) != 0;
  static_assert(ORDER >= 1 && ORDER <= 7, "orders 1..7");
  // The user route runs the predictor without a device workspace: every time
  // slice of the Picard iterate fits LDS or registers (orders up to 6 in FP64).
  static_assert(PredictorLayout<B, V, D, USE_NCP, double>::ws_bytes == 0 &&
                    PredictorLayout<B, V, D, USE_NCP, float>::ws_bytes == 0,
                "this order needs the predictor's device workspace, which user solvers do not have");

  static int predictor(const void* args, long ncells, int precision, void* stream) {
    if (ncells <= 0) return 0;
    const PredictorArgs& a = *static_cast<const PredictorArgs*>(args);
    const dim3 grid((unsigned)ncells), block(node_threads(B));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == EXA_PRECISION_FP32_BULK)
      hipLaunchKernelGGL((predictor_kernel<B, PDE, TERMS, false, float>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((predictor_kernel<B, PDE, TERMS, false, double>), grid, block, 0, s, a);
    return (int)hipGetLastError();
  }

  static int corrector(const void* args, int precision, void* stream) {
    const CorrectorArgs& a = *static_cast<const CorrectorArgs*>(args);
    const long ncells = (long)a.nx * a.ny * a.nz;
    if (ncells <= 0) return 0;
    const long per = (ncells + 7) / 8;  // XCD-ordered cells: eight runs
    const dim3 grid((unsigned)(8 * per)), block(node_threads(B));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == EXA_PRECISION_FP32_BULK)
      hipLaunchKernelGGL((corrector_kernel<B, PDE, USE_NCP, float>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((corrector_kernel<B, PDE, USE_NCP, double>), grid, block, 0, s, a);
    return (int)hipGetLastError();
  }

  // per-cell stable dt of the current solution into the dt slots (the
  // initial time step size)
  static int stable_dt(const void* args, long ncells, void* stream) {
    if (ncells <= 0) return 0;
    const UpdateArgs& a = *static_cast<const UpdateArgs*>(args);
    hipLaunchKernelGGL((update_kernel<B, PDE, false, false, false, double>), dim3((unsigned)ncells),
                       dim3(node_threads(B)), 0, static_cast<hipStream_t>(stream), a);
    return (int)hipGetLastError();
  }

  static int adjust(const void* args, long ncells, void* stream) {
    if (ncells <= 0) return 0;
    if constexpr (!HasAdjustPointSolution<PDE>) return (int)hipErrorNotSupported;
    const AdjustArgs& a = *static_cast<const AdjustArgs*>(args);
    hipLaunchKernelGGL((adjust_kernel<B, PDE>), dim3((unsigned)ncells), dim3(node_threads(B)), 0,
                       static_cast<hipStream_t>(stream), a);
    return (int)hipGetLastError();
  }

  static const exa_user_solver* table() {
    static const exa_user_solver t = {EXA_USER_SOLVER_ABI,
                                      ORDER,
                                      PDE::V,
                                      PDE::P,
                                      TERMS,
                                      cfl(),
                                      HasBoundaryValues<PDE, double> && HasBoundaryValues<PDE, float> ? 1 : 0,
                                      HasAdjustPointSolution<PDE> ? 1 : 0,
                                      &predictor,
                                      &corrector,
                                      &stable_dt,
                                      &adjust};
    return &t;
  }

  static constexpr double cfl() {
    if constexpr (HasCfl<PDE>) return PDE::CFL;
    return 0.9;
  }
};

}  // namespace exa

// Exports `extern "C" const exa_user_solver* NAME(void)` for PDE at ORDER
// with TERMS (exa::kFlux | exa::kSource | exa::kNCP, the spec file's terms).
#define EXA_USER_SOLVER(NAME, PDE, ORDER, TERMS) \
  extern "C" const exa_user_solver* NAME(void) { return exa::UserSolver<PDE, ORDER, TERMS>::table(); }
