// corrector_kernels.hip -- instantiation and dispatch of the one-kernel
// corrector (corrector_kernel, include/exahype_amd/device/corrector.hpp) for
// the library's PDEs.
#include "aderdg.hpp"
#include "corrector.hpp"
#include "exa_device.hpp"

#include <type_traits>

namespace exa {

template <int B, class PDE>
static hipError_t launch_corrector_t(const SolverDesc& sd, const CorrectorArgs& a, hipStream_t s) {
  if (a.boundary == kBoundaryUser && !HasBoundaryValues<PDE, double>) return hipErrorNotSupported;
  if (a.adjust && !HasAdjustPointSolution<PDE>) return hipErrorNotSupported;
  const long ncells = (long)a.nx * a.ny * a.nz;
  const long per = (ncells + 7) / 8;
  const dim3 grid((unsigned)(8 * per)), block(node_threads(B));
  constexpr bool ELASTIC = std::is_same_v<PDE, Elastic>;
  const bool ncp = (sd.terms & kNCP) != 0;
  if (ELASTIC && !ncp) return hipErrorNotSupported;
  if (sd.precision == kFP32Bulk) {
    if (ncp) {
      if constexpr (ELASTIC) hipLaunchKernelGGL((corrector_kernel<B, PDE, true, float>), grid, block, 0, s, a);
      else return hipErrorNotSupported;
    } else {
      hipLaunchKernelGGL((corrector_kernel<B, PDE, false, float>), grid, block, 0, s, a);
    }
  } else if (ncp) {
    hipLaunchKernelGGL((corrector_kernel<B, PDE, true, double>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((corrector_kernel<B, PDE, false, double>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_corrector(const SolverDesc& sd, const CorrectorArgs& a, hipStream_t s) {
  if (sd.linear) return hipErrorNotSupported;
  if ((long)a.nx * a.ny * a.nz <= 0) return hipSuccess;
  if (sd.pde == kPdeEuler && sd.npar == 0) {
    switch (sd.order) {
      case 2: return launch_corrector_t<3, Euler<0>>(sd, a, s);
      case 3: return launch_corrector_t<4, Euler<0>>(sd, a, s);
      case 4: return launch_corrector_t<5, Euler<0>>(sd, a, s);
      case 5: return launch_corrector_t<6, Euler<0>>(sd, a, s);
      case 6: return launch_corrector_t<7, Euler<0>>(sd, a, s);
      case 7: return launch_corrector_t<8, Euler<0>>(sd, a, s);
      default: return hipErrorNotSupported;
    }
  }
  if (sd.pde == kPdeEuler && sd.npar == 1 && sd.order == 3) return launch_corrector_t<4, Euler<1>>(sd, a, s);
  if (sd.pde == kPdeElastic && sd.nvar == Elastic::V && sd.npar == Elastic::P) {
    switch (sd.order) {
      case 2: return launch_corrector_t<3, Elastic>(sd, a, s);
      case 3: return launch_corrector_t<4, Elastic>(sd, a, s);
      case 4: return launch_corrector_t<5, Elastic>(sd, a, s);
      case 5: return launch_corrector_t<6, Elastic>(sd, a, s);
      default: return hipErrorNotSupported;
    }
  }
  return hipErrorNotSupported;
}

// The library's PDEs (Euler, Elastic) have neither boundary_values nor
// adjust_point_solution: those hooks are for PDEs a user compiles in
// (exahype_amd/user_solver.hpp).
bool corrector_supports(const SolverDesc& sd, int boundary, bool adjust) {
  (void)sd;
  return boundary != kBoundaryUser && !adjust;
}

}  // namespace exa
