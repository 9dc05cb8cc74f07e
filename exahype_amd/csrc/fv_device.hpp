// fv_device.hpp -- device building blocks of the finite-volume kernels
// (Euler, FP64): the directional Euler flux and the pointwise Rusanov flux of
// the reference's FV Riemann solver, minmod, and std::min/std::max with the
// reference's operand order.
//   RUS = ExaHyPE/kernels/finitevolumes/riemannsolvers/c/rusanov.cpph
//   MH  = ExaHyPE/kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph
#pragma once

#include <hip/hip_runtime.h>

#include "exa_device.hpp"

namespace exa {

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min(a, b)
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max(a, b)

// Euler flux in direction d only (the d-th row of the Euler functor's flux,
// same expression tree).
__device__ inline void euler_flux_dir(const double* Q, int d, double* F) {
  const double GAMMA = 1.4;
  const double irho = 1.0 / Q[0];
  const double p = (GAMMA - 1.0) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
  const double qn = Q[d + 1];
  F[0] = qn;
#pragma unroll
  for (int m = 1; m < 4; ++m) F[m] = (m == d + 1) ? irho * qn * Q[m] + p : irho * qn * Q[m];
  F[4] = irho * qn * (Q[4] + p);
}

__device__ inline double minmod(double a, double b) {  // MH:27-40
  if (a * b < 0.0) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

// Rusanov flux of RUS:98-156 (useFlux, no NCP) from each side's directional
// flux and largest |eigenvalue|: s_max = max over both sides, and
// fn = 0.5 s_max (qL - qR) + 0.5 (FL + FR), written to fnL and fnR alike.
__device__ inline void rusanov_from_parts(const double* qL, const double* FL, double lamL, const double* qR,
                                          const double* FR, double lamR, double* fn) {
  const double smax = dmax(lamR, lamL);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double f = 0.5 * smax * (qL[v] - qR[v]);
    f += 0.5 * (FL[v] + FR[v]);
    fn[v] = f;
  }
}

}  // namespace exa
