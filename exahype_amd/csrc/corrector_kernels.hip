// corrector_kernels.hip -- the corrector of one time step in one kernel:
// Rusanov Riemann solves on the six faces of a cell, surface integral,
// solution update and the cell's stable time step.
//
// Reference semantics (K = ExaHyPE/kernels/aderdg/generic/c/3d):
//   K/riemannSolverNonlinear.cpph:37-117 (face averages :49-66, smax :68-78,
//     flux :102-112, NCP :81-109), applied at every face once
//     (ADERDGSolver::solveRiemannProblemAtInterface, ADERDGSolver.cpp:2789);
//   transmissive boundary faces: boundaryValues returns the inside state, so
//     the Rusanov flux there is the inner flux (ADERDGSolver.cpp:2893-2966);
//   K/surfaceIntegralNonlinear.cpph:21-90, K/solutionUpdate.cpph:23-47,
//   K/stableTimeStepSize.cpph:21-63 (as update_kernel).
//
// A face is shared by two cells.  Instead of a face pass that writes the flux
// into both cells' lFhbnd blocks and an update pass that reads them back, each
// cell solves the Riemann problems of its own six faces from its face blocks
// and the neighbours' opposite ones.  Both cells of a face form the face
// averages in the reference's (i, j) node order from the same inputs and get
// the same flux bit for bit, so conservation is unchanged; every face is solved
// twice, which costs a few FLOPs per face node and saves the flux write and
// re-read.  Workgroups are ordered so that each XCD walks one contiguous run
// of cells: the x- and y-neighbours' face blocks are read again through the
// same L2.
#include "aderdg.hpp"
#include "exa_device.hpp"

#include <cfloat>
#include <type_traits>

namespace exa {

// Cell of a workgroup: the dispatcher deals workgroups to the 8 XCDs in turn,
// so workgroup b runs on XCD b % 8; XCD x takes cells [x per, (x+1) per).
__device__ inline long xcd_ordered_cell(long ncells) {
  const long per = (ncells + 7) / 8;
  return (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

template <int B, class PDE, bool USE_NCP, typename TF>
__global__ __launch_bounds__(node_threads(B)) void corrector_kernel(CorrectorArgs a) {
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B, B3 = B2 * B;
  constexpr int NT = node_threads(B);
  constexpr int NFN = 6 * B2;                 // face nodes of a cell
  constexpr int R = (NFN + NT - 1) / NT;      // face nodes per thread
  __shared__ TF sPart[6 * 2 * D * B2];        // w_i w_j Q of both sides, [face][side][var][face node]
  __shared__ TF sAvg[6 * 2 * D];              // face averages [face][side][var]
  __shared__ TF sSmax[6];
  __shared__ TF sFlux[6 * B2 * V];            // the cell's face fluxes [face][face node][var]
  __shared__ double sRed[NT / kWave];

  const long ncells = (long)a.nx * a.ny * a.nz;
  const long cell = xcd_ordered_cell(ncells);
  if (cell >= ncells) return;  // whole workgroup
  const int tid = threadIdx.x;
  const int ci = (int)(cell % a.nx), cj = (int)((cell / a.nx) % a.ny), ck = (int)(cell / ((long)a.nx * a.ny));
  // neighbour across face f (-x, +x, -y, +y, -z, +z); -1 at a transmissive boundary
  auto neighbour = [&](int f) -> long {
    int i = ci, j = cj, k = ck;
    const int s = (f & 1) ? 1 : -1;
    if (f < 2) i += s; else if (f < 4) j += s; else k += s;
    if (!a.periodic && (i < 0 || i >= a.nx || j < 0 || j >= a.ny || k < 0 || k >= a.nz)) return -1;
    i = (i + a.nx) % a.nx;
    j = (j + a.ny) % a.ny;
    k = (k + a.nz) % a.nz;
    return ((long)k * a.ny + j) * a.nx + i;
  };
  const TF* const Qb = static_cast<const TF*>(a.lQhbnd);
  const TF* const Fb = static_cast<const TF*>(a.lFhbnd);

  // ---- face nodes: load both sides, stage the weighted states ----
  TF qL[R][D], qR[R][D], fL[R][V], fR[R][V];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = tid + r * NT;
    const bool valid = e < NFN;
    const int ec = valid ? e : NFN - 1;
    const int f = ec / B2, fn = ec - (ec / B2) * B2;
    const bool own_left = (f & 1) != 0;  // face 2d+1: this cell is the left one
    const long nb = neighbour(f);
    const TF* qo = Qb + ((size_t)cell * 6 + f) * (B2 * D) + (size_t)fn * D;
    const TF* fo = Fb + ((size_t)cell * 6 + f) * (B2 * V) + (size_t)fn * V;
    // transmissive boundary: the outside state is the inside one
    const TF* qn = nb >= 0 ? Qb + ((size_t)nb * 6 + (f ^ 1)) * (B2 * D) + (size_t)fn * D : qo;
    const TF* fnb = nb >= 0 ? Fb + ((size_t)nb * 6 + (f ^ 1)) * (B2 * V) + (size_t)fn * V : fo;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const TF o = qo[k], n = qn[k];
      qL[r][k] = own_left ? o : n;
      qR[r][k] = own_left ? n : o;
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const TF o = fo[k], n = fnb[k];
      fL[r][k] = own_left ? o : n;
      fR[r][k] = own_left ? n : o;
    }
    const TF weight = static_cast<TF>(exa_gl_weights[N][fn / B]) * static_cast<TF>(exa_gl_weights[N][fn % B]);
    if (valid) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        sPart[((f * 2 + 0) * D + k) * B2 + fn] = weight * qL[r][k];
        sPart[((f * 2 + 1) * D + k) * B2 + fn] = weight * qR[r][k];
      }
    }
  }
  __syncthreads();
  // ---- face averages in the reference's node order (:49-66), one thread per
  // (face, side, variable), then smax per face (:68-78) ----
  for (int t = tid; t < 12 * D; t += NT) {
    TF acc = TF(0);
#pragma unroll
    for (int fn = 0; fn < B2; ++fn) acc += sPart[t * B2 + fn];
    sAvg[t] = acc;
  }
  __syncthreads();
  if (tid < 6) {
    const int f = tid, d = f >> 1;
    TF avL[D], avR[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      avL[k] = sAvg[(f * 2 + 0) * D + k];
      avR[k] = sAvg[(f * 2 + 1) * D + k];
    }
    const TF sL = PDE::max_abs_eigenvalue(avL, d);
    const TF sR = PDE::max_abs_eigenvalue(avR, d);
    sSmax[f] = (sL < sR) ? sR : sL;
  }
  __syncthreads();
  // ---- Rusanov flux of every face node (:92-112) ----
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = tid + r * NT;
    if (e < NFN) {
      const int f = e / B2, fn = e - (e / B2) * B2;
      const int d = f >> 1;
      const bool own_left = (f & 1) != 0;
      const TF smax = sSmax[f];
      TF ncp[V];
#pragma unroll
      for (int k = 0; k < V; ++k) ncp[k] = TF(0);
      if constexpr (USE_NCP) {
        TF gradQ[3][V];
        TF Qavg[D];
#pragma unroll
        for (int k = 0; k < V; ++k) {
          const TF jump = qR[r][k] - qL[r][k];
          gradQ[0][k] = d == 0 ? jump : TF(0);
          gradQ[1][k] = d == 1 ? jump : TF(0);
          gradQ[2][k] = d == 2 ? jump : TF(0);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) Qavg[k] = TF(0.5) * (qR[r][k] + qL[r][k]);
        PDE::ncp(Qavg, gradQ, ncp);
      }
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const TF F = TF(0.5) * (fR[r][k] + fL[r][k]) - TF(0.5) * smax * (qR[r][k] - qL[r][k]);
        // the left cell's block gets F + ncp / 2, the right cell's F - ncp / 2
        TF v = F;
        if constexpr (USE_NCP) v = own_left ? F + TF(0.5) * ncp[k] : F - TF(0.5) * ncp[k];
        sFlux[(f * B2 + fn) * V + k] = v;
      }
    }
  }
  __syncthreads();
  // ---- surface integral + update + stable dt (as update_kernel) ----
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const double invDx[3] = {1.0 / a.dx[0], 1.0 / a.dx[1], 1.0 / a.dx[2]};
  double* luh = a.luh + (size_t)cell * (B3 * D) + node * D;
  double q[D];
#pragma unroll
  for (int m = 0; m < D; ++m) q[m] = luh[m];
  const double wx = exa_gl_weights[N][x], wy = exa_gl_weights[N][y], wz = exa_gl_weights[N][z];
  const TF* lduhp = static_cast<const TF*>(a.lduh) + (size_t)cell * (B3 * V) + node * V;
  double du[V];
#pragma unroll
  for (int m = 0; m < V; ++m) du[m] = lduhp[m];
  {  // x faces, (z,y)
    const double us = wz * wy * invDx[0];
    const double fr = exa_fr_coeff[N][x], fl = exa_fl_coeff[N][x];
#pragma unroll
    for (int m = 0; m < V; ++m)
      du[m] -= ((double)sFlux[(1 * B2 + z * B + y) * V + m] * fr - (double)sFlux[(0 * B2 + z * B + y) * V + m] * fl) * us;
  }
  {  // y faces, (z,x)
    const double us = wz * wx * invDx[1];
    const double fr = exa_fr_coeff[N][y], fl = exa_fl_coeff[N][y];
#pragma unroll
    for (int m = 0; m < V; ++m)
      du[m] -= ((double)sFlux[(3 * B2 + z * B + x) * V + m] * fr - (double)sFlux[(2 * B2 + z * B + x) * V + m] * fl) * us;
  }
  {  // z faces, (y,x)
    const double us = wy * wx * invDx[2];
    const double fr = exa_fr_coeff[N][z], fl = exa_fl_coeff[N][z];
#pragma unroll
    for (int m = 0; m < V; ++m)
      du[m] -= ((double)sFlux[(5 * B2 + y * B + x) * V + m] * fr - (double)sFlux[(4 * B2 + y * B + x) * V + m] * fl) * us;
  }
  const double updateSize = a.dt[0] / (wz * wy * wx);
#pragma unroll
  for (int m = 0; m < V; ++m) q[m] += du[m] * updateSize;
  if (active) {
#pragma unroll
    for (int m = 0; m < V; ++m) luh[m] = q[m];
  }
  double dtn = active ? node_stable_dt<PDE>(q, invDx, a.cfl_pnpm) : DBL_MAX;
  dtn = dtn < DBL_MAX ? dtn : DBL_MAX;  // a NaN or inf candidate never wins (std::min order)
  dtn = block_min<NT>(dtn, sRed);
  if (tid == 0) atomic_min_positive(a.dt_slots + cell % kDtSlots, dtn);
}

template <int B, class PDE>
static hipError_t launch_corrector_t(const SolverDesc& sd, const CorrectorArgs& a, hipStream_t s) {
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

}  // namespace exa
