// corrector.hpp -- the corrector kernels as templates over the PDE functor:
// the Rusanov Riemann solver over a face list, the surface integral + solution
// update + stable time step, and the one-kernel corrector of a Cartesian grid
// (for the library's own PDEs and for a user's __device__ PDE).
//
// Reference semantics (K = ExaHyPE/kernels/aderdg/generic/c/3d):
//   riemann_kernel     K/riemannSolverNonlinear.cpph:37-117
//   update_kernel      K/surfaceIntegralNonlinear.cpph:21-90, K/solutionUpdate.cpph:23-47,
//                      K/stableTimeStepSize.cpph:21-63 (per-cell minimum)
#pragma once

#include "aderdg.hpp"
#include "exa_device.hpp"
#include "predictor.hpp"

#include <cfloat>
#include <type_traits>

namespace exa {
// ---------------------------------------------------------------------------
// Rusanov flux on a list of faces.  One wavefront per face, lane = face node.
// ---------------------------------------------------------------------------
// Lanes per face in the Riemann kernel: several faces share a wavefront for
// B^2 <= 16 (measured faster at N=3); from N=4 on one face per wavefront
// (packing two N=4 faces was measured slower).
__host__ __device__ constexpr int riemann_segment(int B) { return B * B <= 4 ? 4 : (B * B <= 16 ? 16 : 64); }

template <int B, class PDE, bool USE_NCP, typename T>
__global__ __launch_bounds__(256) void riemann_kernel(RiemannArgs a) {
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B;
  static_assert(B2 <= kWave, "one face per wavefront at most");
  // faces per wavefront: aligned segments of SEG lanes, one face node per lane
  constexpr int SEG = riemann_segment(B);
  constexpr int FPW = kWave / SEG;
  const int lane = threadIdx.x & (kWave - 1);
  const long f0 = ((long)blockIdx.x * (256 / kWave) + threadIdx.x / kWave) * FPW;
  if (f0 >= a.nfaces) return;  // whole wavefront exits together
  const long f = f0 + lane / SEG;
  const int sl = lane & (SEG - 1);
  // faces beyond the list, or between two FV-active limiter cells, take part
  // in the segment sums with zeros and store nothing
  bool live = f < a.nfaces;
  const FaceDesc fd = a.faces[live ? f : f0];
  if (live && a.status && a.status[fd.left] >= a.status_limit && a.status[fd.right] >= a.status_limit) live = false;
  const bool active = live && sl < B2;
  const int node = sl < B2 ? sl : 0;
  const int d = fd.dir;
  T* FL = static_cast<T*>(a.lFhbnd) + ((size_t)fd.left * 6 + 2 * d + 1) * (B2 * V);
  T* FR = static_cast<T*>(a.lFhbnd) + ((size_t)fd.right * 6 + 2 * d) * (B2 * V);
  const T* QL = static_cast<const T*>(a.lQhbnd) + ((size_t)fd.left * 6 + 2 * d + 1) * (B2 * D);
  const T* QR = static_cast<const T*>(a.lQhbnd) + ((size_t)fd.right * 6 + 2 * d) * (B2 * D);
  // One face per wavefront: its four blocks go through LDS, read and written
  // with contiguous lanes (a lane's own D values lie 8 D bytes apart).
  constexpr bool STAGE = SEG == kWave;
  constexpr int SW = 2 * B2 * D + 2 * B2 * V;  // staged values per wavefront
  __shared__ T sStage[STAGE ? (256 / kWave) * SW : 1];
  T* const sQL = sStage + (STAGE ? (threadIdx.x / kWave) * SW : 0);
  T* const sQR = sQL + B2 * D;
  T* const sFL = sQR + B2 * D;
  T* const sFR = sFL + B2 * V;
  if constexpr (STAGE) {
    if (!live) return;  // the wavefront's only face
    for (int e = lane; e < B2 * D; e += kWave) {
      sQL[e] = QL[e];
      sQR[e] = QR[e];
    }
    for (int e = lane; e < B2 * V; e += kWave) {
      sFL[e] = FL[e];
      sFR[e] = FR[e];
    }
    wave_lds_sync();
  }
  const T* const qLs = STAGE ? sQL : QL;
  const T* const qRs = STAGE ? sQR : QR;
  const int i = node / B, j = node % B;
  const T weight = tc<T>(exa_gl_weights[N][i]) * tc<T>(exa_gl_weights[N][j]);
  T qL[D], qR[D], QavL[D], QavR[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    qL[k] = qLs[node * D + k];
    qR[k] = qRs[node * D + k];
    QavL[k] = segment_sum<SEG>(active ? weight * qL[k] : T(0));  // :49-66
    QavR[k] = segment_sum<SEG>(active ? weight * qR[k] : T(0));
  }
  const T sL = pde_max_abs_eigenvalue<PDE>(QavL, d);  // :68-78
  const T sR = pde_max_abs_eigenvalue<PDE>(QavR, d);
  const T smax = (sL < sR) ? sR : sL;
  T ncp[V];
#pragma unroll
  for (int k = 0; k < V; ++k) ncp[k] = T(0);
  if (USE_NCP) {
    T gradQ[3][V];
    T Qavg[D];
    // only the normal row is non-zero; rows selected without a run-time
    // index into the private array (which would place it in scratch memory)
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const T jump = qR[k] - qL[k];
      gradQ[0][k] = d == 0 ? jump : T(0);
      gradQ[1][k] = d == 1 ? jump : T(0);
      gradQ[2][k] = d == 2 ? jump : T(0);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) Qavg[k] = T(0.5) * (qR[k] + qL[k]);
    PDE::ncp(Qavg, gradQ, ncp);
  }
  T* const fLd = STAGE ? sFL : FL;
  T* const fRd = STAGE ? sFR : FR;
  if (active) {
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const T fl = fLd[node * V + k], fr = fRd[node * V + k];
      const T F = T(0.5) * (fr + fl) - T(0.5) * smax * (qR[k] - qL[k]);  // :102-112
      if (USE_NCP) {
        fRd[node * V + k] = F - T(0.5) * ncp[k];
        fLd[node * V + k] = F + T(0.5) * ncp[k];
      } else {
        fRd[node * V + k] = F;
        fLd[node * V + k] = F;
      }
    }
  }
  if constexpr (STAGE) {
    wave_lds_sync();
    for (int e = lane; e < B2 * V; e += kWave) {
      FL[e] = sFL[e];
      FR[e] = sFR[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Surface integral + solution update + the cell's stable time step, in double
// on the double solution; TF is the storage type of lduh and the face fluxes.
// ---------------------------------------------------------------------------
template <int B, class PDE, bool SURFACE, bool UPDATE, bool WRITE_LDUH, typename TF, bool LINEAR = false>
__global__ __launch_bounds__(node_threads(B)) void update_kernel(UpdateArgs a) {
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B, B3 = B2 * B;
  constexpr int NT = node_threads(B);
  __shared__ double sRed[NT / kWave];
  const size_t cell = (size_t)blockIdx.x + a.cell0;
  const int tid = threadIdx.x;
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const int patch = a.cell_patch ? a.cell_patch[cell] : 0;
  const double* dxp = a.cell_patch ? a.dx + (size_t)patch * 3 : a.dx + cell * a.dx_stride;
  const double invDx[3] = {1.0 / dxp[0], 1.0 / dxp[1], 1.0 / dxp[2]};
  double* luh = a.luh + cell * (size_t)(B3 * D) + node * D;
  double q[D];
#pragma unroll
  for (int m = 0; m < D; ++m) q[m] = luh[m];
  if (SURFACE || UPDATE) {
    const double wx = exa_gl_weights[N][x], wy = exa_gl_weights[N][y], wz = exa_gl_weights[N][z];
    TF* lduhp = static_cast<TF*>(a.lduh) + cell * (size_t)(B3 * V) + node * V;
    double du[V];
#pragma unroll
    for (int m = 0; m < V; ++m) du[m] = lduhp[m];
    if (SURFACE) {
      const TF* lF = static_cast<const TF*>(a.lFhbnd) + cell * (size_t)(6 * B2 * V);
      {  // x faces, (z,y)
        const double us = wz * wy * invDx[0];
        const double usl = (wz * wy) / dxp[0];  // surfaceIntegralLinear.cpph:21-90
        const double fr = exa_fr_coeff[N][x], fl = exa_fl_coeff[N][x];
#pragma unroll
        for (int m = 0; m < V; ++m)
          if constexpr (LINEAR)
            du[m] -= ((double)lF[(1 * B2 + z * B + y) * V + m] * fr + (double)lF[(0 * B2 + z * B + y) * V + m] * fl) * usl;
          else
            du[m] -= ((double)lF[(1 * B2 + z * B + y) * V + m] * fr - (double)lF[(0 * B2 + z * B + y) * V + m] * fl) * us;
      }
      {  // y faces, (z,x)
        const double us = wz * wx * invDx[1];
        const double usl = (wz * wx) / dxp[1];  // surfaceIntegralLinear.cpph:21-90
        const double fr = exa_fr_coeff[N][y], fl = exa_fl_coeff[N][y];
#pragma unroll
        for (int m = 0; m < V; ++m)
          if constexpr (LINEAR)
            du[m] -= ((double)lF[(3 * B2 + z * B + x) * V + m] * fr + (double)lF[(2 * B2 + z * B + x) * V + m] * fl) * usl;
          else
            du[m] -= ((double)lF[(3 * B2 + z * B + x) * V + m] * fr - (double)lF[(2 * B2 + z * B + x) * V + m] * fl) * us;
      }
      {  // z faces, (y,x)
        const double us = wy * wx * invDx[2];
        const double usl = (wy * wx) / dxp[2];  // surfaceIntegralLinear.cpph:21-90
        const double fr = exa_fr_coeff[N][z], fl = exa_fl_coeff[N][z];
#pragma unroll
        for (int m = 0; m < V; ++m)
          if constexpr (LINEAR)
            du[m] -= ((double)lF[(5 * B2 + y * B + x) * V + m] * fr + (double)lF[(4 * B2 + y * B + x) * V + m] * fl) * usl;
          else
            du[m] -= ((double)lF[(5 * B2 + y * B + x) * V + m] * fr - (double)lF[(4 * B2 + y * B + x) * V + m] * fl) * us;
      }
    }
    if (UPDATE) {
      const double dt = a.dt[patch];
      const double updateSize = dt / (wz * wy * wx);
#pragma unroll
      for (int m = 0; m < V; ++m) q[m] += du[m] * updateSize;
      if (active) {
#pragma unroll
        for (int m = 0; m < V; ++m) luh[m] = q[m];
      }
    }
    if (WRITE_LDUH && active) {
#pragma unroll
      for (int m = 0; m < V; ++m) lduhp[m] = TF(du[m]);
    }
  }
  if (a.dt_slots || a.dt_cell) {
    double dtn = active ? node_stable_dt<PDE>(q, invDx, a.cfl_pnpm) : DBL_MAX;
    dtn = dtn < DBL_MAX ? dtn : DBL_MAX;  // a NaN or inf candidate never wins (std::min order)
    dtn = block_min<NT>(dtn, sRed);
    if (tid == 0) {
      if (a.dt_cell) a.dt_cell[cell] = dtn;
      if (a.dt_slots)
        atomic_min_positive(a.dt_slots + (a.cell_patch ? (size_t)patch * a.slots_per_patch + cell % a.slots_per_patch
                                                       : cell % kDtSlots),
                            dtn);
    }
  }
}
// The corrector of one time step in one kernel (corrector_kernel):
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
// Cell of a workgroup: the dispatcher deals workgroups to the 8 XCDs in turn,
// so workgroup b runs on XCD b % 8; XCD x takes cells [x per, (x+1) per).
__device__ inline long xcd_ordered_cell(long ncells) {
  const long per = (ncells + 7) / 8;
  return (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

// A walk of an nx x ny x nz grid that keeps face neighbours close in time:
// blocks of T planes, in them blocks of T rows, in them the x columns one
// after the other, each column's T x T cells together.  A cell's x, y and z
// neighbours inside its block are then 16, 1 and 4 positions away (instead
// of 1, nx and nx ny), so the neighbour's face blocks are still in the L2.
// Partial blocks at the grid's high ends keep the walk a permutation.
__device__ inline long blocked_cell(long idx, int nx, int ny, int nz) {
  constexpr int T = 4;
  const long plane_block = (long)T * ny * nx;
  const int kb = (int)(idx / plane_block);
  long r = idx - kb * plane_block;
  const int kdim = nz - kb * T < T ? nz - kb * T : T;
  const long row_block = (long)T * nx * kdim;
  const int jb = (int)(r / row_block);
  r -= jb * row_block;
  const int jdim = ny - jb * T < T ? ny - jb * T : T;
  const int i = (int)(r / (jdim * kdim));
  const int rest = (int)(r - (long)i * jdim * kdim);
  const int kk = rest / jdim, jj = rest - kk * jdim;
  return ((long)(kb * T + kk) * ny + (jb * T + jj)) * nx + i;
}

// boundaryValues at face node fn of face f of cell (ci, cj, ck), normal
// direction DN (boundaryConditions.cpph:66-83): x on the face, the two free
// coordinates from the node (i, j) along (d2, d1); Fout/Qout hold the inside
// flux and state on entry.
template <int DN, int N, class PDE, typename TF>
__device__ inline void user_boundary_values(const CorrectorArgs& a, int f, int ci, int cj, int ck, int fn,
                                            const TF* Fin, const TF* Qin, TF* Fout, TF* Qout) {
  constexpr int B = N + 1;
  constexpr int d1 = DN == 0 ? 1 : 0, d2 = DN == 2 ? 1 : 2;
  const int idx[3] = {ci, cj, ck};
  double x[3];
  x[DN] = a.lo[DN] + (idx[DN] + 0.5) * a.hdx[DN] + (-0.5 + (f & 1)) * a.hdx[DN];
  x[d2] = a.lo[d2] + (idx[d2] + 0.5) * a.hdx[d2] + a.hdx[d2] * (exa_gl_nodes[N][fn / B] - 0.5);
  x[d1] = a.lo[d1] + (idx[d1] + 0.5) * a.hdx[d1] + a.hdx[d1] * (exa_gl_nodes[N][fn % B] - 0.5);
  PDE::boundary_values(x, a.t[0], a.dt[0], f, DN, Fin, Qin, Fout, Qout);
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
  const long slot = xcd_ordered_cell(ncells);
  if (slot >= ncells) return;  // whole workgroup
  const long cell = a.cell_patch ? slot : blocked_cell(slot, a.nx, a.ny, a.nz);
  const int tid = threadIdx.x;
  // the cell's grid (the whole grid, or its patch of a batch)
  int nx = a.nx, ny = a.ny, nz = a.nz, patch = 0;
  long base = 0;
  if (a.cell_patch) {
    patch = a.cell_patch[cell];
    base = a.patch_first[patch];
    nx = ny = nz = a.patch_n[patch];
  }
  const long local = cell - base;
  const int ci = (int)(local % nx), cj = (int)((local / nx) % ny), ck = (int)(local / ((long)nx * ny));
  // neighbour across face f (-x, +x, -y, +y, -z, +z); -1 at a transmissive boundary
  auto neighbour = [&](int f) -> long {
    int i = ci, j = cj, k = ck;
    const int s = (f & 1) ? 1 : -1;
    if (f < 2) i += s; else if (f < 4) j += s; else k += s;
    if (a.boundary != kBoundaryPeriodic && (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz))
      return -1;
    i = (i + nx) % nx;
    j = (j + ny) % ny;
    k = (k + nz) % nz;
    return base + ((long)k * ny + j) * nx + i;
  };
  const TF* const Qb = static_cast<const TF*>(a.lQhbnd);
  const TF* const Fb = static_cast<const TF*>(a.lFhbnd);
  // the node's solution and volume update: a large node state (elastic, 12
  // quantities) is loaded before the face phase, so that its latency overlaps
  // that phase (5.67 -> 4.97 ms per launch at 81^3, N=5); a small one (Euler,
  // 5) after it, where the early load's registers would cost an occupancy
  // step for no gain (3.69 -> 3.75 ms)
  constexpr bool kEarlyNodeLoads = D > 8;
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  double* luh = a.luh + (size_t)cell * (B3 * D) + node * D;
  const TF* lduhp = static_cast<const TF*>(a.lduh) + (size_t)cell * (B3 * V) + node * V;
  double q[D], du[V];
  auto load_node = [&] {
#pragma unroll
    for (int m = 0; m < D; ++m) q[m] = luh[m];
#pragma unroll
    for (int m = 0; m < V; ++m) du[m] = lduhp[m];
  };
  if constexpr (kEarlyNodeLoads) load_node();

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
    // the outside state and flux in registers (at a boundary: the inside ones)
    TF qnv[D], fnv[V];
#pragma unroll
    for (int k = 0; k < D; ++k) qnv[k] = qn[k];
#pragma unroll
    for (int k = 0; k < V; ++k) fnv[k] = fnb[k];
    if constexpr (HasBoundaryValues<PDE, TF>) {
      if (nb < 0 && a.boundary == kBoundaryUser) {
        // one call site per normal direction, so that the hook's indexing by
        // normalNonZero is by a constant and its arrays stay in registers
        switch (f >> 1) {
          case 0: user_boundary_values<0, N, PDE>(a, f, ci, cj, ck, fn, fo, qo, fnv, qnv); break;
          case 1: user_boundary_values<1, N, PDE>(a, f, ci, cj, ck, fn, fo, qo, fnv, qnv); break;
          default: user_boundary_values<2, N, PDE>(a, f, ci, cj, ck, fn, fo, qo, fnv, qnv); break;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const TF o = qo[k], n = qnv[k];
      qL[r][k] = own_left ? o : n;
      qR[r][k] = own_left ? n : o;
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const TF o = fo[k], n = fnv[k];
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
    const TF sL = pde_max_abs_eigenvalue<PDE>(avL, d);
    const TF sR = pde_max_abs_eigenvalue<PDE>(avR, d);
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
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const double* const dxp = a.dx + 3 * patch;
  const double invDx[3] = {1.0 / dxp[0], 1.0 / dxp[1], 1.0 / dxp[2]};
  const double dt = a.dt[patch];
  if constexpr (!kEarlyNodeLoads) load_node();
  const double wx = exa_gl_weights[N][x], wy = exa_gl_weights[N][y], wz = exa_gl_weights[N][z];
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
  const double updateSize = dt / (wz * wy * wx);
#pragma unroll
  for (int m = 0; m < V; ++m) q[m] += du[m] * updateSize;
  if constexpr (HasAdjustPointSolution<PDE>) {
    if (a.adjust) {
      // adjustSolution(t + dt, dt) after the update (ADERDGSolver.cpp:2231-2236),
      // node coordinates of solutionAdjustment.cpph:32-43
      const double xv[3] = {a.lo[0] + (ci + 0.5) * a.hdx[0] + a.hdx[0] * (exa_gl_nodes[N][x] - 0.5),
                            a.lo[1] + (cj + 0.5) * a.hdx[1] + a.hdx[1] * (exa_gl_nodes[N][y] - 0.5),
                            a.lo[2] + (ck + 0.5) * a.hdx[2] + a.hdx[2] * (exa_gl_nodes[N][z] - 0.5)};
      PDE::adjust_point_solution(xv, a.t[0] + a.dt[0], a.dt[0], q);
      if (active) {
#pragma unroll
        for (int m = V; m < D; ++m) luh[m] = q[m];
      }
    }
  }
  if (active) {
#pragma unroll
    for (int m = 0; m < V; ++m) luh[m] = q[m];
  }
  double dtn = active ? node_stable_dt<PDE>(q, invDx, a.cfl_pnpm) : DBL_MAX;
  dtn = dtn < DBL_MAX ? dtn : DBL_MAX;  // a NaN or inf candidate never wins (std::min order)
  dtn = block_min<NT>(dtn, sRed);
  if (tid == 0)
    atomic_min_positive(a.dt_slots + (a.cell_patch ? (size_t)patch * a.slots_per_patch + cell % a.slots_per_patch
                                                   : (size_t)(cell % kDtSlots)),
                        dtn);
}

// solutionAdjustment (solutionAdjustment.cpph:17-47) over a grid: one
// workgroup per cell, one thread per node, x = centre + dx (xi - 1/2).
template <int B, class PDE>
__global__ __launch_bounds__(node_threads(B)) void adjust_kernel(AdjustArgs a) {
  constexpr int D = PDE::V + PDE::P, N = B - 1, B2 = B * B, B3 = B2 * B;
  const long cell = blockIdx.x;
  const int node = threadIdx.x;
  if (node >= B3) return;
  if constexpr (HasAdjustPointSolution<PDE>) {
    const int z = node / B2, y = (node / B) % B, x = node % B;
    const int ci = (int)(cell % a.nx), cj = (int)((cell / a.nx) % a.ny), ck = (int)(cell / ((long)a.nx * a.ny));
    const double xv[3] = {a.lo[0] + (ci + 0.5) * a.dx[0] + a.dx[0] * (exa_gl_nodes[N][x] - 0.5),
                          a.lo[1] + (cj + 0.5) * a.dx[1] + a.dx[1] * (exa_gl_nodes[N][y] - 0.5),
                          a.lo[2] + (ck + 0.5) * a.dx[2] + a.dx[2] * (exa_gl_nodes[N][z] - 0.5)};
    double q[D];
    double* luh = a.luh + (size_t)cell * (B3 * D) + node * D;
#pragma unroll
    for (int m = 0; m < D; ++m) q[m] = luh[m];
    PDE::adjust_point_solution(xv, a.t[0], a.dt, q);
#pragma unroll
    for (int m = 0; m < D; ++m) luh[m] = q[m];
  }
}

}  // namespace exa
