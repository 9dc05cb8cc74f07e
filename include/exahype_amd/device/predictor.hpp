// predictor.hpp -- the space-time predictor kernel (Picard iteration, time
// averaging, extrapolation, volume integral) as a template over the PDE
// functor, for the library's own PDEs and for a user's __device__ PDE
// (exahype_amd/user_solver.hpp).
//
// Reference semantics (K = ExaHyPE/kernels/aderdg/generic/c/3d):
//   predictor_kernel   K/spaceTimePredictorNonlinear.cpph:42-528,661-717
//                      + K/volumeIntegralNonlinear.cpph:136-225 (fused epilogue)
#pragma once

#include "aderdg.hpp"
#include "exa_device.hpp"

#include <cfloat>
#include <type_traits>

namespace exa {

// Picard stopping rule (spaceTimePredictorNonlinear.cpph:314-318): stop when
// sum (lQi_new - lQi)^2 < 1e-14.  In float that absolute bound lies below the
// rounding floor of the iterate (about (6e-8 |q|)^2 per entry), so the
// FP32-bulk form stops when the root-mean-square change falls below
// kPicardRelTol32 times the root-mean-square of the iterate (or below the
// reference's bound, whichever is reached first).
constexpr double kPicardTol2 = 1e-7 * 1e-7;
constexpr float kPicardRelTol32 = 1e-6f;

template <typename T>
__device__ constexpr T tc(double v) {
  return static_cast<T>(v);
}

// S = algebraicSource(Q) - NCP(Q, gradQ) at node (z,y,x) of a time slice whose
// variables are staged in sQ[var][node] (spaceTimePredictorNonlinear.cpph:238-266).
template <int B, class PDE, bool USE_SOURCE, bool USE_NCP, typename T>
__device__ inline void source_ncp(const T* Q, const T* sQ, const T* sDu, const T invDx[3], int z, int y,
                                  int x, T* S) {
  constexpr int V = PDE::V;
  constexpr int B2 = B * B, B3 = B2 * B;
  if (USE_SOURCE) {
    PDE::algebraic_source(Q, S);
  } else {
#pragma unroll
    for (int m = 0; m < V; ++m) S[m] = T(0);
  }
  if (USE_NCP) {
    // gradQ(dir, var) = sum_n invDx * lQi(..n..) * dudx[l][n] (:170-173, 199-202, 228-231)
    T gradQ[3][V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
      T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
      for (int n = 0; n < B; ++n) {
        gx += T(1) * invDx[0] * sQ[m * B3 + (z * B + y) * B + n] * sDu[x * B + n];
        gy += T(1) * invDx[1] * sQ[m * B3 + (z * B + n) * B + x] * sDu[y * B + n];
        gz += T(1) * invDx[2] * sQ[m * B3 + (n * B + y) * B + x] * sDu[z * B + n];
      }
      gradQ[0][m] = gx;
      gradQ[1][m] = gy;
      gradQ[2][m] = gz;
    }
    T ncp[V];
    PDE::ncp(Q, gradQ, ncp);
#pragma unroll
    for (int m = 0; m < V; ++m) S[m] -= ncp[m];
  }
}

// Where the predictor keeps the Picard iterate lQi^k of a cell (B^4 V values):
// its first NQL time slices live in LDS (as many as keep the workgroup's LDS
// within 144 KB next to one slice of fluxes sF), the other NQR slices in
// registers or, when registers run out (WSQ), in a per-cell device workspace
// that each thread reads and writes only at its own nodes.  Large sizes run
// the loop over time slices rolled (ROLLED).
template <int B, int V, int D, bool USE_NCP, typename T, int LDSK = 144>
struct PredictorLayout {
  static constexpr int B2 = B * B, B3 = B2 * B;
  static constexpr int NT = node_threads(B);
  static constexpr bool IS_FP64 = sizeof(T) == sizeof(double);
  // LDS: one slice of fluxes [dir][var][node] (reused for lQhi / lFhi staging)
  static constexpr int SF = (3 * V > D ? 3 * V : D) * B3;
  static constexpr long kLdsBudget = (long)LDSK * 1024;
  static constexpr long kSliceBytes = (long)V * B3 * sizeof(T);
  static constexpr long kOther =
      (long)(SF + (USE_NCP ? V * B3 : 0) + 2 * B2 + B + 3 + 2 * (NT / kWave)) * sizeof(T);
  static constexpr long kFree = kLdsBudget - kOther;
  static constexpr int NQL = kFree <= 0 ? 0 : (kFree / kSliceBytes >= B ? B : (int)(kFree / kSliceBytes));
  static constexpr int NQR = B - NQL;
  // Rolled slices: the large sizes and the 9-variable PDE, whose unrolled
  // slices exhaust the registers, and the float form at N=4 and N=5, where the
  // rolled loop needs fewer registers (N=5: 110 instead of 166, one more
  // wavefront per SIMD) at the cost of uniform table loads per slice
  // (measured: N=4 -2%, N=5 -4% predictor time; at N=6 the rolled form was
  // 19% slower, so N=6 stays unrolled), and the double form of the
  // 5-variable PDE at N=3, where the rolled loop leaves room for the
  // accumulated time averages (the predictor kernel's ACC_FH).
  static constexpr bool ROLLED = (B >= 7 && (IS_FP64 || B >= 8)) || V >= 9 || (!IS_FP64 && (B == 5 || B == 6)) ||
                                 (IS_FP64 && B == 4 && V == 5);
  // 8+ waves per workgroup leave 256 registers per lane, no room for slices;
  // a rolled loop cannot index registers.
  static constexpr bool WSQ = NQR > 0 && (NT >= 8 * kWave || ROLLED);
  // unroll factor of the loop over time slices
  static constexpr int UNROLL_T = ROLLED ? 1 : B;
  static constexpr bool SPLIT = ROLLED && WSQ && NQL > 0;
  static constexpr size_t ws_bytes = WSQ ? (size_t)NQR * V * B3 * sizeof(T) : 0;
};

// ---------------------------------------------------------------------------
// Space-time predictor + time averaging + extrapolation + volume integral.
//
// Per Picard iteration and time slice t, three phases separated by barriers:
//  1. node owners evaluate the fluxes of their node and stage them in LDS as
//     sF[dir][var][node];
//  2. line owners (one task = one direction, variable and grid line) contract
//     the line with Kxi in place -- every staged value is read once, Kxi comes
//     in as scalar constants;
//  3. node owners form rhs(t) from the directional contributions and fold it
//     into the next iterate through iK1 (the reference's n order), so rhs is
//     never stored for all t.
// ---------------------------------------------------------------------------
// Waves per SIMD the register allocation must leave room for: the float
// flux-only form at N=5 and N=7 accumulates its time-averaged fluxes during
// the last iteration (3V more live registers) within the four (N=5) or three
// (N=7, 64 KB layout) waves that its LDS allows anyway; elsewhere no
// constraint.
template <int B, int TERMS, typename T>
__host__ __device__ constexpr int predictor_min_waves() {
  return (sizeof(T) == sizeof(float) && TERMS == kFlux) ? (B == 6 ? 4 : (B == 8 ? 3 : 1)) : 1;
}

template <int B, class PDE, int TERMS, bool EXPORT, typename T, int LDSK = 144>
__global__ __launch_bounds__(node_threads(B), (predictor_min_waves<B, TERMS, T>())) void predictor_kernel(
    PredictorArgs a) {
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B, B3 = B2 * B;
  constexpr int NT = node_threads(B);
  constexpr bool USE_FLUX = (TERMS & kFlux) != 0;
  constexpr bool USE_SOURCE = (TERMS & kSource) != 0;
  constexpr bool USE_NCP = (TERMS & kNCP) != 0;
  constexpr bool USE_SRC_NCP = USE_SOURCE || USE_NCP;
  constexpr bool IS_FP64 = sizeof(T) == sizeof(double);
  constexpr int MAX_ITER = 2 * B;  // spaceTimePredictorNonlinear.cpph:85
  using L = PredictorLayout<B, V, D, USE_NCP, T, LDSK>;
  constexpr int SF = L::SF, NQL = L::NQL, NQR = L::NQR, UNROLL_T = L::UNROLL_T;
  constexpr bool SPLIT = L::SPLIT;
  constexpr bool WSQ = L::WSQ;
  __shared__ T sF[SF];
  __shared__ T sQi[NQL > 0 ? NQL * V * B3 : 1];  // lQi^k slices t < NQL, [t][var][node]
  __shared__ T sQ[USE_NCP ? V * B3 : 1];
  __shared__ T sK[B2];   // Kxi[n][l]
  __shared__ T sDu[B2];  // dudx[l][n]
  __shared__ T sW[B];    // Gauss-Legendre weights
  __shared__ T sRed[2 * (NT / kWave)];

  T* const out_lduh = static_cast<T*>(a.lduh);
  T* const out_lQhbnd = static_cast<T*>(a.lQhbnd);
  T* const out_lFhbnd = static_cast<T*>(a.lFhbnd);

  // Lanes beyond the B^3 nodes compute node 0 again and write the same values,
  // and loops over outputs run whole rounds with the index clamped (surplus
  // lanes redo the last item), so control flow stays uniform; only the in-place
  // line contraction guards its stores.  (No instantiation may spill: the build
  // rejects any kernel with scratch memory.)
  const size_t cell = (size_t)blockIdx.x + a.cell0;
  const int tid = threadIdx.x;
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  if (a.run_if && *a.run_if == 0) return;  // whole launch
  if (a.status && a.status[cell] >= a.status_limit) {  // whole workgroup
    if (a.iters && tid == 0) a.iters[cell] = 0;
    return;
  }

  for (int i = tid; i < B2; i += NT) {
    sK[i] = tc<T>(exa_kxi[N][i / B][i % B]);
    sDu[i] = tc<T>(exa_dudx[N][i / B][i % B]);
  }
  for (int i = tid; i < B; i += NT) sW[i] = tc<T>(exa_gl_weights[N][i]);

  const int patch = a.cell_patch ? a.cell_patch[cell] : 0;
  const double* dxp = a.cell_patch ? a.dx + (size_t)patch * 3 : a.dx + cell * a.dx_stride;
  const T invDx[3] = {T(1.0) / T(dxp[0]), T(1.0) / T(dxp[1]), T(1.0) / T(dxp[2])};
  const T dt = T(a.dt[patch]);

  T q0[D];
  {
    const double* luh = a.luh + cell * (size_t)(B3 * D) + node * D;
#pragma unroll
    for (int m = 0; m < D; ++m) q0[m] = T(luh[m]);
  }
  const T wx = tc<T>(exa_gl_weights[N][x]), wy = tc<T>(exa_gl_weights[N][y]), wz = tc<T>(exa_gl_weights[N][z]);
  const T w3 = wz * wy * wx;  // weight of rhs0 (:130-132) and of iK1 (:286-288)

  // Picard state: the current iterate lQi^k of this node (in LDS, each thread
  // touching only its own node's entries so no barriers are needed, or in
  // registers); the next iterate qn = lQi^{k+1} is accumulated in registers
  // slice by slice (iK1 applied in the reference's n order).
  T qi_reg[(NQR > 0 && !WSQ) ? NQR : 1][V];
  T* const ws = WSQ ? static_cast<T*>(a.ws) + (size_t)blockIdx.x * (NQR * V * B3) : nullptr;
  // Register slices are only used with compile-time t (unrolled loops).  Loads
  // and stores go through separate accessors with a uniform branch on t, so no
  // pointer that may address either LDS or global memory is ever formed.
  auto qi = [&](int t, int m) __device__ -> T {
    if (t < NQL)
      return sQi[(t * V + m) * B3 + node];
    else if (WSQ)
      return ws[((t - NQL) * V + m) * B3 + node];
    else
      return qi_reg[t - NQL][m];
  };
  auto qi_set = [&](int t, int m, T v) __device__ {
    if (t < NQL)
      sQi[(t * V + m) * B3 + node] = v;
    else if (WSQ)
      ws[((t - NQL) * V + m) * B3 + node] = v;
    else
      qi_reg[t - NQL][m] = v;
  };
#pragma unroll
  for (int t = 0; t < B; ++t)
#pragma unroll
    for (int m = 0; m < V; ++m)
      qi_set(t, m, q0[m]);
  T qn[B][V];
  int iters = MAX_ITER;
  // Float form of a flux-only PDE: the first Picard iteration in closed form
  // and the time-averaged fluxes accumulated while the iterations run (see
  // below); the double form keeps the reference's order of operations.
  constexpr bool FLOAT_FLUX_ONLY = !IS_FP64 && USE_FLUX && !USE_SRC_NCP;
  // accumulating the averages needs 3V more registers: only at B = 5, 6 and
  // 8, where they fit the wavefronts per SIMD that LDS allows (B = 6 and 8
  // with the register cap of predictor_min_waves: 124 and 166 registers; the
  // epilogue's B flux evaluations go, 4% of the predictor at 81^3 N=5; B = 3
  // and 4 would drop a wavefront per SIMD, B = 7 spills under the cap), and
  // the double form of the 5-variable PDE at B = 4 with its slice loop rolled
  // (170 registers, the two wavefronts per SIMD its LDS allows; 3.44 -> 3.36
  // ms at 81^3 in the limiter workload, where one iteration is the rule)
  constexpr bool ACC_FH = (FLOAT_FLUX_ONLY && (B == 5 || B == 6 || B == 8)) ||
                          (IS_FP64 && USE_FLUX && !USE_SRC_NCP && B == 4 && V == 5);
  T fh[3][V];
  // Float form of a PDE without fluxes (source and/or NCP only: the elastic
  // PDE): the time-averaged source term likewise accumulates during the
  // iterations, from the S each slice's phase 3 evaluates anyway, so the
  // epilogue re-evaluates no slice (V more registers; the elastic kernels keep
  // their occupancy).
  constexpr bool ACC_SH = USE_SRC_NCP && !USE_FLUX;
  T sh[V];
  const T iweight = T(1.0) / w3;
  __syncthreads();

  // Float form: the quadrature weights and dt/dx of the line contraction are
  // applied per node in phase 3 instead of per line value in phase 2:
  // us_d / w^3 = w_t dt / (dx_d w_d), w_d the node's weight along d.
  const T g_node[3] = {dt * invDx[0] / wx, dt * invDx[1] / wy, dt * invDx[2] / wz};
  // Phase 2 tasks per direction: one (variable m, line) each, NTD of them in
  // RD rounds over the workgroup.
  constexpr int NTD = B2 * V;
  constexpr int RD = (NTD + NT - 1) / NT;

  // Phase 2 of a time slice (line owners): per direction d, one task = one
  // (variable, line); out(l) = sum_n (us * F(n)) * Kxi[n][l] along the line,
  // with us = w_t w_a w_b dt / dx_d (:152-236) -- in the float form us is
  // applied per node in phase 3 instead.  Directions run one after the other
  // with compile-time strides; Kxi comes in as constants.  A task reads its
  // whole line before writing the results back to the same positions, and no
  // other task touches them: the contraction is in place.  Ends with a barrier.
  auto contract_lines = [&](const T wt) __device__ __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      if (tid + r * NT < NTD) {
        const int task = tid + r * NT;
        const int m = task / B2, line = task - (task / B2) * B2;
        const int la = line / B, lb = line - (line / B) * B;
        // line starts: x-lines (z,y) stride 1, y-lines (z,x) stride B,
        // z-lines (y,x) stride B^2; the weights are w_la, w_lb
        const int off0 = m * B3 + line * B, off1 = m * B3 + la * B2 + lb, off2 = m * B3 + line;
        const T wa = sW[la], wb = sW[lb];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int stride = d == 0 ? 1 : (d == 1 ? B : B2);
          const int o = d * V * B3 + (d == 0 ? off0 : (d == 1 ? off1 : off2));
          T f[B];
          if constexpr (IS_FP64) {
            const T us = wt * wa * wb * dt * invDx[d];
#pragma unroll
            for (int n = 0; n < B; ++n) f[n] = us * sF[o + n * stride];
          } else {
#pragma unroll
            for (int n = 0; n < B; ++n) f[n] = sF[o + n * stride];
          }
#pragma unroll
          for (int l = 0; l < B; ++l) {
            T acc = T(0);
#pragma unroll
            for (int n = 0; n < B; ++n) acc += f[n] * tc<T>(exa_kxi[N][n][l]);
            sF[o + l * stride] = acc;
          }
        }
      }
    }
    __syncthreads();
  };

#pragma unroll 1
  for (int iter = 0; iter < MAX_ITER; ++iter) {
    // rhs0 = w^3 FLCoeff[t] luh is loop invariant; keeping B*V copies of it live
    // across iterations costs more registers than recomputing it per slice.
    T w3i = w3;
    asm volatile("" : "+v"(w3i));
    T q0i[V];  // the same for the float form's FLCoeff[t] luh
#pragma unroll
    for (int m = 0; m < V; ++m) {
      q0i[m] = q0[m];
      asm volatile("" : "+v"(q0i[m]));
    }
#pragma unroll
    for (int l = 0; l < B; ++l)
#pragma unroll
      for (int m = 0; m < V; ++m) qn[l][m] = T(0);
    if constexpr (ACC_FH) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < V; ++m) fh[d][m] = T(0);
    }
    if constexpr (ACC_SH) {
#pragma unroll
      for (int m = 0; m < V; ++m) sh[m] = T(0);
    }
    // Rolled (large sizes): table entries come from constant memory with a
    // uniform index; unrolled: they are immediates.
    // One time slice; with SPLIT (rolled loops over slices that live partly
    // in LDS and partly in the workspace) the LDS slices and the workspace
    // slices run in two loops, so that within each loop a slice's location is
    // known and every access is a plain LDS or global one (never a generic
    // pointer that may address either).
    auto picard_slice = [&](const int t) __device__ __attribute__((always_inline)) {
      const T wt = tc<T>(exa_gl_weights[N][t]);
      T Q[D];
#pragma unroll
      for (int m = 0; m < V; ++m) Q[m] = qi(t, m);
#pragma unroll
      for (int m = V; m < D; ++m) Q[m] = q0[m];
      // Float form, first iteration: every slice of the initial iterate equals
      // luh, and phase 2 does not depend on t there, so the fluxes and line
      // contractions of slice 0 (still in LDS) serve every slice (the same
      // inputs give the same values).  The double form keeps w_t inside the
      // contraction, as the reference, and evaluates every slice.
      const bool reuse = !IS_FP64 && iter == 0 && t > 0;
      if (!reuse) {
      // phase 1 (node owners): fluxes of this time slice -> sF[dir][var][node]
      if (USE_FLUX) {
        T F[3][V];
        PDE::flux(Q, F);
        if constexpr (ACC_FH) {
          // time average of the slice fluxes of this iterate (:367-396)
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < V; ++m) fh[d][m] += F[d][m] * wt;
        }
        {
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < V; ++m) sF[(d * V + m) * B3 + node] = F[d][m];
          if (EXPORT) {
            double* lFi = a.lFi + cell * (size_t)(4 * B3 * B * V) + (((size_t)t * B3 + node) * 4) * V;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
              for (int m = 0; m < V; ++m) lFi[d * V + m] = F[d][m];
          }
        }
      }
      if (USE_NCP) {
#pragma unroll
        for (int m = 0; m < V; ++m) sQ[m * B3 + node] = Q[m];
      }
      __syncthreads();

      if (USE_FLUX) contract_lines(wt);
      }  // !reuse

      // phase 3 (node owners): rhs = w^3 FLCoeff[t] luh - contributions (+ S dt w^4)
      // (:124-141, :238-276); lQi_new(l) += iweight * rhs * iK1[l][t] (:292-299, n = t)
      const T fl_t = tc<T>(exa_fl_coeff[N][t]);
      if constexpr (!IS_FP64) {
        // float form: ir = rhs / w^3 directly, weights folded per node
        T ir[V];
        const T wg[3] = {wt * g_node[0], wt * g_node[1], wt * g_node[2]};
#pragma unroll
        for (int m = 0; m < V; ++m) {
          ir[m] = fl_t * q0i[m];
          if (USE_FLUX) {
            ir[m] -= wg[0] * sF[(0 * V + m) * B3 + node];
            ir[m] -= wg[1] * sF[(1 * V + m) * B3 + node];
            ir[m] -= wg[2] * sF[(2 * V + m) * B3 + node];
          }
        }
        if (USE_SRC_NCP) {
          T S[V];
          source_ncp<B, PDE, USE_SOURCE, USE_NCP>(Q, sQ, sDu, invDx, z, y, x, S);
          const T us = wt * dt;
#pragma unroll
          for (int m = 0; m < V; ++m) ir[m] += us * S[m];
          if constexpr (ACC_SH) {
            // time average of the slice sources of this iterate (:367-396)
#pragma unroll
            for (int m = 0; m < V; ++m) sh[m] += S[m] * wt;
          }
        }
#pragma unroll
        for (int m = 0; m < V; ++m)
#pragma unroll
          for (int l = 0; l < B; ++l) qn[l][m] += ir[m] * tc<T>(exa_ik1[N][l][t]);
        if (USE_FLUX || USE_NCP) __syncthreads();  // the next slice rewrites sF / sQ
        return;
      }
      T r[V];
#pragma unroll
      for (int m = 0; m < V; ++m) {
        r[m] = w3i * fl_t * q0[m];
        if (USE_FLUX) {
          r[m] -= sF[(0 * V + m) * B3 + node];
          r[m] -= sF[(1 * V + m) * B3 + node];
          r[m] -= sF[(2 * V + m) * B3 + node];
        }
      }
      if (USE_SRC_NCP) {
        T S[V];
        source_ncp<B, PDE, USE_SOURCE, USE_NCP>(Q, sQ, sDu, invDx, z, y, x, S);
        const T us = wt * wz * wy * wx * dt;
#pragma unroll
        for (int m = 0; m < V; ++m) r[m] += us * S[m];
        if constexpr (ACC_SH) {
#pragma unroll
          for (int m = 0; m < V; ++m) sh[m] += S[m] * wt;
        }
        if (EXPORT) {
          double* lFi = a.lFi + cell * (size_t)(4 * B3 * B * V) + (((size_t)t * B3 + node) * 4 + 3) * V;
#pragma unroll
          for (int m = 0; m < V; ++m) lFi[m] = S[m];
        }
      }
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const T ir = iweight * r[m];
#pragma unroll
        for (int l = 0; l < B; ++l) qn[l][m] += ir * tc<T>(exa_ik1[N][l][t]);
      }
      if (USE_FLUX || USE_NCP) __syncthreads();  // the next slice rewrites sF / sQ
    };
    if (FLOAT_FLUX_ONLY && iter == 0) {
      // First iteration of the float form: every slice of the initial
      // iterate equals luh, so the slices share one flux contraction S and
      // lQi^1(l) = sum_t iK1[l][t] (FLCoeff[t] luh - w_t S)
      //          = (sum_t iK1[l][t] FLCoeff[t]) luh - (sum_t iK1[l][t] w_t) S,
      // coefficients formed in double from the tables.
      T F[3][V];
      {
        T Q[D];
#pragma unroll
        for (int m = 0; m < D; ++m) Q[m] = m < V ? q0i[m < V ? m : 0] : q0[m];
        PDE::flux(Q, F);
      }
      T sumw = T(0);
      {
        double sw = 0.0;
#pragma unroll
        for (int t = 0; t < B; ++t) sw += exa_gl_weights[N][t];
        sumw = tc<T>(sw);
      }
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < V; ++m) {
          sF[(d * V + m) * B3 + node] = F[d][m];
          if constexpr (ACC_FH) fh[d][m] = F[d][m] * sumw;
        }
      __syncthreads();
      contract_lines(T(0));
      T S[V];
#pragma unroll
      for (int m = 0; m < V; ++m)
        S[m] = g_node[0] * sF[(0 * V + m) * B3 + node] + g_node[1] * sF[(1 * V + m) * B3 + node] +
               g_node[2] * sF[(2 * V + m) * B3 + node];
#pragma unroll
      for (int l = 0; l < B; ++l) {
        double al = 0.0, bl = 0.0;
#pragma unroll
        for (int t = 0; t < B; ++t) {
          al += exa_ik1[N][l][t] * exa_fl_coeff[N][t];
          bl += exa_ik1[N][l][t] * exa_gl_weights[N][t];
        }
#pragma unroll
        for (int m = 0; m < V; ++m) qn[l][m] = tc<T>(al) * q0i[m] - tc<T>(bl) * S[m];
      }
      __syncthreads();  // the next iteration rewrites sF
    } else if constexpr (SPLIT) {
#pragma unroll 1
      for (int t = 0; t < NQL; ++t) picard_slice(t);
#pragma unroll 1
      for (int t = NQL; t < B; ++t) picard_slice(t);
    } else {
#pragma unroll UNROLL_T
      for (int t = 0; t < B; ++t) picard_slice(t);
    }

    // squared residual (:303) and exit test (:314-318)
    T sq = T(0), qq = T(0);
#pragma unroll
    for (int l = 0; l < B; ++l)
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const T dq = qn[l][m] - qi(l, m);
        sq += dq * dq;
        if (!IS_FP64) qq += qn[l][m] * qn[l][m];
      }
    if (!active) {
      sq = T(0);
      qq = T(0);
    }
    bool done;
    if (IS_FP64) {
      const T sq_res = block_sum<NT>(sq, sRed);
      done = sq_res < T(kPicardTol2);
    } else {
      const T sq_res = block_sum<NT>(sq, sRed);
      const T qq_sum = block_sum<NT>(qq, sRed + NT / kWave);
      done = sq_res < T(kPicardTol2) || sq_res < kPicardRelTol32 * kPicardRelTol32 * qq_sum;
    }
    if (done || iter + 1 == MAX_ITER) {
      iters = iter + 1;
      break;  // qi still holds lQi^k, the iterate the last fluxes were taken from
    }
#pragma unroll
    for (int l = 0; l < B; ++l)
#pragma unroll
      for (int m = 0; m < V; ++m)
        qi_set(l, m, qn[l][m]);
  }

  // ---- epilogue: time average, extrapolation, volume integral ----
  const size_t cQb = cell * (size_t)(6 * B2 * D);
  const size_t cFb = cell * (size_t)(6 * B2 * V);
  {
    // lQhi = sum_t lQi(t) w_t of the final iterate lQi^{k+1} = qn (:397-404)
    T qh[D];
#pragma unroll
    for (int m = 0; m < D; ++m) {
      T acc = T(0);
#pragma unroll
      for (int t = 0; t < B; ++t) acc += (m < V ? qn[t][m < V ? m : 0] : q0[m]) * tc<T>(exa_gl_weights[N][t]);
      qh[m] = acc;
    }
    if (EXPORT) {
      double* lQi = a.lQi + cell * (size_t)(B3 * B * D) + (size_t)node * B * D;
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int m = 0; m < D; ++m) lQi[t * D + m] = (m < V) ? qn[t][m < V ? m : 0] : q0[m];
      double* lQhi = a.lQhi + cell * (size_t)(B3 * D) + (size_t)node * D;
#pragma unroll
      for (int m = 0; m < D; ++m) lQhi[m] = qh[m];
    }
#pragma unroll
    for (int m = 0; m < D; ++m) sF[m * B3 + node] = qh[m];
  }
  __syncthreads();
  // Face extrapolation (lQhbnd / lFhbnd, :434-527) as line tasks: per
  // direction d, one task = (face node (la, lb), datum k), k fastest so that
  // the stores of a wavefront are contiguous; a task reads its normal line
  // once and forms both faces of the direction (FL and FR coefficients, as
  // immediates).  Source values: sF[k][node] for every direction, or with
  // PER_DIR the direction's own block sF[d][k][node].
  auto extrapolate = [&]<int NK, bool PER_DIR>(T* out, size_t coff) __device__ {
    constexpr int NTK = B2 * NK;
#pragma unroll
    for (int r = 0; r < (NTK + NT - 1) / NT; ++r) {
      const int task = tid + r * NT;
      if (task < NTK) {
        const int k = task % NK, line = task / NK;
        const int la = line / B, lb = line - (line / B) * B;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int stride = d == 0 ? 1 : (d == 1 ? B : B2);
          const int o = (PER_DIR ? d * V + k : k) * B3 + (d == 0 ? line * B : (d == 1 ? la * B2 + lb : line));
          T accL = T(0), accR = T(0);
#pragma unroll
          for (int l = 0; l < B; ++l) {
            const T v = sF[o + l * stride];
            accL += v * tc<T>(exa_fl_coeff[N][l]);
            accR += v * tc<T>(exa_fr_coeff[N][l]);
          }
          out[coff + (size_t)(2 * d) * NTK + task] = accL;
          out[coff + (size_t)(2 * d + 1) * NTK + task] = accR;
        }
      }
    }
  };
  extrapolate.template operator()<D, false>(out_lQhbnd, cQb);
  __syncthreads();

  // Time-averaged fluxes of the last flux evaluation, i.e. of lQi^k = qi (the
  // reference computes lFi before the final lQi update, :96-122 vs :306):
  // lFhi = sum_t w_t F(lQi^k(t)), lShi likewise (:367-396).
#pragma unroll
  for (int m = 0; m < V; ++m) {
    if constexpr (!ACC_FH) {
      fh[0][m] = T(0);
      fh[1][m] = T(0);
      fh[2][m] = T(0);
    }
    if constexpr (!ACC_SH) sh[m] = T(0);
  }
  auto average_slice = [&](const int t) __device__ __attribute__((always_inline)) {
    T Q[D];
#pragma unroll
    for (int m = 0; m < V; ++m) Q[m] = qi(t, m);
#pragma unroll
    for (int m = V; m < D; ++m) Q[m] = q0[m];
    const T wt = tc<T>(exa_gl_weights[N][t]);
    if (USE_FLUX) {
      T F[3][V];
      PDE::flux(Q, F);
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < V; ++m) fh[d][m] += F[d][m] * wt;
    }
    if (USE_SRC_NCP) {
      if (USE_NCP) {
#pragma unroll
        for (int m = 0; m < V; ++m) sQ[m * B3 + node] = Q[m];
      }
      if (USE_NCP) __syncthreads();
      T S[V];
      source_ncp<B, PDE, USE_SOURCE, USE_NCP>(Q, sQ, sDu, invDx, z, y, x, S);
#pragma unroll
      for (int m = 0; m < V; ++m) sh[m] += S[m] * wt;
      if (USE_NCP) __syncthreads();
    }
  };
  if constexpr (ACC_FH || ACC_SH) {
    // fh (flux-only) or sh (no flux) holds the time average of the last
    // iteration's fluxes or sources already
  } else if constexpr (SPLIT) {
#pragma unroll 1
    for (int t = 0; t < NQL; ++t) average_slice(t);
#pragma unroll 1
    for (int t = NQL; t < B; ++t) average_slice(t);
  } else {
#pragma unroll UNROLL_T
    for (int t = 0; t < B; ++t) average_slice(t);
  }
  if (EXPORT) {
    // lFhi blocks: x (z,y,x), y (z,x,y), z (y,x,z), src (z,y,x)  (:375-393)
    double* lFhi = a.lFhi + cell * (size_t)(4 * B3 * V);
#pragma unroll
    for (int m = 0; m < V; ++m) {
      lFhi[0 * B3 * V + ((z * B + y) * B + x) * V + m] = fh[0][m];
      lFhi[1 * B3 * V + ((z * B + x) * B + y) * V + m] = fh[1][m];
      lFhi[2 * B3 * V + ((y * B + x) * B + z) * V + m] = fh[2][m];
      lFhi[3 * B3 * V + ((z * B + y) * B + x) * V + m] = sh[m];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int m = 0; m < V; ++m) sF[(d * V + m) * B3 + node] = fh[d][m];
  __syncthreads();
  // lFhbnd from lFhi_{x,y,z} (:434-527)
  if (USE_FLUX) {
    extrapolate.template operator()<V, true>(out_lFhbnd, cFb);
  } else {
#pragma unroll 1
    for (int round = 0; round < (6 * B2 * V + NT - 1) / NT; ++round)
      out_lFhbnd[cFb + min(tid + round * NT, 6 * B2 * V - 1)] = T(0);
  }
  // volume integral (volumeIntegralNonlinear.cpph:136-225), time-averaged form
  if (out_lduh) {
    T acc[V];
#pragma unroll
    for (int m = 0; m < V; ++m) acc[m] = T(0);
    if constexpr (!IS_FP64) {
      // float form: the directional contractions sum_n Kxi[l][n] lFhi_d(n) as
      // line tasks in place (as in phase 2), then each node sums its three
      // directions with their weights
      if (USE_FLUX) {
        __syncthreads();  // the extrapolation has read sF
#pragma unroll
        for (int r = 0; r < RD; ++r) {
          if (tid + r * NT < NTD) {
            const int task = tid + r * NT;
            const int m = task / B2, line = task - (task / B2) * B2;
            const int la = line / B, lb = line - (line / B) * B;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const int stride = d == 0 ? 1 : (d == 1 ? B : B2);
              const int o = (d * V + m) * B3 + (d == 0 ? line * B : (d == 1 ? la * B2 + lb : line));
              T f[B];
#pragma unroll
              for (int n = 0; n < B; ++n) f[n] = sF[o + n * stride];
#pragma unroll
              for (int l = 0; l < B; ++l) {
                T c = T(0);
#pragma unroll
                for (int n = 0; n < B; ++n) c += f[n] * tc<T>(exa_kxi[N][l][n]);
                sF[o + l * stride] = c;
              }
            }
          }
        }
        __syncthreads();
        const T us[3] = {wz * wy * invDx[0], wz * wx * invDx[1], wy * wx * invDx[2]};
#pragma unroll
        for (int m = 0; m < V; ++m)
          acc[m] = us[0] * sF[(0 * V + m) * B3 + node] + us[1] * sF[(1 * V + m) * B3 + node] +
                   us[2] * sF[(2 * V + m) * B3 + node];
      }
    } else if (USE_FLUX) {
      {
        const T us = wz * wy * invDx[0];
        const T* row = sF + (0 * V) * B3 + (z * B + y) * B;
#pragma unroll
        for (int m = 0; m < V; ++m)
#pragma unroll
          for (int n = 0; n < B; ++n) acc[m] += sK[x * B + n] * row[m * B3 + n] * us;
      }
      {
        const T us = wz * wx * invDx[1];
        const T* col = sF + (1 * V) * B3 + z * B2 + x;
#pragma unroll
        for (int m = 0; m < V; ++m)
#pragma unroll
          for (int n = 0; n < B; ++n) acc[m] += sK[y * B + n] * col[m * B3 + n * B] * us;
      }
      {
        const T us = wy * wx * invDx[2];
        const T* col = sF + (2 * V) * B3 + y * B + x;
#pragma unroll
        for (int m = 0; m < V; ++m)
#pragma unroll
          for (int n = 0; n < B; ++n) acc[m] += sK[z * B + n] * col[m * B3 + n * B2] * us;
      }
    }
    if (USE_SRC_NCP) {
#pragma unroll
      for (int m = 0; m < V; ++m) acc[m] += w3 * sh[m];
    }
    T* lduh = out_lduh + cell * (size_t)(B3 * V) + node * V;
#pragma unroll
    for (int m = 0; m < V; ++m) lduh[m] = acc[m];
  }
  if (a.iters && tid == 0) a.iters[cell] = iters;
}

}  // namespace exa
