// aderdg_kernels.hip -- ADER-DG time-step kernels for gfx950 (CDNA4), templated
// on the arithmetic type T of the bulk work (double: the plain form; float:
// the FP32-bulk form).
//
// One workgroup per cell for the cell-local stages, one thread per spatial
// node (z,y,x) of the B^3 Gauss-Legendre nodes, padded to whole 64-wide
// wavefronts.  All HBM arrays use the reference's per-cell layouts (variable
// innermost), so a cell is one contiguous, coalesced block.
//
// Reference semantics followed (K = ExaHyPE/kernels/aderdg/generic/c/3d):
//   predictor_kernel   K/spaceTimePredictorNonlinear.cpph:42-528,661-717
//                      + K/volumeIntegralNonlinear.cpph:136-225 (fused epilogue)
//   riemann_kernel     K/riemannSolverNonlinear.cpph:37-117
//   update_kernel      K/surfaceIntegralNonlinear.cpph:21-90, K/solutionUpdate.cpph:23-47,
//                      K/stableTimeStepSize.cpph:21-63 (per-cell minimum)
//   dt_finalize_kernel global min of the cell time steps (mappings/SolutionUpdate.cpp:269-327)
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
  // 19% slower, so N=6 stays unrolled).
  static constexpr bool ROLLED = (B >= 7 && (IS_FP64 || B >= 8)) || V >= 9 || (!IS_FP64 && (B == 5 || B == 6));
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
template <int B, class PDE, int TERMS, bool EXPORT, typename T, int LDSK = 144>
__global__ __launch_bounds__(node_threads(B)) void predictor_kernel(PredictorArgs a) {
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
  // accumulating the averages needs 3V more registers: only at B = 5, where
  // they fit the budget of four wavefronts per SIMD (measured register counts:
  // B = 3, 4 and 6 would drop a wavefront per SIMD, B = 7 would spill)
  constexpr bool ACC_FH = FLOAT_FLUX_ONLY && B == 5;
  T fh[3][V];
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
  T sh[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    if constexpr (!ACC_FH) {
      fh[0][m] = T(0);
      fh[1][m] = T(0);
      fh[2][m] = T(0);
    }
    sh[m] = T(0);
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
  if constexpr (ACC_FH) {
    // fh holds the time average of the last iteration's fluxes already
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
  const T sL = PDE::max_abs_eigenvalue(QavL, d);  // :68-78
  const T sR = PDE::max_abs_eigenvalue(QavR, d);
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

// dt_next = min(slots); t += dt (unless initialising); slots reset.
__global__ __launch_bounds__(256) void dt_finalize_kernel(StepState* st, int advance_time) {
  __shared__ double sRed[256 / kWave];
  double m = 1.0e308;
  for (int i = threadIdx.x; i < kDtSlots; i += 256) {
    const double v = __longlong_as_double((long long)atomic_take(&st->dt_slots[i], kDtSlotEmpty));
    m = v < m ? v : m;
  }
  m = block_min<256>(m, sRed);
  if (threadIdx.x == 0) {
    if (advance_time) {
      st->t += st->dt;
      st->steps += 1;
    }
    st->dt = m;
    if (st->dt_hist && st->steps < st->dt_hist_len) st->dt_hist[st->steps] = m;
  }
}

// ---------------------------------------------------------------------------
// Stand-alone per-cell kernels for the drop-in single-cell / batch API (FP64).
// ---------------------------------------------------------------------------

// volumeIntegralNonlinear (time-averaged form) from caller-provided lFhi.
template <int B, int V, bool USE_FLUX, bool USE_SRC_NCP>
__global__ __launch_bounds__(node_threads(B)) void volume_integral_kernel(VolumeArgs a) {
  constexpr int N = B - 1;
  constexpr int B2 = B * B, B3 = B2 * B;
  const size_t cell = blockIdx.x;
  const int node = threadIdx.x;
  if (node >= B3) return;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const double* dxp = a.dx + cell * a.dx_stride;
  const double invDx[3] = {1.0 / dxp[0], 1.0 / dxp[1], 1.0 / dxp[2]};
  const double* lFhi = a.lFhi + cell * (size_t)(4 * B3 * V);
  const double wx = exa_gl_weights[N][x], wy = exa_gl_weights[N][y], wz = exa_gl_weights[N][z];
  double acc[V];
#pragma unroll
  for (int m = 0; m < V; ++m) acc[m] = 0.0;
  if (USE_FLUX) {
    const double usx = wz * wy * invDx[0], usy = wz * wx * invDx[1], usz = wy * wx * invDx[2];
#pragma unroll
    for (int m = 0; m < V; ++m)
#pragma unroll
      for (int n = 0; n < B; ++n) acc[m] += exa_kxi[N][x][n] * lFhi[0 * B3 * V + ((z * B + y) * B + n) * V + m] * usx;
#pragma unroll
    for (int m = 0; m < V; ++m)
#pragma unroll
      for (int n = 0; n < B; ++n) acc[m] += exa_kxi[N][y][n] * lFhi[1 * B3 * V + ((z * B + x) * B + n) * V + m] * usy;
#pragma unroll
    for (int m = 0; m < V; ++m)
#pragma unroll
      for (int n = 0; n < B; ++n) acc[m] += exa_kxi[N][z][n] * lFhi[2 * B3 * V + ((y * B + x) * B + n) * V + m] * usz;
  }
  if (USE_SRC_NCP) {
    const double w3 = wz * wy * wx;
#pragma unroll
    for (int m = 0; m < V; ++m) acc[m] += w3 * lFhi[3 * B3 * V + node * V + m];
  }
  double* lduh = a.lduh + cell * (size_t)(B3 * V) + node * V;
#pragma unroll
  for (int m = 0; m < V; ++m) lduh[m] = acc[m];
}

// ---------------------------------------------------------------------------
// Initial conditions (evaluated at the Gauss-Legendre nodes of each cell,
// x = centre + dx (xi - 1/2), like solutionAdjustment.cpph:32-43).
// ---------------------------------------------------------------------------
__global__ void euler_ic_kernel(ICArgs a) {
  const int B = a.order + 1;
  const int B3 = B * B * B;
  const int D = a.nvar + a.npar;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)a.ncells * B3;
  if (gid >= total) return;
  const size_t cell = gid / B3;
  const int node = (int)(gid % B3);
  const int iz = node / (B * B), iy = (node / B) % B, ix = node % B;
  const int ci = (int)(cell % a.nx), cj = (int)((cell / a.nx) % a.ny), ck = (int)(cell / ((size_t)a.nx * a.ny));
  const double X[3] = {a.lo[0] + (ci + 0.5) * a.dx[0] + a.dx[0] * (exa_gl_nodes[a.order][ix] - 0.5),
                       a.lo[1] + (cj + 0.5) * a.dx[1] + a.dx[1] * (exa_gl_nodes[a.order][iy] - 0.5),
                       a.lo[2] + (ck + 0.5) * a.dx[2] + a.dx[2] * (exa_gl_nodes[a.order][iz] - 0.5)};
  const double twopi = 6.283185307179586476925286766559;
  double rho = a.rho0 + a.amp * sin(twopi * (a.k[0] * X[0] + a.k[1] * X[1] + a.k[2] * X[2]));
  double u[3] = {a.vel[0], a.vel[1], a.vel[2]};
  for (int mo = 0; mo < a.nmodes; ++mo) {
    const double arg = twopi * (a.mode_k[mo][0] * X[0] + a.mode_k[mo][1] * X[1] + a.mode_k[mo][2] * X[2]);
    rho += a.mode_amp * sin(arg + a.mode_phase[mo][0]);
    u[0] += a.mode_amp * sin(arg + a.mode_phase[mo][1]);
    u[1] += a.mode_amp * sin(arg + a.mode_phase[mo][2]);
    u[2] += a.mode_amp * sin(arg + a.mode_phase[mo][3]);
  }
  const double GAMMA = 1.4;
  double* Q = a.luh + gid * D;
  Q[0] = rho;
  Q[1] = rho * u[0];
  Q[2] = rho * u[1];
  Q[3] = rho * u[2];
  Q[4] = a.p / (GAMMA - 1.0) + 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int m = 5; m < D; ++m) Q[m] = 0.0;
}

__global__ __launch_bounds__(256) void elastic_ic_kernel(ElasticICArgs a) {
  const int B = a.order + 1;
  const int B3 = B * B * B;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)a.ncells * B3) return;
  const size_t cell = gid / B3;
  const int node = (int)(gid % B3);
  const int iz = node / (B * B), iy = (node / B) % B, ix = node % B;
  const int ci = (int)(cell % a.nx), cj = (int)((cell / a.nx) % a.ny), ck = (int)(cell / ((size_t)a.nx * a.ny));
  const double x = (a.lo[0] + (ci + 0.5) * a.dx[0]) + a.dx[0] * (exa_gl_nodes[a.order][ix] - 0.5);
  const double y = (a.lo[1] + (cj + 0.5) * a.dx[1]) + a.dx[1] * (exa_gl_nodes[a.order][iy] - 0.5);
  const double z = (a.lo[2] + (ck + 0.5) * a.dx[2]) + a.dx[2] * (exa_gl_nodes[a.order][iz] - 0.5);
  const double twopi = 6.283185307179586476925286766559;
  const double z0 = 0.5 + 0.1 * sin(twopi * x + twopi * a.r[0]) * sin(twopi * y + twopi * a.r[1]);
  const bool upper = z > z0;
  const double rho = (upper ? 2.05 + 0.9 * a.r[2] : 2.05 + 0.9 * a.r[3]) + 0.05 * sin(twopi * ((x + y) + z));
  const double cp = (upper ? 4.1 + 1.8 * a.r[4] : 4.1 + 1.8 * a.r[5]) + 0.1 * sin(twopi * ((x - y) + z));
  const double cs = cp / sqrt(3.0);
  const double mu = rho * cs * cs;
  const double lam = rho * cp * cp - 2.0 * mu;
  const double r2 = ((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5)) + (z - 0.5) * (z - 0.5);
  const double g = a.amp * exp(-r2 / (2.0 * a.width * a.width));
  double* Q = a.luh + gid * 12;
  for (int m = 0; m < 6; ++m) Q[m] = 0.0;
  Q[6] = g;
  Q[7] = g;
  Q[8] = g;
  Q[9] = lam;
  Q[10] = mu;
  Q[11] = rho;
}

hipError_t launch_elastic_ic(const ElasticICArgs& a, hipStream_t s) {
  const size_t total = (size_t)a.ncells * (a.order + 1) * (a.order + 1) * (a.order + 1);
  hipLaunchKernelGGL(elastic_ic_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Sum of the per-cell Picard counts of one predictor launch (for the roofline).
__global__ __launch_bounds__(256) void count_iterations_kernel(const int* iters, long ncells, StepState* st) {
  __shared__ double sRed[256 / kWave];
  double acc = 0.0;
  for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < ncells; c += (long)gridDim.x * 256) acc += iters[c];
  acc = block_sum<256>(acc, sRed);
  if (threadIdx.x == 0) {
    atomicAdd(&st->picard_total, (unsigned long long)acc);
    if (blockIdx.x == 0) atomicAdd(&st->predictor_cells, (unsigned long long)ncells);
  }
}

// ---------------------------------------------------------------------------
// Host-side dispatch: runtime (pde, order, npar, terms, precision) -> instantiation.
// ---------------------------------------------------------------------------
#define EXA_ORDERS(X) X(2) X(3) X(4) X(5) X(6) X(7)

// Calls F.template operator()<B, PDE>() for the compiled (pde, order, npar).
template <class F>
static hipError_t dispatch_pde(const SolverDesc& sd, F&& f) {
  if (sd.pde == kPdeEuler && sd.npar == 0) {
#define EXA_CASE(N) \
  if (sd.order == N) return f.template operator()<N + 1, Euler<0>>();
    EXA_ORDERS(EXA_CASE)
#undef EXA_CASE
  }
  if (sd.pde == kPdeEuler && sd.npar == 1 && sd.order == 3) return f.template operator()<4, Euler<1>>();
  if (sd.pde == kPdeElastic && sd.nvar == Elastic::V && sd.npar == Elastic::P) {
    switch (sd.order) {
      case 2: return f.template operator()<3, Elastic>();
      case 3: return f.template operator()<4, Elastic>();
      case 4: return f.template operator()<5, Elastic>();
      case 5: return f.template operator()<6, Elastic>();
      default: break;
    }
  }
  return hipErrorNotSupported;
}

hipError_t launch_predictor(const SolverDesc& sd, bool export_temps, const PredictorArgs& a, long ncells,
                            hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (export_temps && sd.precision != kFP64) return hipErrorNotSupported;
  if (a.variant == kPredictorMfma) {
    if (export_temps) return hipErrorNotSupported;
    return launch_predictor_mfma(sd, a, ncells, s);
  }
  const size_t ws = predictor_ws_bytes_per_cell(sd, a.variant);
  if (ws > 0 && (a.ws == nullptr || a.ws_cells <= 0)) return hipErrorInvalidValue;
  const long chunk = ws > 0 ? a.ws_cells : ncells;
  return dispatch_pde(sd, [&]<int B, class PDE>() -> hipError_t {
    // chunks of at most ws_cells cells share the workspace (stream order)
    for (long c0 = 0; c0 < ncells; c0 += chunk) {
      PredictorArgs ac = a;
      ac.cell0 = a.cell0 + c0;
      const dim3 grid((unsigned)(ncells - c0 < chunk ? ncells - c0 : chunk)), block(node_threads(B));
      if constexpr (std::is_same_v<PDE, Elastic>) {  // non-conservative product only
        if (sd.terms != kNCP) return hipErrorNotSupported;
        if (sd.precision == kFP32Bulk)
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kNCP, false, float>), grid, block, 0, s, ac);
        else if (export_temps)
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kNCP, true, double>), grid, block, 0, s, ac);
        else
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kNCP, false, double>), grid, block, 0, s, ac);
      } else if (sd.precision == kFP32Bulk) {
        if (sd.terms != kFlux) return hipErrorNotSupported;
        if (a.variant == kPredictorValu64 && PDE::P == 0)
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux, false, float, 64>), grid, block, 0, s, ac);
        else
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux, false, float>), grid, block, 0, s, ac);
      } else if (sd.terms == kFlux) {
        if (a.variant == kPredictorValu64 && PDE::P == 0 && !export_temps)
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux, false, double, 64>), grid, block, 0, s, ac);
        else if (export_temps)
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux, true, double>), grid, block, 0, s, ac);
        else
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux, false, double>), grid, block, 0, s, ac);
      } else if (sd.terms == (kFlux | kSource | kNCP)) {
        if (export_temps)
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux | kSource | kNCP, true, double>), grid, block, 0, s, ac);
        else
          hipLaunchKernelGGL((predictor_kernel<B, PDE, kFlux | kSource | kNCP, false, double>), grid, block, 0, s,
                             ac);
      } else {
        return hipErrorNotSupported;
      }
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  });
}

static bool predictor_has_lds64(const SolverDesc& sd) {
  return sd.pde == kPdeEuler && sd.npar == 0 && sd.terms == kFlux;
}

unsigned predictor_variants(const SolverDesc& sd) {
  unsigned v = 1u << kPredictorAuto;
  if (predictor_has_lds64(sd)) v |= 1u << kPredictorValu64;
  if (predictor_mfma_supported(sd)) v |= 1u << kPredictorMfma;
  return v;
}

bool predictor_variant_available(const SolverDesc& sd, int variant) {
  if (sd.linear) return variant == kPredictorAuto;
  if (variant == kPredictorAuto || variant == kPredictorValu144) return true;
  return variant >= 0 && variant < kPredictorVariants && (predictor_variants(sd) & (1u << variant)) != 0;
}

size_t predictor_ws_bytes_per_cell(const SolverDesc& sd, int variant) {
  size_t bytes = 0;
  (void)dispatch_pde(sd, [&]<int B, class PDE>() -> hipError_t {
    constexpr int V = PDE::V, D = PDE::V + PDE::P;
    if (variant == kPredictorValu64 && predictor_has_lds64(sd)) {
      bytes = sd.precision == kFP32Bulk ? PredictorLayout<B, V, D, false, float, 64>::ws_bytes
                                        : PredictorLayout<B, V, D, false, double, 64>::ws_bytes;
      return hipSuccess;
    }
    if (sd.precision == kFP32Bulk)
      bytes = (sd.terms & kNCP) ? PredictorLayout<B, V, D, true, float>::ws_bytes
                                : PredictorLayout<B, V, D, false, float>::ws_bytes;
    else if (sd.terms & kNCP)
      bytes = PredictorLayout<B, V, D, true, double>::ws_bytes;
    else
      bytes = PredictorLayout<B, V, D, false, double>::ws_bytes;
    return hipSuccess;
  });
  return bytes;
}

hipError_t launch_riemann(const SolverDesc& sd, const RiemannArgs& a, hipStream_t s) {
  if (a.nfaces <= 0) return hipSuccess;
  return dispatch_pde(sd, [&]<int B, class PDE>() -> hipError_t {
    const long per_block = 4L * (kWave / riemann_segment(B));  // 4 wavefronts
    const dim3 grid((unsigned)((a.nfaces + per_block - 1) / per_block)), block(256);
    const bool ncp = (sd.terms & kNCP) != 0;
    if constexpr (std::is_same_v<PDE, Elastic>) {
      if (sd.precision == kFP32Bulk)
        hipLaunchKernelGGL((riemann_kernel<B, PDE, true, float>), grid, block, 0, s, a);
      else
        hipLaunchKernelGGL((riemann_kernel<B, PDE, true, double>), grid, block, 0, s, a);
    } else if (sd.precision == kFP32Bulk) {
      if (ncp) return hipErrorNotSupported;
      hipLaunchKernelGGL((riemann_kernel<B, PDE, false, float>), grid, block, 0, s, a);
    } else if (ncp) {
      hipLaunchKernelGGL((riemann_kernel<B, PDE, true, double>), grid, block, 0, s, a);
    } else {
      hipLaunchKernelGGL((riemann_kernel<B, PDE, false, double>), grid, block, 0, s, a);
    }
    return hipGetLastError();
  });
}

hipError_t launch_update(const SolverDesc& sd, int mode, const UpdateArgs& a, long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  return dispatch_pde(sd, [&]<int B, class PDE>() -> hipError_t {
    const dim3 grid((unsigned)ncells), block(node_threads(B));
    if (sd.linear && (mode & kUpdSurface)) {  // the linear solvers' surface integral
      if constexpr (std::is_same_v<PDE, Elastic> || std::is_same_v<PDE, Euler<1>>) {
        if (sd.precision != kFP64) return hipErrorNotSupported;
        if (mode == (kUpdSurface | kUpdUpdate))
          hipLaunchKernelGGL((update_kernel<B, PDE, true, true, false, double, true>), grid, block, 0, s, a);
        else if (mode == (kUpdSurface | kUpdWriteLduh))
          hipLaunchKernelGGL((update_kernel<B, PDE, true, false, true, double, true>), grid, block, 0, s, a);
        else
          return hipErrorInvalidValue;
        return hipGetLastError();
      } else {
        return hipErrorNotSupported;
      }
    }
    if (sd.precision == kFP32Bulk) {
      if (mode == (kUpdSurface | kUpdUpdate))
        hipLaunchKernelGGL((update_kernel<B, PDE, true, true, false, float>), grid, block, 0, s, a);
      else if (mode == 0)
        hipLaunchKernelGGL((update_kernel<B, PDE, false, false, false, float>), grid, block, 0, s, a);
      else
        return hipErrorNotSupported;
      return hipGetLastError();
    }
    switch (mode) {
      case kUpdSurface | kUpdUpdate:
        hipLaunchKernelGGL((update_kernel<B, PDE, true, true, false, double>), grid, block, 0, s, a);
        break;
      case kUpdSurface | kUpdWriteLduh:
        hipLaunchKernelGGL((update_kernel<B, PDE, true, false, true, double>), grid, block, 0, s, a);
        break;
      case kUpdUpdate:
        hipLaunchKernelGGL((update_kernel<B, PDE, false, true, false, double>), grid, block, 0, s, a);
        break;
      case 0:
        hipLaunchKernelGGL((update_kernel<B, PDE, false, false, false, double>), grid, block, 0, s, a);
        break;
      default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
  });
}

__global__ __launch_bounds__(256) void patch_dt_finalize_kernel(double* dt, double* t, unsigned long long* slots,
                                                                int npatches, int sp, int advance) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= npatches) return;
  double m = DBL_MAX;
  for (int i = 0; i < sp; ++i) {
    const double v = __longlong_as_double((long long)atomic_take(&slots[(size_t)p * sp + i], kDtSlotEmpty));
    m = v < m ? v : m;
  }
  if (advance) t[p] += dt[p];
  dt[p] = m;
}

hipError_t launch_patch_dt_finalize(double* dt, double* t, unsigned long long* slots, int npatches,
                                    int slots_per_patch, bool advance_time, hipStream_t s) {
  hipLaunchKernelGGL(patch_dt_finalize_kernel, dim3((npatches + 255) / 256), dim3(256), 0, s, dt, t, slots, npatches,
                     slots_per_patch, advance_time ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_dt_finalize(StepState* st, bool advance_time, hipStream_t s) {
  hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(256), 0, s, st, advance_time ? 1 : 0);
  return hipGetLastError();
}

// lFhi (x (z,y,x), y (z,x,y), z (y,x,z), src (z,y,x)) = sum_t w_t lFi(t,z,y,x,{x,y,z,src})
// (aderPredictorNonlinear, spaceTimePredictorNonlinear.cpph:367-396): the
// time average the noTimeAveraging form leaves to the volume integral.
template <int B>
__global__ __launch_bounds__(node_threads(B)) void time_average_lfi_kernel(const double* lFi, double* lFhi, int V) {
  constexpr int N = B - 1, B2 = B * B, B3 = B2 * B;
  const int node = threadIdx.x;
  if (node >= B3) return;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const int dst[4] = {(z * B + y) * B + x, (z * B + x) * B + y, (y * B + x) * B + z, (z * B + y) * B + x};
  for (int blk = 0; blk < 4; ++blk)
    for (int m = 0; m < V; ++m) {
      double acc = 0.0;
      for (int t = 0; t < B; ++t) acc += lFi[(((size_t)t * B3 + node) * 4 + blk) * V + m] * exa_gl_weights[N][t];
      lFhi[(size_t)blk * B3 * V + dst[blk] * V + m] = acc;
    }
}

hipError_t launch_time_average_lfi(int order, int V, const double* lFi, double* lFhi, hipStream_t s) {
  switch (order + 1) {
#define EXA_TA(Bv) \
  case Bv:         \
    hipLaunchKernelGGL(time_average_lfi_kernel<Bv>, dim3(1), dim3(node_threads(Bv)), 0, s, lFi, lFhi, V); break;
    EXA_TA(3) EXA_TA(4) EXA_TA(5) EXA_TA(6) EXA_TA(7) EXA_TA(8)
#undef EXA_TA
    default: return hipErrorNotSupported;
  }
  return hipGetLastError();
}

template <int B, int V>
static hipError_t launch_volume_t(int terms, const VolumeArgs& a, long ncells, hipStream_t s) {
  const dim3 grid((unsigned)ncells), block(node_threads(B));
  const bool f = terms & kFlux, sn = terms & (kSource | kNCP);
  if (f && sn)
    hipLaunchKernelGGL((volume_integral_kernel<B, V, true, true>), grid, block, 0, s, a);
  else if (f)
    hipLaunchKernelGGL((volume_integral_kernel<B, V, true, false>), grid, block, 0, s, a);
  else if (sn)
    hipLaunchKernelGGL((volume_integral_kernel<B, V, false, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((volume_integral_kernel<B, V, false, false>), grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_volume_integral(const SolverDesc& sd, const VolumeArgs& a, long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (sd.nvar == 5) {
#define EXA_CASE(N) \
  if (sd.order == N) return launch_volume_t<N + 1, 5>(sd.terms, a, ncells, s);
    EXA_ORDERS(EXA_CASE)
#undef EXA_CASE
  }
  return hipErrorNotSupported;
}

hipError_t launch_euler_ic(const ICArgs& a, hipStream_t s) {
  const size_t total = (size_t)a.ncells * (a.order + 1) * (a.order + 1) * (a.order + 1);
  const unsigned blocks = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(euler_ic_kernel, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_count_iterations(const int* iters, long ncells, StepState* st, hipStream_t s) {
  hipLaunchKernelGGL(count_iterations_kernel, dim3(512), dim3(256), 0, s, iters, ncells, st);
  return hipGetLastError();
}

// noTimeAveraging (spaceTimePredictorNonlinear.cpph:690-695) selects where
// the time average is taken: the reference extrapolates and integrates the
// space-time data directly; the device averages first.  The operations are
// the same linear maps summed in another order, so both settings run the same
// kernels.
bool supported_sizes(const SolverDesc& sd) { return sd.nvar == 5 && sd.order >= 2 && sd.order <= 7; }

bool supported(const SolverDesc& sd) {
  if (sd.pde == kPdeElastic)
    return sd.nvar == Elastic::V && sd.npar == Elastic::P && sd.terms == kNCP && sd.order >= 2 && sd.order <= 5 &&
           (sd.precision == kFP64 || sd.precision == kFP32Bulk);
  if (sd.pde != kPdeEuler || sd.nvar != 5) return false;
  if (sd.precision != kFP64 && sd.precision != kFP32Bulk) return false;
  if (sd.npar == 0) return sd.order >= 2 && sd.order <= 7 && sd.terms == kFlux;
  if (sd.npar == 1)
    return sd.precision == kFP64 && sd.order == 3 && (sd.terms == kFlux || sd.terms == (kFlux | kSource | kNCP));
  return false;
}

}  // namespace exa
