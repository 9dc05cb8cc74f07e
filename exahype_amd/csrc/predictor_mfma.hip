// predictor_mfma.hip -- the FP32-bulk space-time predictor (Picard iteration,
// time averaging, extrapolation, volume integral) with its tensor contractions
// on the CDNA4 matrix cores.
//
// Reference semantics: ExaHyPE/kernels/aderdg/generic/c/3d/
//   spaceTimePredictorNonlinear.cpph:42-528 (aderPicardLoopNonlinear,
//   aderPredictorNonlinear, aderExtrapolatorNonlinear) and
//   volumeIntegralNonlinear.cpph:136-225, as in predictor_kernel; the float
//   form's stopping rule is that of predictor_kernel.
//
// Matrix-core mapping.  Every contraction of the predictor applies a B x B
// operator along one index: the line contraction with Kxi (per direction, per
// variable and grid line, :149-236), the update of the next iterate with iK1
// over the time slices (:280-299), the face extrapolation with FLCoeff and
// FRCoeff (:434-527) and the volume integral's Kxi^T (:138-204).
// v_mfma_f32_4x4x1f32 runs 16 independent 4x4x1 outer products per
// wavefront: with CBSZ = 4 the A operand of block ABID goes to all 16 blocks,
// so lane L computes D_L[j] += A[j] * b_L (j = 0..3) -- four output rows of
// the operator for the lane's own input column.  A lane that owns one grid
// line (or one node) therefore applies a B x B operator to it with B * G
// instructions, G = ceil(B / 4) four-row groups: 75% of the matrix-core rate
// at B = 6, without moving data between lanes.  The A operands (the operator
// columns) are packed 16 to a register: lane 4 s + r holds row r of slot s,
// slot s = n * G + g for input index n and row group g.
#include "aderdg.hpp"
#include "exa_device.hpp"

#include <type_traits>
#include <utility>

namespace exa {

// Calls f(std::integral_constant<int, i>{}) for i = 0 .. N-1 (compile-time
// indices for the matrix-core slot immediates).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

typedef float v4f __attribute__((ext_vector_type(4)));

// D_L[j] = C_L[j] + A_s[j] * b_L: slot s of the packed operator register
// (broadcast to every block) times the lane's own value.
template <int S>
__device__ __forceinline__ v4f mfma_slot(float packed, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(packed, b, c, 4, S, 0);
}

// D_L[j] = C_L[j] + a_{L - L%4 + j} * b_L: every block reads its own four
// lanes' A values (used with the same value in every block).
__device__ __forceinline__ v4f mfma_own(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

constexpr float kPicardRelTol32Mfma = 1e-6f;
constexpr double kPicardTol2Mfma = 1e-7 * 1e-7;

template <int B>
struct MfmaShape {
  static constexpr int G = (B + 3) / 4;  // four-row groups of the B outputs
  static_assert(B * G <= 16, "operator columns must fit the 16 packed slots");
};

// Packed operator register for a contraction out[l] = sum_n M(n, l) in[n]:
// lane 4 s + r holds M(n, 4 g + r) for slot s = n G + g (zero past B).
template <int B, class F>
__device__ inline float pack_operator(int lane, F m) {
  constexpr int G = MfmaShape<B>::G;
  const int s = lane >> 2, r = lane & 3;
  const int n = s / G, l = 4 * (s % G) + r;
  return (n < B && l < B) ? m(n, l) : 0.0f;
}

// Wavefronts per SIMD the register budget is sized for: four workgroups of
// four wavefronts per CU up to B = 6, two of eight from B = 7.
__host__ __device__ constexpr int mfma_predictor_waves_per_eu(int B) { return B <= 6 ? 4 : 2; }

// In-place line contraction on the matrix cores.  sF holds NS slices of three
// blocks [s][d][var][node]; block (s, d) is contracted along direction d:
// out(l) = sum_n M_d(n, l) in(n) for every (variable, line), M_d the packed
// operator op_d, one line per lane, 64 lines per wavefront group.  Groups
// are dealt to the NW wavefronts round-robin.  A line is read whole before its
// outputs overwrite it, and no other lane touches it.  Ends without a barrier.
template <int B, int V, int NW, int NS>
__device__ inline void contract_lines_mfma(float* sF, float op0, float op1, float op2, int lane, int wave) {
  constexpr int G = MfmaShape<B>::G;
  constexpr int B2 = B * B, B3 = B2 * B;
  constexpr int NTD = V * B2;                     // lines per (slice, direction) block
  constexpr int GPD = (NTD + kWave - 1) / kWave;  // wavefront groups per block
  constexpr int NG = NS * 3 * GPD;                // groups in all
  constexpr int KMAX = (NG + NW - 1) / NW;        // groups per wavefront, at most
  // one group of a block along compile-time direction DIR (the strides and
  // the operator are then constants of the code, not a run-time selection)
  auto group = [&]<int DIR>(int sl, int j, float op) __device__ {
    constexpr int stride = DIR == 0 ? 1 : (DIR == 1 ? B : B2);
    const int task = j * kWave + lane;
    const bool valid = task < NTD;
    const int tc = valid ? task : NTD - 1;
    const int m = tc / B2, line = tc - (tc / B2) * B2;
    const int la = line / B, lb = line - (line / B) * B;
    const int o = ((sl * 3 + DIR) * V + m) * B3 + (DIR == 0 ? line * B : (DIR == 1 ? la * B2 + lb : line));
    float f[B];
#pragma unroll
    for (int n = 0; n < B; ++n) f[n] = sF[o + n * stride];
    v4f acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = v4f{0.f, 0.f, 0.f, 0.f};
    static_for<B>([&](auto nc) __device__ {
      static_for<G>([&](auto gc) __device__ { acc[gc] = mfma_slot<nc * G + gc>(op, f[nc], acc[gc]); });
    });
    if (valid) {
#pragma unroll
      for (int l = 0; l < B; ++l) sF[o + l * stride] = acc[l / 4][l % 4];
    }
  };
#pragma unroll 1
  for (int k = 0; k < KMAX; ++k) {
    const int q = wave + k * NW;
    if (q >= NG) break;  // uniform per wavefront
    const int blk = q / GPD, j = q - (q / GPD) * GPD;
    const int sl = blk / 3, d = blk - (blk / 3) * 3;
    if (d == 0)
      group.template operator()<0>(sl, j, op0);
    else if (d == 1)
      group.template operator()<1>(sl, j, op1);
    else
      group.template operator()<2>(sl, j, op2);
  }
}

// Face extrapolation (spaceTimePredictorNonlinear.cpph:434-527) on the matrix
// cores: per direction d one task = (face node (la, lb), datum k), k fastest so
// that a wavefront's stores are contiguous; the task's normal line is
// contracted with (FLCoeff, FRCoeff) (packed operator cE: slot n holds
// FLCoeff[n], FRCoeff[n], 0, 0), D[0] the left and D[1] the right face value.
// Source values: sF[k][node] for every direction, or with PER_DIR the
// direction's own block sF[d][k][node].  Ends without a barrier.
template <int B, int V, int NK, bool PER_DIR, int NW>
__device__ inline void extrapolate_mfma(const float* sF, float* out, size_t coff, float cE, int lane, int wave) {
  constexpr int B2 = B * B, B3 = B2 * B;
  constexpr int NTK = B2 * NK;                    // tasks per direction
  constexpr int GPD = (NTK + kWave - 1) / kWave;  // groups per direction
  constexpr int NG = 3 * GPD;
  auto group = [&]<int DIR>(int j) __device__ {
    constexpr int stride = DIR == 0 ? 1 : (DIR == 1 ? B : B2);
    const int task = j * kWave + lane;
    const bool valid = task < NTK;
    const int tc = valid ? task : NTK - 1;
    const int kk = tc % NK, line = tc / NK;
    const int la = line / B, lb = line - (line / B) * B;
    const int o = (PER_DIR ? DIR * V + kk : kk) * B3 + (DIR == 0 ? line * B : (DIR == 1 ? la * B2 + lb : line));
    float f[B];
#pragma unroll
    for (int n = 0; n < B; ++n) f[n] = sF[o + n * stride];
    v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
    static_for<B>([&](auto nc) __device__ { acc = mfma_slot<nc>(cE, f[nc], acc); });
    if (valid) {
      out[coff + (size_t)(2 * DIR) * NTK + task] = acc[0];
      out[coff + (size_t)(2 * DIR + 1) * NTK + task] = acc[1];
    }
  };
#pragma unroll 1
  for (int k = 0; k < (NG + NW - 1) / NW; ++k) {
    const int q = wave + k * NW;
    if (q >= NG) break;  // uniform per wavefront
    const int d = q / GPD, j = q - (q / GPD) * GPD;
    if (d == 0)
      group.template operator()<0>(j);
    else if (d == 1)
      group.template operator()<1>(j);
    else
      group.template operator()<2>(j);
  }
}

// ---------------------------------------------------------------------------
// One workgroup per cell, one thread per node (as predictor_kernel).  The
// Picard iterate lQi^k lives in LDS ([t][var][node], each thread touching only
// its own node).  Per time slice: the node owners stage the slice's fluxes in
// LDS, the line owners contract them with the per-direction operators
// Kxi[n][l] dt / (dx_d w_l) (the node weights and dt / dx of :152-236 folded
// in), and the node owners fold the slice into the next iterate on the matrix
// cores: qn[m][g] (rows l = 4g..4g+3 of variable m at the lane's node) +=
// iK1[l][t] * ir(t), ir = rhs / w^3 (:280-299).
// ---------------------------------------------------------------------------
template <int B, class PDE, int TERMS>
__global__ __launch_bounds__(node_threads(B)) __attribute__((amdgpu_waves_per_eu(mfma_predictor_waves_per_eu(B))))
void predictor_mfma_kernel(PredictorArgs a) {
  static_assert(TERMS == kFlux, "the matrix-core predictor is built for flux-only PDEs");
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B, B3 = B2 * B;
  constexpr int NT = node_threads(B);
  constexpr int NW = NT / kWave;
  constexpr int G = MfmaShape<B>::G;
  static_assert(D <= 3 * V, "lQhi staging fits the flux buffer");
  static_assert(2 * G <= 16, "first-iteration coefficients fit the packed slots");
  constexpr int MAX_ITER = 2 * B;  // spaceTimePredictorNonlinear.cpph:85
  // rolled loop over time slices from B = 5 (registers)
  constexpr bool ROLLED = B >= 5;
  __shared__ float sF[3 * V * B3];    // one slice's fluxes [d][var][node]
  __shared__ float sQi[B * V * B3];   // lQi^k, [t][var][node]
  __shared__ float sIK[B * G * 4];    // [t][g][r] = iK1[4g + r][t] (zero past B)
  __shared__ float sRed[2 * NW];

  float* const out_lduh = static_cast<float*>(a.lduh);
  float* const out_lQhbnd = static_cast<float*>(a.lQhbnd);
  float* const out_lFhbnd = static_cast<float*>(a.lFhbnd);

  const size_t cell = (size_t)blockIdx.x + a.cell0;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  if (a.run_if && *a.run_if == 0) return;  // whole launch
  if (a.status && a.status[cell] >= a.status_limit) {  // whole workgroup
    if (a.iters && tid == 0) a.iters[cell] = 0;
    return;
  }
  for (int i = tid; i < B * G * 4; i += NT) {
    const int t = i / (4 * G), l = 4 * ((i / 4) % G) + i % 4;
    sIK[i] = l < B ? (float)exa_ik1[N][l][t] : 0.f;
  }
  const int patch = a.cell_patch ? a.cell_patch[cell] : 0;
  const double* dxp = a.cell_patch ? a.dx + (size_t)patch * 3 : a.dx + cell * a.dx_stride;
  const float dt = (float)a.dt[patch];
  // packed operators of the Picard line contractions, per direction:
  // Kxi[n][l] dt / (dx_d w_l)
  float cK[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float s = dt * (1.0f / (float)dxp[d]);
    cK[d] = pack_operator<B>(
        lane, [&](int n, int l) __device__ { return (float)exa_kxi[N][n][l] * s / (float)exa_gl_weights[N][l]; });
  }

  float q0[D];
  {
    const double* luh = a.luh + cell * (size_t)(B3 * D) + node * D;
#pragma unroll
    for (int m = 0; m < D; ++m) q0[m] = (float)luh[m];
  }
#pragma unroll
  for (int t = 0; t < B; ++t)
#pragma unroll
    for (int m = 0; m < V; ++m) sQi[(t * V + m) * B3 + node] = q0[m];
  v4f qn[V][G];
  int iters = MAX_ITER;
  // From B = 7 (two or three wavefronts per SIMD, set by LDS) the registers
  // have room to accumulate the time-averaged fluxes of each iteration as its
  // slices are evaluated, so the epilogue evaluates no slice again.
  constexpr bool ACC_FH = B >= 7;
  float fh[3][V];
  __syncthreads();

  // phase 1: fluxes of Q -> sF[d][var][node]; with ACC_FH also fh += wt F
  auto stage_flux = [&](const float* Qv, float wt) __device__ __attribute__((always_inline)) {
    float Q[D];
#pragma unroll
    for (int m = 0; m < V; ++m) Q[m] = Qv[m];
#pragma unroll
    for (int m = V; m < D; ++m) Q[m] = q0[m];
    float F[3][V];
    PDE::flux(Q, F);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < V; ++m) {
        sF[(d * V + m) * B3 + node] = F[d][m];
        if constexpr (ACC_FH) fh[d][m] += F[d][m] * wt;
      }
  };

#pragma unroll 1
  for (int iter = 0; iter < MAX_ITER; ++iter) {
    // FLCoeff[t] luh is loop invariant; keeping its B V products live across
    // iterations costs more registers than forming them per slice, so the
    // copy is made opaque to loop-invariant code motion
    float q0i[V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
      q0i[m] = q0[m];
      asm volatile("" : "+v"(q0i[m]));
    }
    if constexpr (ACC_FH) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < V; ++m) fh[d][m] = 0.f;
    }
    if (iter == 0) {
      // First iteration: every slice of the initial iterate equals luh, so the
      // slices share one contraction S = sum_d C'_d and
      // lQi^1(l) = (sum_t iK1[l][t] FLCoeff[t]) luh - (sum_t iK1[l][t] w_t) S,
      // coefficients formed in double from the tables (packed: slot g holds
      // rows 4g.. of the first sum, slot G + g of the second, negated).
      {
        // every slice's flux: the average is sum_t w_t F(luh)
        double sw = 0.0;
        for (int t = 0; t < B; ++t) sw += exa_gl_weights[N][t];
        stage_flux(q0i, (float)sw);
      }
      __syncthreads();
      contract_lines_mfma<B, V, NW, 1>(sF, cK[0], cK[1], cK[2], lane, wave);
      __syncthreads();
      float cAB;
      {
        const int sl = lane >> 2, r = lane & 3;
        const int g = sl % G, l = 4 * g + r;
        double acc = 0.0;
        if (sl < 2 * G && l < B) {
          for (int t = 0; t < B; ++t)
            acc += exa_ik1[N][l][t] * (sl < G ? exa_fl_coeff[N][t] : -exa_gl_weights[N][t]);
        }
        cAB = (float)acc;
      }
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const float S = sF[(0 * V + m) * B3 + node] + sF[(1 * V + m) * B3 + node] + sF[(2 * V + m) * B3 + node];
        static_for<G>([&](auto gc) __device__ {
          qn[m][gc] = mfma_slot<gc>(cAB, q0i[m], v4f{0.f, 0.f, 0.f, 0.f});
          qn[m][gc] = mfma_slot<G + gc>(cAB, S, qn[m][gc]);
        });
      }
    } else {
#pragma unroll
      for (int m = 0; m < V; ++m)
#pragma unroll
        for (int g = 0; g < G; ++g) qn[m][g] = v4f{0.f, 0.f, 0.f, 0.f};
      auto picard_slice = [&](const int t) __device__ __attribute__((always_inline)) {
        float Q[V];
#pragma unroll
        for (int m = 0; m < V; ++m) Q[m] = sQi[(t * V + m) * B3 + node];
        stage_flux(Q, (float)exa_gl_weights[N][t]);
        __syncthreads();
        contract_lines_mfma<B, V, NW, 1>(sF, cK[0], cK[1], cK[2], lane, wave);
        __syncthreads();
        // phase 3: ir = rhs / w^3 = FLCoeff[t] luh - w_t sum_d C'_d, folded
        // into the next iterate (:292-299, n = t)
        const float fl_t = (float)exa_fl_coeff[N][t], wt = (float)exa_gl_weights[N][t];
        float aI[G];
#pragma unroll
        for (int g = 0; g < G; ++g) aI[g] = sIK[(t * G + g) * 4 + (lane & 3)];
#pragma unroll
        for (int m = 0; m < V; ++m) {
          const float c = sF[(0 * V + m) * B3 + node] + sF[(1 * V + m) * B3 + node] + sF[(2 * V + m) * B3 + node];
          const float ir = fl_t * q0i[m] - wt * c;
#pragma unroll
          for (int g = 0; g < G; ++g) qn[m][g] = mfma_own(aI[g], ir, qn[m][g]);
        }
        __syncthreads();  // the next slice rewrites sF
      };
      if constexpr (ROLLED) {
#pragma unroll 1
        for (int t = 0; t < B; ++t) picard_slice(t);
      } else {
#pragma unroll
        for (int t = 0; t < B; ++t) picard_slice(t);
      }
    }
    // squared residual (:303) and exit test (:314-318), float form
    float sq = 0.f, qq = 0.f;
#pragma unroll
    for (int l = 0; l < B; ++l)
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const float v = qn[m][l / 4][l % 4];
        const float dq = v - sQi[(l * V + m) * B3 + node];
        sq += dq * dq;
        qq += v * v;
      }
    if (!active) {
      sq = 0.f;
      qq = 0.f;
    }
    const float sq_res = block_sum<NT>(sq, sRed);
    const float qq_sum = block_sum<NT>(qq, sRed + NW);
    const bool done = sq_res < (float)kPicardTol2Mfma || sq_res < kPicardRelTol32Mfma * kPicardRelTol32Mfma * qq_sum;
    if (done || iter + 1 == MAX_ITER) {
      iters = iter + 1;
      break;  // sQi still holds lQi^k, the iterate the last fluxes were taken from
    }
#pragma unroll
    for (int l = 0; l < B; ++l)
#pragma unroll
      for (int m = 0; m < V; ++m) sQi[(l * V + m) * B3 + node] = qn[m][l / 4][l % 4];
  }

  // ---- epilogue: time average, extrapolation, volume integral ----
  const float cV = pack_operator<B>(lane, [](int n, int l) __device__ { return (float)exa_kxi[N][l][n]; });
  float cE;
  {
    const int n = lane >> 2, r = lane & 3;
    cE = (n < B && r < 2) ? (float)(r == 0 ? exa_fl_coeff[N][n] : exa_fr_coeff[N][n]) : 0.f;
  }
  const size_t cQb = cell * (size_t)(6 * B2 * D);
  const size_t cFb = cell * (size_t)(6 * B2 * V);
  // lQhi = sum_t lQi(t) w_t of the final iterate lQi^{k+1} = qn (:397-404)
#pragma unroll
  for (int m = 0; m < D; ++m) {
    float acc = 0.f;
    if (m < V) {
#pragma unroll
      for (int t = 0; t < B; ++t) acc += qn[m < V ? m : 0][t / 4][t % 4] * (float)exa_gl_weights[N][t];
    } else {
      acc = q0[m];
    }
    sF[m * B3 + node] = acc;
  }
  __syncthreads();
  extrapolate_mfma<B, V, D, false, NW>(sF, out_lQhbnd, cQb, cE, lane, wave);
  // Time-averaged fluxes of the last flux evaluation, i.e. of lQi^k (the
  // reference computes lFi before the final lQi update, :96-122 vs :306):
  // lFhi = sum_t w_t F(lQi^k(t)) (:367-396)
  if constexpr (!ACC_FH) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < V; ++m) fh[d][m] = 0.f;
  }
  auto average_slice = [&](const int t) __device__ __attribute__((always_inline)) {
    float Q[D];
#pragma unroll
    for (int m = 0; m < V; ++m) Q[m] = sQi[(t * V + m) * B3 + node];
#pragma unroll
    for (int m = V; m < D; ++m) Q[m] = q0[m];
    float F[3][V];
    PDE::flux(Q, F);
    const float wt = (float)exa_gl_weights[N][t];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < V; ++m) fh[d][m] += F[d][m] * wt;
  };
  if constexpr (ACC_FH) {
    // fh holds the time average of the last iteration's fluxes already
  } else if constexpr (ROLLED) {
#pragma unroll 1
    for (int t = 0; t < B; ++t) average_slice(t);
  } else {
#pragma unroll
    for (int t = 0; t < B; ++t) average_slice(t);
  }
  __syncthreads();  // the extrapolation has read sF
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int m = 0; m < V; ++m) sF[(d * V + m) * B3 + node] = fh[d][m];
  __syncthreads();
  extrapolate_mfma<B, V, V, true, NW>(sF, out_lFhbnd, cFb, cE, lane, wave);
  // volume integral (volumeIntegralNonlinear.cpph:136-225): the directional
  // contractions sum_n Kxi[l][n] lFhi_d(n) on the matrix cores, in place,
  // then each node sums its three directions with their weights
  if (out_lduh) {
    __syncthreads();  // the extrapolation has read sF
    contract_lines_mfma<B, V, NW, 1>(sF, cV, cV, cV, lane, wave);
    __syncthreads();
    // node weights and 1/dx formed here, not kept live across the Picard loop
    const float wx = (float)exa_gl_weights[N][x], wy = (float)exa_gl_weights[N][y],
                wz = (float)exa_gl_weights[N][z];
    const float ivx = 1.0f / (float)dxp[0], ivy = 1.0f / (float)dxp[1], ivz = 1.0f / (float)dxp[2];
    const float us[3] = {wz * wy * ivx, wz * wx * ivy, wy * wx * ivz};
    float* lduh = out_lduh + cell * (size_t)(B3 * V) + node * V;
    if (active) {
#pragma unroll
      for (int m = 0; m < V; ++m)
        lduh[m] = us[0] * sF[(0 * V + m) * B3 + node] + us[1] * sF[(1 * V + m) * B3 + node] +
                  us[2] * sF[(2 * V + m) * B3 + node];
    }
  }
  if (a.iters && tid == 0) a.iters[cell] = iters;
}

// Whether the matrix-core predictor is compiled for a solver (FP32-bulk).
template <int B, class PDE>
static constexpr bool mfma_predictor_compiled() {
  if constexpr (std::is_same_v<PDE, Euler<0>>) return B >= 3 && B <= 8;
  return false;
}

template <int B, class PDE>
hipError_t launch_predictor_mfma_t(const PredictorArgs& a, long ncells, hipStream_t s) {
  if constexpr (mfma_predictor_compiled<B, PDE>()) {
    const dim3 grid((unsigned)ncells), block(node_threads(B));
    hipLaunchKernelGGL((predictor_mfma_kernel<B, PDE, kFlux>), grid, block, 0, s, a);
    return hipGetLastError();
  } else {
    return hipErrorNotSupported;
  }
}

bool predictor_mfma_supported(const SolverDesc& sd) {
  return sd.precision == kFP32Bulk && sd.pde == kPdeEuler && sd.nvar == 5 && sd.npar == 0 && sd.terms == kFlux &&
         sd.order >= 2 && sd.order <= 7 && !sd.linear;
}

hipError_t launch_predictor_mfma(const SolverDesc& sd, const PredictorArgs& a, long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (!predictor_mfma_supported(sd)) return hipErrorNotSupported;
  switch (sd.order) {
    case 2: return launch_predictor_mfma_t<3, Euler<0>>(a, ncells, s);
    case 3: return launch_predictor_mfma_t<4, Euler<0>>(a, ncells, s);
    case 4: return launch_predictor_mfma_t<5, Euler<0>>(a, ncells, s);
    case 5: return launch_predictor_mfma_t<6, Euler<0>>(a, ncells, s);
    case 6: return launch_predictor_mfma_t<7, Euler<0>>(a, ncells, s);
    case 7: return launch_predictor_mfma_t<8, Euler<0>>(a, ncells, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace exa
