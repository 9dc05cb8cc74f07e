// exa_device.hpp -- device-side building blocks shared by the ADER-DG kernels:
// DG tables in constant memory, PDE functors, wavefront/workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace exa {

// ---------------------------------------------------------------------------
// Tables.  The reference's printed literals (DGMatrices.cpp, Gauss*Quadrature.cpp),
// [order][i] / [order][row][col], zero padded to 10.  Indexed with compile-time
// indices after unrolling, they become scalar constants (converted to the
// kernel's arithmetic type at compile time).
// ---------------------------------------------------------------------------
#define EXA_TABLE static __constant__
#include "dg_tables.inc"
#undef EXA_TABLE

// Stable time step size factors PNPM[order], stableTimeStepSize.cpph:33-34.
static __constant__ double exa_pnpm[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};

constexpr int kWave = 64;

__host__ __device__ constexpr int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Threads per workgroup for a kernel that maps one thread to one of the B^3
// nodes of a cell: a whole number of 64-wide wavefronts.
__host__ __device__ constexpr int node_threads(int B) { return round_up(B * B * B, kWave); }

// ---------------------------------------------------------------------------
// Reductions over a 64-wide wavefront (xor butterfly: every lane ends with the
// identical value because each level adds the same two operands).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Make a wavefront's LDS writes visible to all of its lanes (no workgroup
// barrier: other wavefronts of the workgroup do not take part).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over aligned segments of SEG lanes (SEG a power of two up to 64), the
// same butterfly restricted to the segment: for data that is zero outside the
// segment's first lanes it equals wave_sum bit for bit.
template <int SEG, typename T>
__device__ inline T segment_sum(T v) {
#pragma unroll
  for (int off = SEG / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Atomic minimum of a non-negative double kept as its bit pattern (for
// non-negative doubles the IEEE bit patterns order like the values).  The
// atomic returns and its result is consumed, so the wavefront waits until
// the atomic has been performed: the slot is read by the next kernel on the
// stream, and a small kernel launched right behind a producer has been seen
// to read a slot before the producer's last no-return atomic landed.
__device__ inline void atomic_min_positive(unsigned long long* slot, double v) {
  const unsigned long long old = atomicMin(slot, (unsigned long long)__double_as_longlong(v));
  asm volatile("" : : "v"(old));
}

// Reads a slot written by atomics of an earlier kernel and resets it, at the
// atomics' coherence point (not through a cache that may hold an old copy).
__device__ inline unsigned long long atomic_take(unsigned long long* slot, unsigned long long reset) {
  return atomicExch(slot, reset);
}

template <typename T>
__device__ inline T wave_min(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

// Workgroup sum; `red` holds NT/64 values of LDS.  All threads return the
// same value (partials are summed in wave order by every thread).
template <int NT, typename T>
__device__ inline T block_sum(T v, T* red) {
  constexpr int NW = NT / kWave;
  v = wave_sum(v);
  if (NW == 1) return v;
  const int wid = threadIdx.x / kWave;
  __syncthreads();
  red[wid] = v;  // every lane of the wave writes the same value: no divergent branch
  __syncthreads();
  T s = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) s += red[w];
  return s;
}

template <int NT, typename T>
__device__ inline T block_min(T v, T* red) {
  constexpr int NW = NT / kWave;
  v = wave_min(v);
  if (NW == 1) return v;
  const int wid = threadIdx.x / kWave;
  __syncthreads();
  red[wid] = v;  // every lane of the wave writes the same value: no divergent branch
  __syncthreads();
  T s = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) s = red[w] < s ? red[w] : s;
  return s;
}

// ---------------------------------------------------------------------------
// PDE functors: the user solver interface (UserSolverInterface.h:64-127)
// compiled for the device, for arithmetic type T.  Q holds V variables
// followed by P parameters.
// ---------------------------------------------------------------------------

// Compressible Euler, gamma = 1.4; the expression trees of the reference's
// kernel test solver (tests/kernels/c/GenericEulerKernelTest3D.cpp:26-82).
// NPAR trailing parameters are carried and ignored (the test solver has 1).
template <int NPAR>
struct Euler {
  static constexpr int V = 5;
  static constexpr int P = NPAR;

  // 1/rho: the correctly rounded division in double (the reference's
  // arithmetic); in the float form the hardware reciprocal (v_rcp_f32, about
  // 1 ulp), whose error is far below the float predictor's 1e-6 stopping bar.
  template <typename T>
  __device__ static inline T inverse(T x) {
    if constexpr (sizeof(T) == sizeof(float))
      return __builtin_amdgcn_rcpf(x);
    else
      return T(1.0) / x;
  }

  template <typename T>
  __device__ static inline void flux(const T* Q, T (*F)[V]) {
    const T GAMMA = T(1.4);
    const T irho = inverse(Q[0]);
    const T p = (GAMMA - T(1)) * (Q[4] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    F[0][0] = Q[1];
    F[0][1] = irho * Q[1] * Q[1] + p;
    F[0][2] = irho * Q[1] * Q[2];
    F[0][3] = irho * Q[1] * Q[3];
    F[0][4] = irho * Q[1] * (Q[4] + p);
    F[1][0] = Q[2];
    F[1][1] = irho * Q[2] * Q[1];
    F[1][2] = irho * Q[2] * Q[2] + p;
    F[1][3] = irho * Q[2] * Q[3];
    F[1][4] = irho * Q[2] * (Q[4] + p);
    F[2][0] = Q[3];
    F[2][1] = irho * Q[3] * Q[1];
    F[2][2] = irho * Q[3] * Q[2];
    F[2][3] = irho * Q[3] * Q[3] + p;
    F[2][4] = irho * Q[3] * (Q[4] + p);
  }

  // Largest |eigenvalue| in direction d (the eigenvalues are u_n-c, u_n x3, u_n+c).
  template <typename T>
  __device__ static inline T max_abs_eigenvalue(const T* Q, int d) {
    const T GAMMA = T(1.4);
    const T irho = T(1.0) / Q[0];
    const T p = (GAMMA - T(1)) * (Q[4] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    const T u_n = Q[d + 1] * irho;
    const T c = sqrt(GAMMA * p * irho);
    const T l0 = fabs(u_n - c), l1 = fabs(u_n), l4 = fabs(u_n + c);
    // std::max_element order: first maximum over (l0, l1, l1, l1, l4)
    T m = l0;
    m = (m < l1) ? l1 : m;
    m = (m < l4) ? l4 : m;
    return m;
  }

  template <typename T>
  __device__ static inline void algebraic_source(const T*, T* S) {
#pragma unroll
    for (int m = 0; m < V; ++m) S[m] = T(0);
  }

  // The test solver's NCP in the state its predictor test uses (zero).
  template <typename T>
  __device__ static inline void ncp(const T*, const T (*)[V], T* BgradQ) {
#pragma unroll
    for (int m = 0; m < V; ++m) BgradQ[m] = T(0);
  }

  // Linear-kernel interface: B_d grad_d Q of one direction (the linear
  // kernels' NCP, and B_d q = Bn^T q for the Riemann solver's coefficient
  // matrix); the test solver's are zero (GenericEulerKernelTest3D.cpp:108-132).
  template <int DIR, typename T>
  __device__ static inline void ncp_dir(const T*, const T*, T* out) {
#pragma unroll
    for (int m = 0; m < V; ++m) out[m] = T(0);
  }
};

// Linear elasticity, velocity-stress form: variables (sxx, syy, szz, sxy,
// syz, sxz, u, v, w), parameters (lambda, mu, rho), in the layout of the
// reference's elasticity test solver (ElasticityKernelTest2D.cpp:101-182,
// extended to 3-D by its z-block).  Pure non-conservative product: no flux,
// no source; the largest |eigenvalue| in any direction is cp.
struct Elastic {
  static constexpr int V = 9;
  static constexpr int P = 3;

  template <typename T>
  __device__ static inline void flux(const T*, T (*F)[V]) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < V; ++m) F[d][m] = T(0);
  }

  template <typename T>
  __device__ static inline T max_abs_eigenvalue(const T* Q, int) {
    const T lam = Q[9], mu = Q[10], rho0 = Q[11];
    return sqrt((lam + T(2) * mu) / rho0);
  }

  template <typename T>
  __device__ static inline void algebraic_source(const T*, T* S) {
#pragma unroll
    for (int m = 0; m < V; ++m) S[m] = T(0);
  }

  // B(Q) grad Q summed over the three directions (x + y) + z.
  template <typename T>
  __device__ static inline void ncp(const T* Q, const T (*g)[V], T* BgradQ) {
    const T lam = Q[9], mu = Q[10], irho = T(1) / Q[11];
    const T l2m = lam + T(2) * mu;
    BgradQ[0] = (-l2m * g[0][6] + -lam * g[1][7]) + -lam * g[2][8];
    BgradQ[1] = (-lam * g[0][6] + -l2m * g[1][7]) + -lam * g[2][8];
    BgradQ[2] = (-lam * g[0][6] + -lam * g[1][7]) + -l2m * g[2][8];
    BgradQ[3] = -mu * g[0][7] + -mu * g[1][6];
    BgradQ[4] = -mu * g[1][8] + -mu * g[2][7];
    BgradQ[5] = -mu * g[0][8] + -mu * g[2][6];
    BgradQ[6] = (-irho * g[0][0] + -irho * g[1][3]) + -irho * g[2][5];
    BgradQ[7] = (-irho * g[0][3] + -irho * g[1][1]) + -irho * g[2][4];
    BgradQ[8] = (-irho * g[0][5] + -irho * g[1][4]) + -irho * g[2][2];
  }

  // Linear-kernel interface (the linear kernels take B_d grad_d Q per
  // direction, spaceTimePredictorLinear.cpph:32): one direction's block of
  // the product above (ElasticityKernelTest2D.cpp:130-181, the z block being
  // the comment there).  B_d q is also Bn^T q for the coefficient matrix
  // Bn[l][k] = B_d(k, l) of ElasticityKernelTest2D.cpp:183-228.
  template <int DIR, typename T>
  __device__ static inline void ncp_dir(const T* Q, const T* g, T* out) {
    const T lam = Q[9], mu = Q[10], irho = T(1) / Q[11];
    const T l2m = lam + T(2) * mu;
    if constexpr (DIR == 0) {
      out[0] = -l2m * g[6];
      out[1] = -lam * g[6];
      out[2] = -lam * g[6];
      out[3] = -mu * g[7];
      out[4] = T(0);
      out[5] = -mu * g[8];
      out[6] = -irho * g[0];
      out[7] = -irho * g[3];
      out[8] = -irho * g[5];
    } else if constexpr (DIR == 1) {
      out[0] = -lam * g[7];
      out[1] = -l2m * g[7];
      out[2] = -lam * g[7];
      out[3] = -mu * g[6];
      out[4] = -mu * g[8];
      out[5] = T(0);
      out[6] = -irho * g[3];
      out[7] = -irho * g[1];
      out[8] = -irho * g[4];
    } else {
      out[0] = -lam * g[8];
      out[1] = -lam * g[8];
      out[2] = -l2m * g[8];
      out[3] = T(0);
      out[4] = -mu * g[7];
      out[5] = -mu * g[6];
      out[6] = -irho * g[5];
      out[7] = -irho * g[4];
      out[8] = -irho * g[2];
    }
  }
};

// ---------------------------------------------------------------------------
// PDE functor interface.  Required: V, P, and, templated on the arithmetic
// type T (float and double), flux(Q, F[3][V]), algebraic_source(Q, S) and
// ncp(Q, gradQ[3][V], BgradQ) as the terms the solver uses need them
// (UserSolverInterface.h:64-127), plus either the reference's
// eigenvalues(Q, d, lambda[V]) or max_abs_eigenvalue(Q, d).  Optional:
//   boundary_values(x[3], t, dt, faceIndex, normalNonZero, Fin, Qin, Fout, Qout)
//     (UserSolverInterface.h:171): the outside state and flux at a boundary
//     face node of a grid with EXA_BC_USER;
//   adjust_point_solution(x[3], t, dt, Q) (UserSolverInterface.h:149): applied
//     at every node after each update of a grid that has it switched on.
//
// The hooks are __device__ functions.  Each test below runs inside a
// __host__ __device__ constexpr function: a requires-expression met first in
// host code sees a __device__ function template as not callable, and the
// concept's cached answer is then false for the whole translation unit.
// ---------------------------------------------------------------------------
namespace hook_detect {
template <class PDE, typename T>
__host__ __device__ constexpr bool max_abs_eigenvalue() {
  return requires(const T* q, int d) { PDE::max_abs_eigenvalue(q, d); };
}
template <class PDE, typename T>
__host__ __device__ constexpr bool eigenvalues() {
  return requires(const T* q, int d, T* l) { PDE::eigenvalues(q, d, l); };
}
template <class PDE, typename T>
__host__ __device__ constexpr bool boundary_values() {
  return requires(const double* x, double t, int f, const T* fin, const T* qin, T* fo, T* qo) {
    PDE::boundary_values(x, t, t, f, f, fin, qin, fo, qo);
  };
}
template <class PDE>
__host__ __device__ constexpr bool adjust_point_solution() {
  return requires(const double* x, double t, double* q) { PDE::adjust_point_solution(x, t, t, q); };
}
}  // namespace hook_detect

template <class PDE, typename T>
concept HasMaxAbsEigenvalue = hook_detect::max_abs_eigenvalue<PDE, T>();
template <class PDE, typename T>
concept HasEigenvalues = hook_detect::eigenvalues<PDE, T>();
template <class PDE, typename T>
concept HasBoundaryValues = hook_detect::boundary_values<PDE, T>();
template <class PDE>
concept HasAdjustPointSolution = hook_detect::adjust_point_solution<PDE>();

// Largest |eigenvalue| in direction d: the PDE's own, or the maximum of
// |lambda| over the eigenvalues (stableTimeStepSize.cpph:44-52,
// riemannSolverNonlinear.cpph:68-78: the same value for finite lambda).
template <class PDE, typename T>
__device__ inline T pde_max_abs_eigenvalue(const T* Q, int d) {
  if constexpr (HasMaxAbsEigenvalue<PDE, T>) {
    return PDE::max_abs_eigenvalue(Q, d);
  } else {
    static_assert(HasEigenvalues<PDE, T>, "the PDE needs eigenvalues() or max_abs_eigenvalue()");
    T lambda[PDE::V];
    PDE::eigenvalues(Q, d, lambda);
    T m = fabs(lambda[0]);
#pragma unroll
    for (int v = 1; v < PDE::V; ++v) {
      const T a = fabs(lambda[v]);
      m = (m < a) ? a : m;
    }
    return m;
  }
}

// Stable time step of one node, stableTimeStepSize.cpph:37-57 (the caller takes
// the minimum).  maxEigenvalue = max over variables of |lambda|.
template <class PDE>
__device__ inline double node_stable_dt(const double* Q, const double invDx[3], double cfl_pnpm) {
  double denominator = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) denominator += pde_max_abs_eigenvalue<PDE>(Q, d) * invDx[d];
  return cfl_pnpm / denominator;
}

}  // namespace exa
