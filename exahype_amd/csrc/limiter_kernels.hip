// limiter_kernels.hip -- a-posteriori subcell limiter on the device (3-D,
// Euler, FP64): DG <-> FV projections, the discrete maximum principle with
// the physical admissibility check and the limiter status machine, FV ghost
// layers, the MUSCL-Hancock patch update, and the rollback / local
// recomputation of an irregular limiter-domain change.
//
// Reference functions restated (paths relative to the reference checkout):
//   LPM = ExaHyPE/kernels/LimiterProjectionMatrices.cpp
//   LIM = ExaHyPE/kernels/limiter/generic/c/Limiter.cpp
//   MH  = ExaHyPE/kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph
//   RUS = ExaHyPE/kernels/finitevolumes/riemannsolvers/c/rusanov.cpph
//   FVC = ExaHyPE/kernels/finitevolumes/commons/c/commons.cpph
//   LS  = ExaHyPE/exahype/solvers/LimitingADERDGSolver.cpp
//   AS  = ExaHyPE/exahype/solvers/ADERDGSolver.cpp
//
// Layout: a cell's FV patch is (S,S,S,V), S = L + 2g, L = 2N+1, g = 2,
// variable fastest (Limiter.cpp:37-38); every cell owns two patch buffers and
// cur[c] names the current one, so swapSolutionAndPreviousSolution is a flip.
// Edge and corner ghost cells are never written (they stay zero from the
// allocation, as FiniteVolumesSolver.cpp:477-481 requires).
//
// Sum-factorised tensor contractions replace the reference's naive B^3-term
// sums of the projections and of the Lobatto / subcell evaluations; results
// agree to rounding (the summation order differs).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>

#include "exa_device.hpp"
#include "fv_device.hpp"
#include "limiter.hpp"

namespace exa {

__constant__ LimTables c_lim;

static unsigned blocks256(long n) { return (unsigned)((n + 255) / 256); }

constexpr int kLimNT = 256;     // per-cell kernels
constexpr int kMusclNT = 448;   // MUSCL-Hancock patch kernel (LDS allows one per CU; seven wavefronts)

template <int N>
struct LimShape {
  static constexpr int B = N + 1, L = 2 * N + 1, G = kLimGhost, S = L + 2 * G;
  static constexpr int B3 = B * B * B, L3 = L * L * L, S3 = S * S * S;
  static constexpr int V = 5;
  static constexpr int PATCH = S3 * V;
  static constexpr int TT = V * B * L * L;  // tensor3 temporaries, either direction
};


// out[v][Z][Y][X] = sum in[v][z][y][x] T[z][Z] T[y][Y] T[x][X]; in/out/t1/t2
// in LDS, T (P x M, row = input index) in LDS.  Ends with a barrier.
template <int P, int M, int V, int NT>
__device__ inline void tensor3(const double* in, const double* T, double* t1, double* t2, double* out) {
  const int tid = threadIdx.x;
  for (int i = tid; i < V * P * P * M; i += NT) {  // t1[v][z][y][X]
    const int X = i % M, r = i / M;
    const double* row = in + r * P;
    double acc = 0.0;
#pragma unroll
    for (int x = 0; x < P; ++x) acc += row[x] * T[x * M + X];
    t1[i] = acc;
  }
  __syncthreads();
  for (int i = tid; i < V * P * M * M; i += NT) {  // t2[v][z][Y][X]
    const int X = i % M, Y = (i / M) % M, vz = i / (M * M);
    double acc = 0.0;
#pragma unroll
    for (int y = 0; y < P; ++y) acc += t1[(vz * P + y) * M + X] * T[y * M + Y];
    t2[i] = acc;
  }
  __syncthreads();
  for (int i = tid; i < V * M * M * M; i += NT) {  // out[v][Z][Y][X]
    const int YX = i % (M * M), Z = (i / (M * M)) % M, v = i / (M * M * M);
    double acc = 0.0;
#pragma unroll
    for (int z = 0; z < P; ++z) acc += t2[(v * P + z) * M * M + YX] * T[z * M + Z];
    out[i] = acc;
  }
  __syncthreads();
}

// Workgroup min and max of NV values per thread (all threads get the result).
template <int NV, int NT>
__device__ inline void block_minmax(double* mn, double* mx, double* red) {
  constexpr int NW = NT / kWave;
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    mn[v] = wave_min(mn[v]);
    mx[v] = -wave_min(-mx[v]);
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      red[(2 * v) * NW + wid] = mn[v];
      red[(2 * v + 1) * NW + wid] = mx[v];
    }
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double a = red[(2 * v) * NW], b = red[(2 * v + 1) * NW];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      a = dmin(a, red[(2 * v) * NW + w]);
      b = dmax(b, red[(2 * v + 1) * NW + w]);
    }
    mn[v] = a;
    mx[v] = b;
  }
  __syncthreads();
}

__device__ inline double* patch_ptr(const LimArgs& a, int c, int buf, int PATCH) {
  return a.lim + ((size_t)c * 2 + buf) * (size_t)PATCH;
}

// Local min/max of a cell's DG solution (LIM:210-368): GL nodes, Lobatto
// nodes, subcell means.  sQ[v][B^3] holds the solution; on return sSub holds
// the subcell means [v][L^3] (= the FV projection's interior).
template <int N>
__device__ inline void dg_local_minmax(const double* sQ, double* sLob, double* sSub, double* t1, double* t2,
                                       const double* sUh2lim, const double* sUh2lob, double* red, double* mn,
                                       double* mx) {
  using Sh = LimShape<N>;
  constexpr int V = Sh::V, B3 = Sh::B3, L3 = Sh::L3;
  tensor3<Sh::B, Sh::B, V, kLimNT>(sQ, sUh2lob, t1, t2, sLob);
  tensor3<Sh::B, Sh::L, V, kLimNT>(sQ, sUh2lim, t1, t2, sSub);
#pragma unroll
  for (int v = 0; v < V; ++v) {
    mn[v] = DBL_MAX;
    mx[v] = -DBL_MAX;
  }
  for (int i = threadIdx.x; i < B3; i += kLimNT)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      mn[v] = dmin(mn[v], sQ[v * B3 + i]);
      mx[v] = dmax(mx[v], sQ[v * B3 + i]);
      mn[v] = dmin(mn[v], sLob[v * B3 + i]);
      mx[v] = dmax(mx[v], sLob[v * B3 + i]);
    }
  for (int i = threadIdx.x; i < L3; i += kLimNT)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      mn[v] = dmin(mn[v], sSub[v * L3 + i]);
      mx[v] = dmax(mx[v], sSub[v * L3 + i]);
    }
  block_minmax<V, kLimNT>(mn, mx, red);
}

// Write the subcell values sSub[v][L^3] into a patch interior.
template <int N>
__device__ inline void store_interior(double* patch, const double* sSub) {
  using Sh = LimShape<N>;
  constexpr int L = Sh::L, G = Sh::G, S = Sh::S, V = Sh::V, L3 = Sh::L3;
  for (int i = threadIdx.x; i < L3 * V; i += kLimNT) {
    const int v = i % V, n = i / V;
    const int z = n / (L * L), y = (n / L) % L, x = n % L;
    patch[(((z + G) * S + (y + G)) * S + (x + G)) * V + v] = sSub[v * L3 + n];
  }
}

// Project a patch interior to the DG solution of the cell (LIM:69-112) and
// return the interior min/max (LIM:167-205) in mn/mx.
template <int N, int NT>
__device__ inline void fv_to_dg(const double* patch, double* luh, double* sIn, double* t1, double* t2, double* sOut,
                                const double* sLim2uh, double* red, double* mn, double* mx) {
  using Sh = LimShape<N>;
  constexpr int L = Sh::L, G = Sh::G, S = Sh::S, V = Sh::V, L3 = Sh::L3, B3 = Sh::B3;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    mn[v] = DBL_MAX;
    mx[v] = -DBL_MAX;
  }
  for (int i = threadIdx.x; i < L3; i += NT) {
    const int z = i / (L * L), y = (i / L) % L, x = i % L;
    const double* q = patch + (((z + G) * S + (y + G)) * S + (x + G)) * V;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      sIn[v * L3 + i] = q[v];
      mn[v] = dmin(mn[v], q[v]);
      mx[v] = dmax(mx[v], q[v]);
    }
  }
  __syncthreads();
  tensor3<Sh::L, Sh::B, V, NT>(sIn, sLim2uh, t1, t2, sOut);
  for (int i = threadIdx.x; i < B3 * V; i += NT) luh[i] = sOut[(i % V) * B3 + i / V];
  block_minmax<V, NT>(mn, mx, red);
}

template <int N>
__device__ inline void load_tables(double* sUh2lim, double* sUh2lob, double* sLim2uh) {
  using Sh = LimShape<N>;
  constexpr int B = Sh::B, L = Sh::L;
  for (int i = threadIdx.x; i < B * L; i += blockDim.x) {
    if (sUh2lim) sUh2lim[i] = c_lim.uh2lim[N][i];
    if (sLim2uh) sLim2uh[i] = c_lim.lim2uh[N][i];
  }
  for (int i = threadIdx.x; i < B * B; i += blockDim.x)
    if (sUh2lob) sUh2lob[i] = c_lim.uh2lob[N][i];
}

// ---------------------------------------------------------------------------
// Step begin: flags cleared, pst = st, list of FV-active cells (st >= A).
// ---------------------------------------------------------------------------
// flags[0..2] are per step; flags[3] counts steps with an irregular change
__global__ void lim_clear_flags_kernel(int* flags) {
  if (threadIdx.x < 3) flags[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void lim_begin_kernel(LimArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncells) return;
  const int s = a.st[c];
  a.pst[c] = s;
  if (s >= a.A) a.fv_list[atomicAdd(&a.flags[kLimFlagFvCount], 1)] = c;
}

// mergeSolutionMinMaxOnFace (LS:1654-1690): one thread per (face, observable).
__global__ __launch_bounds__(256) void lim_minmax_merge_kernel(LimArgs a, const FaceDesc* faces, long nfaces) {
  constexpr int V = 5;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nfaces * V) return;
  const FaceDesc f = faces[i / V];
  const int v = (int)(i % V);
  const size_t b1 = ((size_t)f.left * 6 + 2 * f.dir + 1) * V + v;
  const size_t b2 = ((size_t)f.right * 6 + 2 * f.dir) * V + v;
  const double lo = dmin(a.mn[b1], a.mn[b2]);
  const double hi = dmax(a.mx[b1], a.mx[b2]);
  a.mn[b1] = lo;
  a.mn[b2] = lo;
  a.mx[b1] = hi;
  a.mx[b2] = hi;
}

// ---------------------------------------------------------------------------
// Ghost layers of the listed FV cells: ghostLayerFilling (FVC:874-980) from
// face neighbours; at boundary faces boundaryLayerExtraction + transmissive
// boundaryValues with the mirror K = g-k-1 + ghostLayerFillingAtBoundary
// (FVC:499-721), keeping the extraction of the lower y/z faces that reads
// layer patchBegin for every k (FVC:558,588).  Persistent over the list.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(kLimNT) void lim_ghost_kernel(LimArgs a, int only_if_irregular) {
  using Sh = LimShape<N>;
  constexpr int L = Sh::L, G = Sh::G, S = Sh::S, V = Sh::V, PATCH = Sh::PATCH;
  if (only_if_irregular && !a.flags[kLimFlagIrregular]) return;
  const int count = a.flags[kLimFlagFvCount];
  for (int it = blockIdx.x; it < count; it += gridDim.x) {
    const int c = a.fv_list[it];
    double* own = patch_ptr(a, c, a.cur[c], PATCH);
    for (int i = threadIdx.x; i < 6 * G * L * L * V; i += kLimNT) {
      const int v = i % V;
      int r = i / V;
      const int b = r % L;
      r /= L;
      const int aa = r % L;
      r /= L;
      const int k = r % G;
      const int f = r / G;
      const int dir = f >> 1, high = f & 1;
      const int nb = a.nbr[c * 6 + f];
      const int gc = high ? G + L + k : k;  // ghost coordinate along dir
      int sc;                               // source coordinate along dir
      const double* src;
      if (nb >= 0) {
        src = patch_ptr(a, nb, a.cur[nb], PATCH);
        sc = high ? G + k : L + k;  // neighbour's pb + k / pe - g + k
      } else {
        src = own;
        const int K = G - k - 1;
        if (high)
          sc = G + L - G + K;  // extraction pe - g + K
        else
          sc = (dir == 0) ? G + K : G;  // extraction pb + K (x) / pb (y, z)
      }
      // transverse (aa, b): x faces (z,y), y faces (z,x), z faces (y,x)
      int zd, yd, xd, zs, ys, xs;
      if (dir == 0) {
        zd = zs = G + aa;
        yd = ys = G + b;
        xd = gc;
        xs = sc;
      } else if (dir == 1) {
        zd = zs = G + aa;
        xd = xs = G + b;
        yd = gc;
        ys = sc;
      } else {
        yd = ys = G + aa;
        xd = xs = G + b;
        zd = gc;
        zs = sc;
      }
      own[((zd * S + yd) * S + xd) * V + v] = src[((zs * S + ys) * S + xs) * V + v];
    }
  }
}

// ---------------------------------------------------------------------------
// MUSCL-Hancock update of one patch (MH:49-369, flux-only PDE) followed by
// the projection to DG (LIM:69-112) and the interior min/max (LIM:167-205).
// Persistent over the FV list.  Only the cells whose evolved face states feed
// a Riemann problem are evolved: the interior and the first ghost layer of
// the six faces (the reference also evolves other ghost cells whose states
// no Riemann problem reads).
// GODUNOV: the first-order update (GOD:35-143) -- the face states are the
// cell values themselves, the face fluxes and their accumulation order are
// those of MUSCL-Hancock (the two reference kernels share both).
// ---------------------------------------------------------------------------
template <int N, bool GODUNOV>
__global__ __launch_bounds__(kMusclNT) void lim_fv_update_kernel(LimArgs a, int only_if_irregular) {
  using Sh = LimShape<N>;
  constexpr int L = Sh::L, G = Sh::G, S = Sh::S, V = Sh::V, PATCH = Sh::PATCH, L3 = Sh::L3, B = Sh::B;
  constexpr int NF = L * L * (L + 1);  // faces per direction
  constexpr int NC = L3 + 6 * L * L;   // cells to evolve
  __shared__ double sP[PATCH];
  __shared__ double sF[3 * 2 * V * NF];  // [d][side][v][a][b][f]; flux into side 0
  __shared__ double sLim2uh[L * B];
  __shared__ double red[2 * V * (kMusclNT / kWave)];
  if (only_if_irregular && !a.flags[kLimFlagIrregular]) return;
  load_tables<N>(nullptr, nullptr, sLim2uh);
  const int count = a.flags[kLimFlagFvCount];
  const double dt = a.st_state->dt;
  const double h[3] = {a.dx[0] / L, a.dx[1] / L, a.dx[2] / L};
  const double ih[3] = {1. / h[0], 1. / h[1], 1. / h[2]};
  const double cdt[3] = {dt * ih[0], dt * ih[1], dt * ih[2]};
  auto fidx = [](int d, int side, int v, int fa, int fb, int f) {
    return ((((d * 2 + side) * V + v) * L + fa) * L + fb) * (L + 1) + f;
  };
  for (int it = blockIdx.x; it < count; it += gridDim.x) {
    const int c = a.fv_list[it];
    const int ob = a.cur[c];
    const double* old = patch_ptr(a, c, ob, PATCH);
    double* nw = patch_ptr(a, c, ob ^ 1, PATCH);
    __syncthreads();  // the previous patch's LDS use is over
    for (int i = threadIdx.x; i < PATCH; i += kMusclNT) sP[i] = old[i];
    __syncthreads();
    // evolve the half time step (MH:77-261)
    for (int j = threadIdx.x; j < NC; j += kMusclNT) {
      int z, y, x, sheet = -1;
      if (j < L3) {
        z = G + j / (L * L);
        y = G + (j / L) % L;
        x = G + j % L;
      } else {
        const int k = j - L3;
        sheet = k / (L * L);
        const int r = k % (L * L), fa = G + r / L, fb = G + r % L;
        const int gc = (sheet & 1) ? G + L : G - 1;
        if ((sheet >> 1) == 0) {
          z = fa;
          y = fb;
          x = gc;
        } else if ((sheet >> 1) == 1) {
          z = fa;
          x = fb;
          y = gc;
        } else {
          y = fa;
          x = fb;
          z = gc;
        }
      }
      const int p0 = (z * S + y) * S + x;
      double u[V], slope[3][V], Qt[V];
#pragma unroll
      for (int v = 0; v < V; ++v) u[v] = sP[p0 * V + v];
      if constexpr (!GODUNOV) {
        const int step[3] = {1, S, S * S};
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int v = 0; v < V; ++v)
            slope[d][v] = minmod(sP[(p0 + step[d]) * V + v] - u[v], u[v] - sP[(p0 - step[d]) * V + v]) * ih[d];
#pragma unroll
        for (int v = 0; v < V; ++v) Qt[v] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double w[V], F[V];
#pragma unroll
          for (int v = 0; v < V; ++v) w[v] = u[v] - 0.5 * h[d] * slope[d][v];
          euler_flux_dir(w, d, F);
#pragma unroll
          for (int v = 0; v < V; ++v) Qt[v] += F[v] * ih[d];
#pragma unroll
          for (int v = 0; v < V; ++v) w[v] = u[v] + 0.5 * h[d] * slope[d][v];
          euler_flux_dir(w, d, F);
#pragma unroll
          for (int v = 0; v < V; ++v) Qt[v] -= F[v] * ih[d];
        }
      }
      // face states into the face table: the evolved wL/wR (+= dt/2 Qt), or
      // the cell value itself for the first-order scheme
      auto face_state = [&](int d, int v, double sign) -> double {
        if constexpr (GODUNOV) return u[v];
        return (u[v] + sign * 0.5 * h[d] * slope[d][v]) + 0.5 * dt * Qt[v];
      };
      const int co[3] = {x - G, y - G, z - G};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int fa = (d == 2) ? co[1] : co[2];  // x,y faces: a = z; z faces: a = y
        const int fb = (d == 0) ? co[1] : co[0];  // x faces: b = y; y,z faces: b = x
        const bool needL = (sheet < 0) || (sheet == 2 * d + 1);  // wL feeds the face on its left
        const bool needR = (sheet < 0) || (sheet == 2 * d);      // wR feeds the face on its right
        if (needL) {
#pragma unroll
          for (int v = 0; v < V; ++v) sF[fidx(d, 1, v, fa, fb, co[d])] = face_state(d, v, -1.0);
        }
        if (needR) {
#pragma unroll
          for (int v = 0; v < V; ++v) sF[fidx(d, 0, v, fa, fb, co[d] + 1)] = face_state(d, v, 1.0);
        }
      }
    }
    __syncthreads();
    // Rusanov on every face (RUS:98-156), flux stored over the left state;
    // dt_max_allowed = min over faces of cflFactor/3 h_d / s_max (MH:267-329)
    double dt_allowed = DBL_MAX;
    for (int j = threadIdx.x; j < 3 * NF; j += kMusclNT) {
      const int d = j / NF, r = j % NF;
      const int fa = r / (L * (L + 1)), fb = (r / (L + 1)) % L, f = r % (L + 1);
      double qL[V], qR[V], FL[V], FR[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        qL[v] = sF[fidx(d, 0, v, fa, fb, f)];
        qR[v] = sF[fidx(d, 1, v, fa, fb, f)];
      }
      const double smax = dmax(Euler<0>::max_abs_eigenvalue(qR, d), Euler<0>::max_abs_eigenvalue(qL, d));
      const double dta = (GODUNOV ? 0.9 : 0.99) / 3 * h[d] / smax;  // FV solver CFL / MH:29
      dt_allowed = dta < dt_allowed ? dta : dt_allowed;
      euler_flux_dir(qL, d, FL);
      euler_flux_dir(qR, d, FR);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        double fn = 0.5 * smax * (qL[v] - qR[v]);
        fn += 0.5 * (FL[v] + FR[v]);
        sF[fidx(d, 0, v, fa, fb, f)] = fn;
      }
    }
    if (a.fv_dt_max) {
      const double m = wave_min(dt_allowed);
      if ((threadIdx.x & (kWave - 1)) == 0) atomic_min_positive(a.fv_dt_max, m);
    }
    __syncthreads();
    // flux differences in the reference's accumulation order (MH:267-329)
    for (int j = threadIdx.x; j < L3; j += kMusclNT) {
      const int iz = j / (L * L), iy = (j / L) % L, ix = j % L;
      const int p0 = ((iz + G) * S + (iy + G)) * S + (ix + G);
      double q[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        q[v] = sP[p0 * V + v];
        q[v] += cdt[0] * sF[fidx(0, 0, v, iz, iy, ix)];
        q[v] -= cdt[0] * sF[fidx(0, 0, v, iz, iy, ix + 1)];
        q[v] += cdt[1] * sF[fidx(1, 0, v, iz, ix, iy)];
        q[v] -= cdt[1] * sF[fidx(1, 0, v, iz, ix, iy + 1)];
        q[v] += cdt[2] * sF[fidx(2, 0, v, iy, ix, iz)];
        q[v] -= cdt[2] * sF[fidx(2, 0, v, iy, ix, iz + 1)];
        nw[p0 * V + v] = q[v];
        sP[p0 * V + v] = q[v];
      }
    }
    __syncthreads();
    // projection to DG and the patch min/max (sF is free now)
    double mn[V], mx[V];
    fv_to_dg<N, kMusclNT>(sP, a.luh + (size_t)c * Sh::B3 * V, sF, sF + L3 * V, sF + 2 * L3 * V, sF + 3 * L3 * V,
                          sLim2uh, red, mn, mx);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        a.limmm[(size_t)c * 2 * V + v] = mn[v];
        a.limmm[(size_t)c * 2 * V + V + v] = mx[v];
      }
      a.cur[c] = (uint8_t)(ob ^ 1);
    }
  }
}

// ---------------------------------------------------------------------------
// SolutionUpdate of every cell (LS:843-937 + AS:2185-2248): DG update for
// status < A (with the previous solution, and for helper cells a fresh FV
// projection into the other patch buffer), then startNewTimeStep's dt, the
// DMP + admissibility (LS:996-1029, 1045-1065), the limiter min/max of
// FV-active cells, the new status (LS:955-994) and patch allocation
// (LS:940-953).
//
// One wavefront per cell (lane = node; N <= 3, so B^3 <= 64), four cells per
// workgroup, synchronised per wavefront only.  The local min/max runs the
// tensor contractions one variable at a time (tensor3_rows: the same sums in
// the same order as tensor3), so a wavefront's LDS holds one variable's
// temporaries; the Lobatto and subcell values are folded into the extrema in
// registers, and the subcell values of the FV projection are formed again,
// per variable, only for the cells that need the projection.
// ---------------------------------------------------------------------------
//
// Each lane owns a row of a phase: in phase 1 lane (z,y) forms t1[z][y][X]
// for all X, in phase 2 lane (z,X) forms t2[z][Y][X] for all Y, in phase 3
// lane (Y,X) forms out[Z][Y][X] for all Z and hands each value to
// emit(Z, Y*M + X, value) in registers.  A lane reads its P inputs once for
// M outputs, and the table index is uniform in every loop, so the table
// entries are scalar loads from constant memory: the status kernel is bound
// by LDS instruction issue, and an output per lane (P input and P table reads
// each) took three times the LDS instructions.  Every output is the sum
// 0 + in[0]*T[0][.] + ... + in[P-1]*T[P-1][.] of tensor3, in that order.
enum { kTabUh2Lim, kTabUh2Lob, kTabLim2Uh };  // the table of tensor3_rows

template <int N, int TAB, int P, int M, class Emit>
__device__ inline void tensor3_rows(const double* in, double* t1, double* t2, int lane, Emit&& emit) {
  static_assert(P * P <= kWave && P * M <= kWave && M * M <= kWave, "one row per lane");
  const double* const T =
      TAB == kTabUh2Lob ? c_lim.uh2lob[N] : (TAB == kTabLim2Uh ? c_lim.lim2uh[N] : c_lim.uh2lim[N]);
  if (lane < P * P) {  // t1[z][y][X], lane = z*P + y
    double row[P];
#pragma unroll
    for (int x = 0; x < P; ++x) row[x] = in[lane * P + x];
#pragma unroll
    for (int X = 0; X < M; ++X) {
      double acc = 0.0;
#pragma unroll
      for (int x = 0; x < P; ++x) acc += row[x] * T[x * M + X];
      t1[lane * M + X] = acc;
    }
  }
  wave_lds_sync();
  if (lane < P * M) {  // t2[z][Y][X], lane = z*M + X
    const int z = lane / M, X = lane % M;
    double col[P];
#pragma unroll
    for (int y = 0; y < P; ++y) col[y] = t1[(z * P + y) * M + X];
#pragma unroll
    for (int Y = 0; Y < M; ++Y) {
      double acc = 0.0;
#pragma unroll
      for (int y = 0; y < P; ++y) acc += col[y] * T[y * M + Y];
      t2[(z * M + Y) * M + X] = acc;
    }
  }
  wave_lds_sync();
  if (lane < M * M) {  // out[Z][Y][X], lane = Y*M + X
    double col[P];
#pragma unroll
    for (int z = 0; z < P; ++z) col[z] = t2[z * M * M + lane];
#pragma unroll
    for (int Z = 0; Z < M; ++Z) {
      double acc = 0.0;
#pragma unroll
      for (int z = 0; z < P; ++z) acc += col[z] * T[z * M + Z];
      emit(Z, lane, acc);
    }
  }
  wave_lds_sync();  // t1, t2 are rewritten by the next call
}

constexpr int kStatusCells = 4;  // cells (wavefronts) per workgroup of the status kernel

template <int N>
// Registers capped for five wavefronts per SIMD (109 -> at most 102 VGPRs,
// no scratch; 2.83 -> 2.73 ms at 81^3 N=3), six spill.
__global__ __launch_bounds__(kStatusCells * kWave, 5) void lim_update_status_kernel(LimArgs a) {
  using Sh = LimShape<N>;
  constexpr int B = Sh::B, B2 = B * B, B3 = Sh::B3, L = Sh::L, L3 = Sh::L3, V = Sh::V, PATCH = Sh::PATCH;
  static_assert(B3 <= kWave, "one node per lane");
  constexpr int T1 = B * B * L, T2 = B * L * L;  // the subcell transform's; the Lobatto one's are smaller
  constexpr int PER = V * B3 + T1 + T2 + 12 * V;  // sQ, t1, t2, face extrema
  __shared__ double sWave[kStatusCells * PER];
  // no workgroup barrier: each wavefront works in its own part of sWave
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  const int c = blockIdx.x * kStatusCells + wid;
  if (c >= a.ncells) return;
  double* const sQ = sWave + wid * PER;
  double* const t1 = sQ + V * B3;
  double* const t2 = t1 + T1;
  double* const sBnd = t2 + T2;
  // loads whose latency the DG update hides: the merged face extrema (one
  // lane per face-block entry) and the neighbours' previous statuses (one
  // lane per face)
  const double bnd = lane < 6 * V ? a.mn[(size_t)c * 6 * V + lane]
                                  : (lane >= 32 && lane < 32 + 6 * V ? a.mx[(size_t)c * 6 * V + lane - 32] : 0.0);
  const int nbf = lane < 6 ? a.nbr[c * 6 + lane] : -1;
  const int onb = nbf >= 0 ? a.pst[nbf] : 0;
  const int itc0 = a.itc[c];
  const bool had_patch = a.has_patch[c] != 0;
  const int cur0 = a.cur[c];
  const int s = a.pst[c];
  const bool active = lane < B3;
  const int node = active ? lane : 0;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  double* luh = a.luh + (size_t)c * B3 * V + node * V;
  double q[V];
#pragma unroll
  for (int m = 0; m < V; ++m) q[m] = luh[m];
  const double invDx[3] = {1.0 / a.dx[0], 1.0 / a.dx[1], 1.0 / a.dx[2]};
  if (s < a.A) {
    // surface integral (surfaceIntegralNonlinear.cpph:21-90) on the volume integral
    constexpr int NO = B - 1;
    const double wx = exa_gl_weights[NO][x], wy = exa_gl_weights[NO][y], wz = exa_gl_weights[NO][z];
    const double* du_p = a.lduh + (size_t)c * B3 * V + node * V;
    const double* lF = a.lFhbnd + (size_t)c * 6 * B2 * V;
    double du[V];
#pragma unroll
    for (int m = 0; m < V; ++m) du[m] = du_p[m];
    {
      const double us = wz * wy * invDx[0], fr = exa_fr_coeff[NO][x], fl = exa_fl_coeff[NO][x];
#pragma unroll
      for (int m = 0; m < V; ++m)
        du[m] -= (lF[(1 * B2 + z * B + y) * V + m] * fr - lF[(0 * B2 + z * B + y) * V + m] * fl) * us;
    }
    {
      const double us = wz * wx * invDx[1], fr = exa_fr_coeff[NO][y], fl = exa_fl_coeff[NO][y];
#pragma unroll
      for (int m = 0; m < V; ++m)
        du[m] -= (lF[(3 * B2 + z * B + x) * V + m] * fr - lF[(2 * B2 + z * B + x) * V + m] * fl) * us;
    }
    {
      const double us = wy * wx * invDx[2], fr = exa_fr_coeff[NO][z], fl = exa_fl_coeff[NO][z];
#pragma unroll
      for (int m = 0; m < V; ++m)
        du[m] -= (lF[(5 * B2 + y * B + x) * V + m] * fr - lF[(4 * B2 + y * B + x) * V + m] * fl) * us;
    }
    const double updateSize = a.st_state->dt / (wz * wy * wx);
    if (active) {
      double* prev = a.luh_prev + (size_t)c * B3 * V + node * V;
#pragma unroll
      for (int m = 0; m < V; ++m) {
        prev[m] = q[m];
        q[m] += du[m] * updateSize;
        luh[m] = q[m];
      }
    }
  }
  if (active) {
#pragma unroll
    for (int m = 0; m < V; ++m) sQ[m * B3 + node] = q[m];
  }
  // ensureNoUnrequiredLimiterPatchIsAllocatedOnComputeCell (status before the update)
  bool patch = had_patch;
  if (s == 0) patch = false;
  // startNewTimeStep: stableTimeStepSize of the new solution.  The slot's
  // returning atomic is waited for at the end of the kernel (its result is
  // consumed there), so its latency overlaps the min/max work.
  unsigned long long dt_old = 0;
  {
    double dtn = active ? node_stable_dt<Euler<0>>(q, invDx, a.cfl_pnpm) : DBL_MAX;
    dtn = dtn < DBL_MAX ? dtn : DBL_MAX;
    dtn = wave_min(dtn);
    if (lane == 0) {
      dt_old = atomicMin(&a.st_state->dt_slots[c % kDtSlots], (unsigned long long)__double_as_longlong(dtn));
      a.dt_cell[c] = dtn;
    }
  }
  wave_lds_sync();
  // local min/max (LIM:210-368): Gauss-Legendre nodes, Lobatto nodes,
  // subcell means, one variable at a time
  double mn[V], mx[V];
#pragma unroll 1
  for (int v = 0; v < V; ++v) {
    double lo = DBL_MAX, hi = -DBL_MAX;
    if (active) {  // B3 <= kWave: one node per lane
      lo = dmin(lo, sQ[v * B3 + lane]);
      hi = dmax(hi, sQ[v * B3 + lane]);
    }
    auto fold = [&](int, int, double u) {
      lo = dmin(lo, u);
      hi = dmax(hi, u);
    };
    tensor3_rows<N, kTabUh2Lob, B, B>(sQ + v * B3, t1, t2, lane, fold);  // Lobatto nodes
    tensor3_rows<N, kTabUh2Lim, B, L>(sQ + v * B3, t1, t2, lane, fold);  // subcell means
    lo = wave_min(lo);
    hi = -wave_min(-hi);
    // a register array indexed by the rolled v would live in scratch memory:
    // select into the unrolled slots instead
#pragma unroll
    for (int w = 0; w < V; ++w) {
      mn[w] = (w == v) ? lo : mn[w];
      mx[w] = (w == v) ? hi : mx[w];
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      a.loc_mm[(size_t)c * 2 * V + v] = mn[v];
      a.loc_mm[(size_t)c * 2 * V + V + v] = mx[v];
    }
  }
  // DMP (LIM:114-165) against the merged boundary extrema, staged in LDS
  double* bmn = a.mn + (size_t)c * 6 * V;
  double* bmx = a.mx + (size_t)c * 6 * V;
  double* const sSub = sBnd;
  if (lane < 6 * V) sSub[lane] = bnd;
  if (lane >= 32 && lane < 32 + 6 * V) sSub[6 * V + lane - 32] = bnd;
  wave_lds_sync();
  bool ok = true;
  double bMin[V], bMax[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    double lo = sSub[v], hi = sSub[6 * V + v];
#pragma unroll
    for (int f = 1; f < 6; ++f) {
      lo = dmin(lo, sSub[f * V + v]);
      hi = dmax(hi, sSub[6 * V + f * V + v]);
    }
    double sd = (hi - lo) * a.scaling;
    sd = dmax(sd, a.relax);
    if ((mn[v] < (lo - sd)) || (mx[v] > (hi + sd))) ok = false;
    bMin[v] = lo;
    bMax[v] = hi;
  }
  // isPhysicallyAdmissible of the Euler solver (face-0 block = local minima)
  if (ok) ok = !(mn[0] <= 0.0) && !(mn[4] <= 0.0);
  if (lane == 0) {  // every lane has its face-block values in registers
    if (s >= a.A) {  // determineLimiterMinAndMax: patch extrema on all faces
      double pmn[V], pmx[V];  // loaded once, before any store
#pragma unroll
      for (int v = 0; v < V; ++v) {
        pmn[v] = a.limmm[(size_t)c * 2 * V + v];
        pmx[v] = a.limmm[(size_t)c * 2 * V + V + v];
      }
#pragma unroll
      for (int f = 0; f < 6; ++f)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          bmn[f * V + v] = pmn[v];
          bmx[f * V + v] = pmx[v];
        }
    } else {
      for (int v = 0; v < V; ++v) {
        bmn[v] = mn[v];
        bmx[v] = mx[v];
        bmn[V + v] = bMin[v];
        bmx[V + v] = bMax[v];
      }
    }
  }
  // determineLimiterStatusAfterSolutionUpdate (LS:955-994); the face-wise
  // statuses come from mergeWithLimiterStatus (AS:2547-2563) of the merge phase
  int ns;
  int itc = itc0;
  if (!ok) {
    itc = a.cure + 1;
    ns = a.T;
    if (s <= a.A && lane == 0) a.flags[kLimFlagIrregular] = 1;
  } else if (s >= a.T) {
    ns = a.T;
    itc -= 1;
    if (itc == 0) {
      ns = a.T - 1;
      itc = a.cure + 1;
    }
  } else {
    // the six face statuses (mergeWithLimiterStatus) from the lanes that
    // loaded them, their maximum over the wavefront
    const int own = s < a.T ? s : a.T;
    const int o = onb < a.T ? onb : a.T;
    const int fs = nbf >= 0 ? (own > o ? own : o) - 1 : 0;
    ns = (int)(-wave_min(-(double)(fs > 0 ? fs : 0)));
  }
  // helper cells: swap + project (LS:861-873); new patches: project (LS:940-953)
  const bool helper = s > 0 && s < a.A;
  const bool fresh = !patch && ns > 0;
  const int buf = cur0 ^ (helper ? 1 : 0);
  if (helper || (fresh && ns < a.A)) {
    double* const dst = patch_ptr(a, c, buf, PATCH);
    constexpr int G = Sh::G, S = Sh::S;
    const int yy = lane / L, xx = lane % L;  // the lane's subcell row (lanes < L*L)
#pragma unroll 1
    for (int v = 0; v < V; ++v)
      tensor3_rows<N, kTabUh2Lim, B, L>(sQ + v * B3, t1, t2, lane, [&](int zz, int, double u) {
        dst[(((zz + G) * S + (yy + G)) * S + (xx + G)) * V + v] = u;
      });
  }
  if (lane == 0) {
    a.st[c] = ns;
    a.itc[c] = itc;
    a.has_patch[c] = (uint8_t)((patch || fresh) ? 1 : 0);
    a.cur[c] = (uint8_t)buf;
  }
  asm volatile("" : : "v"(dt_old));  // the dt atomic has been performed
}

// ---------------------------------------------------------------------------
// Irregular change: status spreading sweep (LimiterStatusSpreading.cpp +
// LS:392-405) in -> out over interior faces; ensureRequiredLimiterPatchIsAllocated.
// Sweep 0 also clears the dt slots and the FV list (they are rebuilt by the
// rollback and the recomputation).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void lim_spread_kernel(LimArgs a, const int* in, int* out, int sweep) {
  if (!a.flags[kLimFlagIrregular]) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (sweep == 0) {
    if (c < kDtSlots) a.st_state->dt_slots[c] = kDtSlotEmpty;
  }
  if (c >= a.ncells) return;
  const int s = in[c];
  int ns;
  if (s < a.T) {
    ns = 0;
    for (int f = 0; f < 6; ++f) {
      const int nb = a.nbr[c * 6 + f];
      if (nb < 0) continue;
      const int o = in[nb] < a.T ? in[nb] : a.T;
      const int fs = (s > o ? s : o) - 1;
      ns = fs > ns ? fs : ns;
    }
  } else {
    ns = a.T;
    a.itc[c] = a.cure + 1;
  }
  out[c] = ns;
  if (ns > 0) a.has_patch[c] = 1;
}

// recomputePredictorLocally's selection (LS:1438-1464) after an irregular
// change in fused time stepping, as a status array for the predictor's gate:
// 0 for the cells to predict again (status < troubled and the solution
// recomputed from an FV patch: FV-active neighbours, cells troubled in the
// previous step), troubled for the rest.
__global__ __launch_bounds__(256) void lim_repredict_mask_kernel(LimArgs a, int* out) {
  if (!a.flags[kLimFlagIrregular]) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncells) return;
  const int s = a.st[c];
  out[c] = (s < a.T && (s >= a.A || a.pst[c] >= a.T)) ? 0 : a.T;
}

__global__ void lim_reset_fv_count_kernel(int* flags, int only_if_irregular) {
  if (only_if_irregular && !flags[kLimFlagIrregular]) return;
  if (threadIdx.x == 0) {
    flags[kLimFlagFvCount] = 0;
    flags[kLimFlagRcCount] = 0;
  }
}

// reinitialiseSolversLocally (LS:1343-1381), list part (thread per cell):
// the FV list and the list of cells with status > 0 (rolled back), the cure
// counter of troubled cells; cells with status 0 release their patch.
__global__ __launch_bounds__(256) void lim_rollback_lists_kernel(LimArgs a) {
  if (!a.flags[kLimFlagIrregular]) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncells) return;
  const int s = a.st[c];
  if (s >= a.A) a.fv_list[atomicAdd(&a.flags[kLimFlagFvCount], 1)] = c;
  if (s > 0) {
    a.rc_list[atomicAdd(&a.flags[kLimFlagRcCount], 1)] = c;
    if (s >= a.T) a.itc[c] = a.cure + 1;
  } else {
    a.has_patch[c] = 0;
  }
}

// reinitialiseSolversLocally, work part: roll the listed cells (status > 0)
// back to the previous time step -- previously FV-active cells to their
// previous patch and its DG projection, the others to their previous DG
// solution and its FV projection.  Persistent over the list.
template <int N>
__global__ __launch_bounds__(kLimNT) void lim_rollback_kernel(LimArgs a) {
  using Sh = LimShape<N>;
  constexpr int B3 = Sh::B3, L3 = Sh::L3, V = Sh::V, PATCH = Sh::PATCH, L = Sh::L;
  __shared__ double sQ[V * B3], sSub[V * L3];
  __shared__ double t1[Sh::TT], t2[Sh::TT];
  __shared__ double sUh2lim[Sh::B * L], sLim2uh[Sh::B * L];
  __shared__ double red[2 * V * (kLimNT / kWave)];
  if (!a.flags[kLimFlagIrregular]) return;
  const int count = a.flags[kLimFlagRcCount];
  load_tables<N>(sUh2lim, nullptr, sLim2uh);
  for (int it = blockIdx.x; it < count; it += gridDim.x) {
    const int c = a.rc_list[it];
    __syncthreads();  // tables loaded / the previous cell's LDS use is over
    if (a.pst[c] >= a.A) {
      const int buf = a.cur[c] ^ 1;
      double mn[V], mx[V];
      fv_to_dg<N, kLimNT>(patch_ptr(a, c, buf, PATCH), a.luh + (size_t)c * B3 * V, sSub, t1, t2, sQ, sLim2uh, red, mn,
                          mx);
      if (threadIdx.x == 0) a.cur[c] = (uint8_t)buf;
    } else {
      double* luh = a.luh + (size_t)c * B3 * V;
      double* prev = a.luh_prev + (size_t)c * B3 * V;
      for (int i = threadIdx.x; i < B3 * V; i += kLimNT) {
        const double t = luh[i];
        luh[i] = prev[i];
        prev[i] = t;
        sQ[(i % V) * B3 + i / V] = luh[i];
      }
      __syncthreads();
      tensor3<Sh::B, L, V, kLimNT>(sQ, sUh2lim, t1, t2, sSub);
      store_interior<N>(patch_ptr(a, c, a.cur[c], PATCH), sSub);
    }
  }
}

// LocalRecomputation enterCell (LocalRecomputation.cpp:236-285) after the FV
// cells were recomputed: recomputeSolutionLocally of the helper cells
// (LS:1396-1436), startNewTimeStep, determineMinAndMax.  Cells with status 0
// were not rolled back: their solution is the one the status kernel saw, so
// its dt candidate and local min/max are reused (lim_recompute_light_kernel);
// this kernel walks the cells with status > 0 (the rollback's list).
//
// One wavefront per cell (lane = node, N <= 3), four cells per workgroup,
// the contractions row-owned as in the status kernel (tensor3_rows, the same
// sums in the same order as tensor3); each wavefront's LDS holds the cell's
// solution [var][node] and one variable's temporaries.
constexpr int kRecomputeCells = 4;  // cells (wavefronts) per workgroup of the recompute kernel

template <int N>
__global__ __launch_bounds__(kRecomputeCells * kWave) void lim_recompute_kernel(LimArgs a) {
  using Sh = LimShape<N>;
  constexpr int B = Sh::B, B2 = B * B, B3 = Sh::B3, L = Sh::L, L3 = Sh::L3, V = Sh::V, PATCH = Sh::PATCH;
  constexpr int G = Sh::G, S = Sh::S;
  static_assert(B3 <= kWave, "one node per lane");
  constexpr int TT = B * L * L;              // the largest phase temporary of the three transforms
  constexpr int PER = V * B3 + 2 * TT + L3;  // sQ, t1, t2, one variable's subcell values
  __shared__ double sWave[kRecomputeCells * PER];
  if (!a.flags[kLimFlagIrregular]) return;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  double* const sQ = sWave + wid * PER;
  double* const t1 = sQ + V * B3;
  double* const t2 = t1 + TT;
  double* const sIn = t2 + TT;
  const int count = a.flags[kLimFlagRcCount];
  const double invDx[3] = {1.0 / a.dx[0], 1.0 / a.dx[1], 1.0 / a.dx[2]};
  const bool active = lane < B3;
  const int node = active ? lane : 0;
  for (int it = blockIdx.x * kRecomputeCells + wid; it < count; it += gridDim.x * kRecomputeCells) {
    const int c = a.rc_list[it];
    const int s = a.st[c];
    const bool helper = s > 0 && s < a.A;
    const bool from_dg = helper && a.pst[c] < a.A;  // rolled back to its previous DG solution
    double* const luh = a.luh + (size_t)c * B3 * V;
    if (from_dg) {
      double* const prev = a.luh_prev + (size_t)c * B3 * V;
      for (int i = lane; i < B3 * V; i += kWave) {
        const double t = luh[i], p = prev[i];
        luh[i] = p;
        prev[i] = t;
        sQ[(i % V) * B3 + i / V] = p;
      }
    } else if (helper) {  // the FV patch of the other buffer projected to DG
      const int buf = a.cur[c] ^ 1;
      const double* const patch = patch_ptr(a, c, buf, PATCH);
#pragma unroll 1
      for (int v = 0; v < V; ++v) {
        for (int i = lane; i < L3; i += kWave) {
          const int zz = i / (L * L), yy = (i / L) % L, xx = i % L;
          sIn[i] = patch[(((zz + G) * S + (yy + G)) * S + (xx + G)) * V + v];
        }
        wave_lds_sync();
        tensor3_rows<N, kTabLim2Uh, L, B>(sIn, t1, t2, lane, [&](int Z, int YX, double u) {
          luh[(Z * B2 + YX) * V + v] = u;
          sQ[v * B3 + Z * B2 + YX] = u;
        });
      }
      if (lane == 0) a.cur[c] = (uint8_t)buf;
    } else {
      for (int i = lane; i < B3 * V; i += kWave) sQ[(i % V) * B3 + i / V] = luh[i];
    }
    wave_lds_sync();
    {  // startNewTimeStep
      double q[V];
#pragma unroll
      for (int m = 0; m < V; ++m) q[m] = sQ[m * B3 + node];
      double dtn = active ? node_stable_dt<Euler<0>>(q, invDx, a.cfl_pnpm) : DBL_MAX;
      dtn = dtn < DBL_MAX ? dtn : DBL_MAX;
      dtn = wave_min(dtn);
      if (lane == 0) atomic_min_positive(&a.st_state->dt_slots[c % kDtSlots], dtn);
    }
    // determineMinAndMax over Gauss-Legendre nodes, Lobatto nodes and subcell
    // means; a helper rolled back to DG also gets its subcell values as its
    // patch interior
    double* const dst = from_dg ? patch_ptr(a, c, a.cur[c], PATCH) : nullptr;
    const int yy = lane / L, xx = lane % L;  // the lane's subcell column in the last phase
    double mn[V], mx[V];
#pragma unroll 1
    for (int v = 0; v < V; ++v) {
      double lo = DBL_MAX, hi = -DBL_MAX;
      if (active) {
        lo = dmin(lo, sQ[v * B3 + lane]);
        hi = dmax(hi, sQ[v * B3 + lane]);
      }
      tensor3_rows<N, kTabUh2Lob, B, B>(sQ + v * B3, t1, t2, lane, [&](int, int, double u) {
        lo = dmin(lo, u);
        hi = dmax(hi, u);
      });
      tensor3_rows<N, kTabUh2Lim, B, L>(sQ + v * B3, t1, t2, lane, [&](int zz, int, double u) {
        lo = dmin(lo, u);
        hi = dmax(hi, u);
        if (from_dg) dst[(((zz + G) * S + (yy + G)) * S + (xx + G)) * V + v] = u;
      });
      lo = wave_min(lo);
      hi = -wave_min(-hi);
#pragma unroll
      for (int w = 0; w < V; ++w) {  // no register array indexed by the rolled v
        mn[w] = (w == v) ? lo : mn[w];
        mx[w] = (w == v) ? hi : mx[w];
      }
    }
    if (lane == 0) {
      double* bmn = a.mn + (size_t)c * 6 * V;
      double* bmx = a.mx + (size_t)c * 6 * V;
      double fmn[V], fmx[V];  // loaded once, before any store
#pragma unroll
      for (int v = 0; v < V; ++v) {
        fmn[v] = (s < a.A) ? mn[v] : a.limmm[(size_t)c * 2 * V + v];
        fmx[v] = (s < a.A) ? mx[v] : a.limmm[(size_t)c * 2 * V + V + v];
      }
#pragma unroll
      for (int f = 0; f < 6; ++f)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          bmn[f * V + v] = fmn[v];
          bmx[f * V + v] = fmx[v];
        }
    }
    wave_lds_sync();  // sQ is rewritten for the next cell
  }
}

// Cells with status 0 after an irregular change: determineMinAndMax writes
// the local extrema on all faces, startNewTimeStep the dt candidate, both as
// the status kernel computed them for the same solution.  One thread per
// face-block value (cell, face, variable), so the stores are contiguous.
__global__ __launch_bounds__(256) void lim_recompute_light_kernel(LimArgs a) {
  if (!a.flags[kLimFlagIrregular]) return;
  constexpr int V = 5;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)a.ncells * 6 * V) return;
  const int c = (int)(i / (6 * V)), r = (int)(i % (6 * V)), v = r % V;
  if (a.st[c] != 0) return;
  a.mn[i] = a.loc_mm[(size_t)c * 2 * V + v];
  a.mx[i] = a.loc_mm[(size_t)c * 2 * V + V + v];
  if (r == 0) atomic_min_positive(&a.st_state->dt_slots[c % kDtSlots], a.dt_cell[c]);
}

// Initial state (LS:1031-1043 and the limiter patches' own initial condition,
// LS:746-762): min/max of the DG initial condition, inadmissible cells
// troubled.  The status spreading of troubled cells reuses the sweep kernel.
template <int N>
__global__ __launch_bounds__(kLimNT) void lim_init_kernel(LimArgs a) {
  using Sh = LimShape<N>;
  constexpr int B3 = Sh::B3, L3 = Sh::L3, V = Sh::V, L = Sh::L, B = Sh::B;
  __shared__ double sQ[V * B3], sLob[V * B3], sSub[V * L3];
  __shared__ double t1[V * B * B * L], t2[V * B * L * L];
  __shared__ double sUh2lim[B * L], sUh2lob[B * B];
  __shared__ double red[2 * V * (kLimNT / kWave)];
  const int c = blockIdx.x;
  load_tables<N>(sUh2lim, sUh2lob, nullptr);
  const double* luh = a.luh + (size_t)c * B3 * V;
  for (int i = threadIdx.x; i < B3 * V; i += kLimNT) sQ[(i % V) * B3 + i / V] = luh[i];
  __syncthreads();
  double mn[V], mx[V];
  dg_local_minmax<N>(sQ, sLob, sSub, t1, t2, sUh2lim, sUh2lob, red, mn, mx);
  if (threadIdx.x == 0) {
    for (int f = 0; f < 6; ++f)
      for (int v = 0; v < V; ++v) {
        a.mn[((size_t)c * 6 + f) * V + v] = mn[v];
        a.mx[((size_t)c * 6 + f) * V + v] = mx[v];
      }
    const bool ok = !(mn[0] <= 0.0) && !(mn[4] <= 0.0);
    a.st[c] = ok ? 0 : a.T;
    a.itc[c] = ok ? 0 : a.cure + 1;
    a.has_patch[c] = 0;
    a.cur[c] = 0;
    if (!ok) a.flags[kLimFlagIrregular] = 1;
  }
}

// Cells with status > 0 after the initial spreading get their patch: the FV
// initial condition at subcell centres (lim_ic [cells][L^3][V]) or, without
// one, the projection of the DG initial condition.
template <int N>
__global__ __launch_bounds__(kLimNT) void lim_init_patches_kernel(LimArgs a, const double* lim_ic) {
  using Sh = LimShape<N>;
  constexpr int B3 = Sh::B3, L3 = Sh::L3, V = Sh::V, L = Sh::L, B = Sh::B, PATCH = Sh::PATCH;
  __shared__ double sQ[V * B3], sSub[V * L3];
  __shared__ double t1[V * B * B * L], t2[V * B * L * L];
  __shared__ double sUh2lim[B * L];
  const int c = blockIdx.x;
  const int s = a.st[c];
  if (threadIdx.x == 0) {
    a.pst[c] = s;
    if (s > 0) a.has_patch[c] = 1;
  }
  if (s <= 0) return;
  double* patch = patch_ptr(a, c, a.cur[c], PATCH);
  if (lim_ic) {
    for (int i = threadIdx.x; i < L3 * V; i += kLimNT) sSub[(i % V) * L3 + i / V] = lim_ic[(size_t)c * L3 * V + i];
    __syncthreads();
  } else {
    load_tables<N>(sUh2lim, nullptr, nullptr);
    const double* luh = a.luh + (size_t)c * B3 * V;
    for (int i = threadIdx.x; i < B3 * V; i += kLimNT) sQ[(i % V) * B3 + i / V] = luh[i];
    __syncthreads();
    tensor3<B, L, V, kLimNT>(sQ, sUh2lim, t1, t2, sSub);
  }
  store_interior<N>(patch, sSub);
}

__global__ void lim_copy_status_kernel(LimArgs a, const int* src, int only_if_irregular) {
  if (only_if_irregular && !a.flags[kLimFlagIrregular]) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < a.ncells) a.st[c] = src[c];
}

// ---------------------------------------------------------------------------
// Batched single-cell entry points
// ---------------------------------------------------------------------------
// projectOnFVLimiterSpace (LIM:23-66): one workgroup per cell; the patch has
// g ghost layers that are left untouched.
template <int N>
__global__ __launch_bounds__(kLimNT) void project_fv_kernel(const double* luh, double* lim, int g) {
  using Sh = LimShape<N>;
  constexpr int B3 = Sh::B3, L3 = Sh::L3, V = Sh::V, L = Sh::L, B = Sh::B;
  __shared__ double sQ[V * B3], sSub[V * L3];
  __shared__ double t1[V * B * B * L], t2[V * B * L * L];
  __shared__ double sUh2lim[B * L];
  const int c = blockIdx.x;
  const int S = L + 2 * g;
  load_tables<N>(sUh2lim, nullptr, nullptr);
  for (int i = threadIdx.x; i < B3 * V; i += kLimNT) sQ[(i % V) * B3 + i / V] = luh[(size_t)c * B3 * V + i];
  __syncthreads();
  tensor3<B, L, V, kLimNT>(sQ, sUh2lim, t1, t2, sSub);
  double* out = lim + (size_t)c * S * S * S * V;
  for (int i = threadIdx.x; i < L3 * V; i += kLimNT) {
    const int v = i % V, n = i / V;
    const int z = n / (L * L), y = (n / L) % L, x = n % L;
    out[(((z + g) * S + (y + g)) * S + (x + g)) * V + v] = sSub[v * L3 + n];
  }
}

// projectOnDGSpace (LIM:69-112)
template <int N>
__global__ __launch_bounds__(kLimNT) void project_dg_kernel(const double* lim, double* luh, int g) {
  using Sh = LimShape<N>;
  constexpr int B3 = Sh::B3, L3 = Sh::L3, V = Sh::V, L = Sh::L, B = Sh::B;
  __shared__ double sIn[V * L3], sOut[V * B3];
  __shared__ double t1[V * L * L * B], t2[V * L * B * B];
  __shared__ double sLim2uh[B * L];
  const int c = blockIdx.x;
  const int S = L + 2 * g;
  load_tables<N>(nullptr, nullptr, sLim2uh);
  const double* in = lim + (size_t)c * S * S * S * V;
  for (int i = threadIdx.x; i < L3 * V; i += kLimNT) {
    const int v = i % V, n = i / V;
    const int z = n / (L * L), y = (n / L) % L, x = n % L;
    sIn[v * L3 + n] = in[(((z + g) * S + (y + g)) * S + (x + g)) * V + v];
  }
  __syncthreads();
  tensor3<L, B, V, kLimNT>(sIn, sLim2uh, t1, t2, sOut);
  for (int i = threadIdx.x; i < B3 * V; i += kLimNT) luh[(size_t)c * B3 * V + i] = sOut[(i % V) * B3 + i / V];
}

// findCellLocalMinAndMax (+ discreteMaximumPrincipleAndMinAndMaxSearch when
// bmin is given), one workgroup per cell.
template <int N>
__global__ __launch_bounds__(kLimNT) void dmp_kernel(const double* luh, double* omn, double* omx, double* bmin,
                                                     double* bmax, int* ok, double relax, double scaling) {
  using Sh = LimShape<N>;
  constexpr int B3 = Sh::B3, L3 = Sh::L3, V = Sh::V, L = Sh::L, B = Sh::B;
  __shared__ double sQ[V * B3], sLob[V * B3], sSub[V * L3];
  __shared__ double t1[V * B * B * L], t2[V * B * L * L];
  __shared__ double sUh2lim[B * L], sUh2lob[B * B];
  __shared__ double red[2 * V * (kLimNT / kWave)];
  const int c = blockIdx.x;
  load_tables<N>(sUh2lim, sUh2lob, nullptr);
  for (int i = threadIdx.x; i < B3 * V; i += kLimNT) sQ[(i % V) * B3 + i / V] = luh[(size_t)c * B3 * V + i];
  __syncthreads();
  double mn[V], mx[V];
  dg_local_minmax<N>(sQ, sLob, sSub, t1, t2, sUh2lim, sUh2lob, red, mn, mx);
  if (threadIdx.x != 0) return;
  for (int v = 0; v < V; ++v) {
    omn[(size_t)c * V + v] = mn[v];
    omx[(size_t)c * V + v] = mx[v];
  }
  if (!bmin) return;
  double* bmn = bmin + (size_t)c * 6 * V;
  double* bmx = bmax + (size_t)c * 6 * V;
  int good = 1;
  for (int v = 0; v < V; ++v) {
    double lo = bmn[v], hi = bmx[v];
    for (int f = 1; f < 6; ++f) {
      lo = dmin(lo, bmn[f * V + v]);
      hi = dmax(hi, bmx[f * V + v]);
    }
    double sd = (hi - lo) * scaling;
    sd = dmax(sd, relax);
    if ((mn[v] < (lo - sd)) || (mx[v] > (hi + sd))) good = 0;
    bmn[v] = mn[v];
    bmx[v] = mx[v];
    bmn[V + v] = lo;
    bmx[V + v] = hi;
  }
  ok[c] = good;
}

// A list covering every patch (the batched MUSCL-Hancock entry point runs the
// grid's FV kernel over it).
__global__ __launch_bounds__(256) void iota_kernel(int* list, int n, int* flags) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) list[i] = i;
  if (i == 0) flags[kLimFlagFvCount] = n;
}


// ---------------------------------------------------------------------------
// Projection tables on the host (LPM:39-223 restated: Lagrange basis on the
// Gauss-Legendre nodes, subcell means by Gauss-Legendre quadrature, Lobatto
// values, and the constrained least-squares inverse through a Gauss-Jordan
// inverse that keeps the reference's pivot bookkeeping, LPM:181).  Evaluated
// without contraction so the tables match a plain x86-64 build bit for bit.
// ---------------------------------------------------------------------------
namespace host_tables {  // the same reference tables, as host data
#define EXA_TABLE static const
#include "dg_tables.inc"
#undef EXA_TABLE
}  // namespace host_tables

#pragma clang fp contract(off)
static void basis_values(double* phi, double xi, int B) {
  const double* xn = host_tables::exa_gl_nodes[B - 1];
  for (int i = 0; i < B; ++i) phi[i] = 1;
  for (int m = 0; m < B; ++m)
    for (int j = 0; j < B; ++j) {
      if (j == m) continue;
      phi[m] = phi[m] * (xi - xn[j]) / (xn[m] - xn[j]);
    }
}

static bool invert(int n, const double* a, double* ia) {
  double c[11][22];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) c[i][j] = a[j * n + i];
    for (int j = n; j < 2 * n; ++j) c[i][j] = 0.;
    c[i][i + n] = 1.;
  }
  for (int i = 0; i < n; ++i) {
    int ml = i;
    double mlV = fabs(c[i][i]);
    for (int j = i + 1; j < n; ++j)
      if (fabs(c[j][i]) > mlV) {
        ml = j;
        mlV = c[j][i];  // signed, as the reference stores it
      }
    for (int k = 0; k < 2 * n; ++k) {
      const double t = c[ml][k];
      c[ml][k] = c[i][k];
      c[i][k] = t;
    }
    if (c[i][i] == 0) return false;
    const double piv = 1. / c[i][i];
    for (int k = 0; k < 2 * n; ++k) c[i][k] *= piv;
    for (int j = i + 1; j < n; ++j) {
      const double t = c[j][i];
      for (int k = 0; k < 2 * n; ++k) c[j][k] -= t * c[i][k];
    }
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = i - 1; j >= 0; --j) {
      const double t = c[j][i];
      for (int k = 0; k < 2 * n; ++k) c[j][k] -= t * c[i][k];
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) ia[j * n + i] = c[i][j + n];
  return true;
}

bool build_limiter_tables(LimTables* t) {
  memset(t, 0, sizeof(*t));
  for (int o = 0; o < 10; ++o) {
    const int B = o + 1, L = 2 * o + 1;
    double phi[10];
    double* u2l = t->uh2lim[o];
    const double dxi = 1. / L;
    for (int i = 0; i < L; ++i) {
      const double xl = i * dxi;
      for (int j = 0; j < B; ++j) {
        basis_values(phi, xl + dxi * host_tables::exa_gl_nodes[B - 1][j], B);
        for (int k = 0; k < B; ++k) u2l[k * L + i] += host_tables::exa_gl_weights[B - 1][j] * phi[k];
      }
    }
    for (int i = 0; i < B; ++i) {
      basis_values(phi, host_tables::exa_lob_nodes[B - 1][i], B);
      for (int j = 0; j < B; ++j) t->uh2lob[o][j * B + i] = phi[j];
    }
    const int n1 = B + 1;
    double m[11 * 11], im[11 * 11], rhs[19 * 11];
    for (int i = 0; i < B; ++i) {
      for (int j = 0; j < B; ++j) {
        m[i * n1 + j] = 0.;
        for (int k = 0; k < L; ++k) m[i * n1 + j] += 2 * u2l[i * L + k] * u2l[j * L + k];
      }
      m[i * n1 + B] = host_tables::exa_gl_weights[B - 1][i];
    }
    for (int i = 0; i < B; ++i) m[B * n1 + i] = -host_tables::exa_gl_weights[B - 1][i];
    m[B * n1 + B] = 0.;
    if (!invert(n1, m, im)) return false;
    for (int i = 0; i < L; ++i) {
      for (int j = 0; j < B; ++j) rhs[i * n1 + j] = 2 * u2l[j * L + i];
      rhs[i * n1 + B] = dxi;
    }
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < B; ++j) {
        double acc = 0.;
        for (int k = 0; k < B + 1; ++k) acc += im[k * n1 + j] * rhs[i * n1 + k];
        t->lim2uh[o][i * B + j] = acc;
      }
  }
  return true;
}
#pragma clang fp contract(on)

hipError_t ensure_limiter_tables() {
  static std::once_flag once;
  static hipError_t status = hipSuccess;
  std::call_once(once, [] {
    static LimTables host;
    if (!build_limiter_tables(&host)) {
      status = hipErrorInvalidValue;
      return;
    }
    status = hipMemcpyToSymbol(HIP_SYMBOL(c_lim), &host, sizeof(host));
  });
  return status;
}

hipError_t launch_lim_iota(int* list, int n, int* flags, hipStream_t s) {
  hipLaunchKernelGGL(iota_kernel, dim3(blocks256(n > 0 ? n : 1)), dim3(256), 0, s, list, n, flags);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <class F>
static hipError_t dispatch_lim(int order, F&& f) {
  switch (order) {
    case 2: return f.template operator()<2>();
    case 3: return f.template operator()<3>();
    default: return hipErrorNotSupported;
  }
}

bool limiter_supported(const SolverDesc& sd) {
  return sd.pde == kPdeEuler && sd.nvar == 5 && sd.npar == 0 && sd.terms == kFlux && sd.precision == kFP64 &&
         sd.order >= 2 && sd.order <= kLimMaxOrder && !sd.no_time_averaging;
}

hipError_t launch_lim_begin_step(const SolverDesc& sd, const LimArgs& a, hipStream_t s) {
  (void)sd;
  hipLaunchKernelGGL(lim_clear_flags_kernel, dim3(1), dim3(64), 0, s, a.flags);
  hipLaunchKernelGGL(lim_begin_kernel, dim3(blocks256(a.ncells)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_lim_minmax_merge(const SolverDesc& sd, const LimArgs& a, const FaceDesc* faces, long nfaces,
                                   hipStream_t s) {
  (void)sd;
  if (nfaces <= 0) return hipSuccess;
  hipLaunchKernelGGL(lim_minmax_merge_kernel, dim3(blocks256(nfaces * 5)), dim3(256), 0, s, a, faces, nfaces);
  return hipGetLastError();
}

static unsigned persistent_grid(long ncells, long per_cu = 2) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const long g = per_cu * cus;
  return (unsigned)(g < ncells ? g : (ncells > 0 ? ncells : 1));
}

hipError_t launch_lim_ghosts(const SolverDesc& sd, const LimArgs& a, bool only_if_irregular, hipStream_t s) {
  return dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    // a pure gather (no LDS): eight workgroups per CU keep more of it in flight
    hipLaunchKernelGGL(lim_ghost_kernel<N>, dim3(persistent_grid(a.ncells, 8)), dim3(kLimNT), 0, s, a,
                       only_if_irregular ? 1 : 0);
    return hipGetLastError();
  });
}

hipError_t launch_lim_fv_update(const SolverDesc& sd, const LimArgs& a, bool only_if_irregular, hipStream_t s) {
  return dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    if (a.fv_godunov)
      hipLaunchKernelGGL((lim_fv_update_kernel<N, true>), dim3(persistent_grid(a.ncells)), dim3(kMusclNT), 0, s, a,
                         only_if_irregular ? 1 : 0);
    else
      hipLaunchKernelGGL((lim_fv_update_kernel<N, false>), dim3(persistent_grid(a.ncells)), dim3(kMusclNT), 0, s, a,
                         only_if_irregular ? 1 : 0);
    return hipGetLastError();
  });
}

hipError_t launch_lim_update_status(const SolverDesc& sd, const LimArgs& a, hipStream_t s) {
  return dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    hipLaunchKernelGGL(lim_update_status_kernel<N>, dim3((a.ncells + kStatusCells - 1) / kStatusCells),
                       dim3(kStatusCells * kWave), 0, s, a);
    return hipGetLastError();
  });
}

hipError_t launch_lim_spread_rollback(const SolverDesc& sd, const LimArgs& a, hipStream_t s) {
  const unsigned g = blocks256(a.ncells > kDtSlots ? a.ncells : kDtSlots);
  const int* in = a.st;
  int* out = a.st2;
  for (int sweep = 0; sweep < a.T; ++sweep) {
    hipLaunchKernelGGL(lim_spread_kernel, dim3(g), dim3(256), 0, s, a, in, out, sweep);
    int* nin = out;
    out = (in == a.st) ? a.st : a.st2;
    in = nin;
  }
  if (in != a.st) hipLaunchKernelGGL(lim_copy_status_kernel, dim3(blocks256(a.ncells)), dim3(256), 0, s, a, in, 1);
  hipLaunchKernelGGL(lim_reset_fv_count_kernel, dim3(1), dim3(64), 0, s, a.flags, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(lim_rollback_lists_kernel, dim3(blocks256(a.ncells)), dim3(256), 0, s, a);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const long g = 8L * cus;
    hipLaunchKernelGGL(lim_rollback_kernel<N>, dim3((unsigned)(g < a.ncells ? g : a.ncells)), dim3(kLimNT), 0, s, a);
    return hipGetLastError();
  });
}

hipError_t launch_lim_repredict_mask(const LimArgs& a, int* out, hipStream_t s) {
  hipLaunchKernelGGL(lim_repredict_mask_kernel, dim3(blocks256(a.ncells)), dim3(256), 0, s, a, out);
  return hipGetLastError();
}

hipError_t launch_lim_recompute_finish(const SolverDesc& sd, const LimArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(lim_recompute_light_kernel, dim3(blocks256((long)a.ncells * 30)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const long g = 8L * cus;  // workgroups of kRecomputeCells cells each
    hipLaunchKernelGGL(lim_recompute_kernel<N>, dim3((unsigned)(g < a.ncells ? g : a.ncells)),
                       dim3(kRecomputeCells * kWave), 0, s, a);
    return hipGetLastError();
  });
}

hipError_t launch_lim_init(const SolverDesc& sd, const LimArgs& a, const double* lim_ic, hipStream_t s) {
  hipLaunchKernelGGL(lim_clear_flags_kernel, dim3(1), dim3(64), 0, s, a.flags);
  hipError_t e = dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    hipLaunchKernelGGL(lim_init_kernel<N>, dim3(a.ncells), dim3(kLimNT), 0, s, a);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  // initial spreading (only when a cell was inadmissible: the flag gates it)
  const unsigned g = blocks256(a.ncells > kDtSlots ? a.ncells : kDtSlots);
  const int* in = a.st;
  int* out = a.st2;
  for (int sweep = 0; sweep < a.T; ++sweep) {
    hipLaunchKernelGGL(lim_spread_kernel, dim3(g), dim3(256), 0, s, a, in, out, sweep);
    int* nin = out;
    out = (in == a.st) ? a.st : a.st2;
    in = nin;
  }
  if (in != a.st) hipLaunchKernelGGL(lim_copy_status_kernel, dim3(blocks256(a.ncells)), dim3(256), 0, s, a, in, 1);
  return dispatch_lim(sd.order, [&]<int N>() -> hipError_t {
    hipLaunchKernelGGL(lim_init_patches_kernel<N>, dim3(a.ncells), dim3(kLimNT), 0, s, a, lim_ic);
    return hipGetLastError();
  });
}

hipError_t launch_project_on_fv(int order, int V, int g, const double* luh, double* lim, long ncells, hipStream_t s) {
  if (V != 5 || ncells <= 0) return ncells <= 0 ? hipSuccess : hipErrorNotSupported;
  return dispatch_lim(order, [&]<int N>() -> hipError_t {
    hipLaunchKernelGGL(project_fv_kernel<N>, dim3((unsigned)ncells), dim3(kLimNT), 0, s, luh, lim, g);
    return hipGetLastError();
  });
}

hipError_t launch_project_on_dg(int order, int V, int g, const double* lim, double* luh, long ncells, hipStream_t s) {
  if (V != 5 || ncells <= 0) return ncells <= 0 ? hipSuccess : hipErrorNotSupported;
  return dispatch_lim(order, [&]<int N>() -> hipError_t {
    hipLaunchKernelGGL(project_dg_kernel<N>, dim3((unsigned)ncells), dim3(kLimNT), 0, s, lim, luh, g);
    return hipGetLastError();
  });
}

hipError_t launch_dmp(int order, const double* luh, double* mn, double* mx, double* bmin, double* bmax, int* ok,
                      double relax, double scaling, long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  return dispatch_lim(order, [&]<int N>() -> hipError_t {
    hipLaunchKernelGGL(dmp_kernel<N>, dim3((unsigned)ncells), dim3(kLimNT), 0, s, luh, mn, mx, bmin, bmax, ok, relax,
                       scaling);
    return hipGetLastError();
  });
}

}  // namespace exa
