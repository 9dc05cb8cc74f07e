// fv_kernels.hip -- finite-volume kernels on the device (3-D, Euler, FP64):
// the first-order Godunov update of one patch in the reference layout, and a
// device-resident finite-volumes grid stored as dense subcell planes.
//
// Reference functions restated (paths relative to the reference checkout):
//   GOD = ExaHyPE/kernels/finitevolumes/godunov/c/3d/godunov.cpph
//   RUS = ExaHyPE/kernels/finitevolumes/riemannsolvers/c/rusanov.cpph
//   FVC = ExaHyPE/kernels/finitevolumes/commons/c/commons.cpph
//   FVS = ExaHyPE/exahype/solvers/FiniteVolumesSolver.cpp
//
// Every subcell's new value is formed in the reference's accumulation order:
// start from the old value, then + dt/h F(x-1/2), - dt/h F(x+1/2), the same
// for y, then for z (GOD:61-123 visits all x faces, then all y faces, then
// all z faces, and each face adds to its right cell before the next face
// subtracts from it).  A face flux is a function of its two cell states
// only, so computing it once per face, or once per patch that shares it,
// gives the same value.
#include <cfloat>
#include <cmath>

#include "exa_device.hpp"
#include "fv.hpp"
#include "fv_device.hpp"

namespace exa {

namespace {

constexpr double kGamma = 1.4;

// Pressure and sound speed of one state, shared by its directional fluxes
// and eigenvalues (the same expression trees as euler_flux_dir and
// Euler::max_abs_eigenvalue, which compute them again per call).
struct CellBase {
  double irho, p, cs;
};

__device__ inline CellBase cell_base(const double* q) {
  CellBase b;
  b.irho = 1.0 / q[0];
  b.p = (kGamma - 1.0) * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * b.irho);
  b.cs = sqrt(kGamma * b.p * b.irho);
  return b;
}

// Largest |eigenvalue| in direction d: max_element of |(u-c, u, u, u, u+c)|.
__device__ inline double cell_lambda(const double* q, const CellBase& b, int d) {
  const double u_n = q[d + 1] * b.irho;
  const double l0 = fabs(u_n - b.cs), l1 = fabs(u_n), l4 = fabs(u_n + b.cs);
  double m = l0;
  m = (m < l1) ? l1 : m;
  m = (m < l4) ? l4 : m;
  return m;
}

__device__ inline void cell_flux(const double* q, const CellBase& b, int d, double* F) {
  const double qn = q[d + 1];
  F[0] = qn;
#pragma unroll
  for (int m = 1; m < 4; ++m) F[m] = (m == d + 1) ? b.irho * qn * q[m] + b.p : b.irho * qn * q[m];
  F[4] = b.irho * qn * (q[4] + b.p);
}

// Rusanov flux across a face between states qa (left) and qb (right) in
// direction d; returns s_max.
__device__ inline double face_flux(const double* qa, const double* qb, int d, double* fn) {
  const CellBase ba = cell_base(qa), bb = cell_base(qb);
  double Fa[5], Fb[5];
  cell_flux(qa, ba, d, Fa);
  cell_flux(qb, bb, d, Fb);
  const double la = cell_lambda(qa, ba, d), lb = cell_lambda(qb, bb, d);
  rusanov_from_parts(qa, Fa, la, qb, Fb, lb, fn);
  return dmax(lb, la);
}

// ---------------------------------------------------------------------------
// Single patch (GOD:35-143).  One thread per subcell of the (S,S,S) patch:
// interior subcells take their six faces in the reference's order; a
// face-ghost subcell of the first layer takes the one face it shares with
// the interior (GOD's loops start one layer early and end one layer late);
// every other ghost is left as it is.  Every face's dt candidate is formed
// once: by the subcell on its right, or for the last face of a line by the
// subcell on its left.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void godunov_patch_kernel(GodunovPatchArgs a) {
  const int L = a.L, g = a.g, S = L + 2 * g;
  const int pb = g, pe = g + L;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  double dtc = DBL_MAX;
  if (id < S * S * S) {
    const int z = id / (S * S), y = (id / S) % S, x = id % S;
    const int co[3] = {x, y, z};
    const int stride[3] = {5, 5 * S, 5 * S * S};
    const bool in[3] = {x >= pb && x < pe, y >= pb && y < pe, z >= pb && z < pe};
    const double cdt[3] = {a.dt / a.h[0], a.dt / a.h[1], a.dt / a.h[2]};
    const double ch[3] = {a.cfl / 3 * a.h[0], a.cfl / 3 * a.h[1], a.cfl / 3 * a.h[2]};
    const double* q = a.luh + (size_t)id * 5;
    double* out = a.luh_new + (size_t)id * 5;
    if (in[0] && in[1] && in[2]) {
      double un[5], fl[5], fr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) un[v] = out[v];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double smax_l = face_flux(q - stride[d], q, d, fl);
        const double smax_r = face_flux(q, q + stride[d], d, fr);
#pragma unroll
        for (int v = 0; v < 5; ++v) un[v] += cdt[d] * fl[v];
#pragma unroll
        for (int v = 0; v < 5; ++v) un[v] -= cdt[d] * fr[v];
        dtc = dmin(dtc, ch[d] / smax_l);
        if (co[d] == pe - 1) dtc = dmin(dtc, ch[d] / smax_r);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] = un[v];
    } else {
      // a face ghost: exactly one coordinate one layer outside, the others inside
      int dir = -1, side = 0, nout = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (in[d]) continue;
        ++nout;
        if (co[d] == pb - 1) {
          dir = d;
          side = -1;
        } else if (co[d] == pe) {
          dir = d;
          side = 1;
        } else {
          nout = 9;
        }
      }
      if (nout == 1) {
        double f[5];
        if (side < 0) {  // left ghost: luh_new -= dt/h F(face to its right)
          face_flux(q, q + stride[dir], dir, f);
#pragma unroll
          for (int v = 0; v < 5; ++v) out[v] -= cdt[dir] * f[v];
        } else {  // right ghost: luh_new += dt/h F(face to its left)
          face_flux(q - stride[dir], q, dir, f);
#pragma unroll
          for (int v = 0; v < 5; ++v) out[v] += cdt[dir] * f[v];
        }
      }
    }
  }
  const double m = wave_min(dtc);
  if ((threadIdx.x & (kWave - 1)) == 0 && m < DBL_MAX) atomic_min_positive(a.dt_max, m);
}

// ---------------------------------------------------------------------------
// Grid update.  A workgroup owns a tile of 62 x 6 subcell columns and
// walks a range of z planes.  Wavefront w holds row y0 - 1 + w, lane l
// column x0 - 1 + l, so the outer rows and lanes carry the neighbours of the
// tile.  Per plane and column the subcell's pressure and sound speed are
// formed once, each face flux once:
//   x faces: lane l forms the face to its right from lane l+1's state,
//            directional flux and eigenvalue (DPP lane shifts), and passes
//            it to lane l+1 as that lane's left face;
//   y faces: rows exchange state, y flux and eigenvalue through LDS; row w
//            forms the face above it and leaves it in LDS for row w+1;
//   z faces: carried along the walk (the face above plane z is the face
//            below plane z+1).
// The update is followed by the new subcell's eigenvalue maxima for the next
// stable dt (startNewTimeStep, FVS:553-574), reduced over the wavefront and
// into state->lam_max.  Reads per step: each subcell once from the tile, plus
// the tile borders (1/31 along x, 1/3 along y, read again by the
// neighbouring tile, partly from L2); writes: each subcell once.  Eight
// wavefronts per workgroup and two workgroups per CU (114 VGPRs, 64 KiB of
// LDS each) measured faster than one workgroup of sixteen, whose barriers
// hold all sixteen wavefronts of the CU.
// ---------------------------------------------------------------------------
__device__ inline int wrap_index(int i, int n, int periodic) {
  if (periodic) {
    i %= n;
    return i < 0 ? i + n : i;
  }
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// The value of lane l+1 (wave_shl:1) / lane l-1 (wave_shr:1) of a double;
// the last / first lane receives 0 (not used).
__device__ inline double from_next_lane(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double from_prev_lane(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// The state's maxima are read by the next kernel on the stream.  A returning
// atomic whose result is consumed makes the wavefront wait until the atomic
// has been performed (a one-wavefront kernel launched right behind has been
// seen to read a maximum before the producer's last no-return atomic landed).
__device__ inline void atomic_max_positive(unsigned long long* slot, double v) {
  const unsigned long long old = atomicMax(slot, (unsigned long long)__double_as_longlong(v));
  asm volatile("" : : "v"(old));
}

// Wavefront maximum of one direction's eigenvalue maxima into the state.
__device__ inline void reduce_lambda(FvState* st, int d, double lm) {
  const double m = -wave_min(-lm);
  if ((threadIdx.x & (kWave - 1)) == 0 && m > 0.0) atomic_max_positive(&st->lam_max[d], m);
}

__global__ __launch_bounds__(kWave * kFvRows) void fv_godunov_kernel(FvGridArgs a) {
  constexpr int ROWS = kFvRows, OUT_ROWS = ROWS - 2;
  __shared__ double sQ[ROWS][11][kWave];  // state (5), y flux (5), y eigenvalue
  __shared__ double sFy[ROWS][5][kWave];  // face above the row
  // XCD-aware tile order: workgroup b runs on XCD b % 8; give each XCD a
  // contiguous run of tiles with the x strip fastest, so that the tiles that
  // share border columns (x neighbours, adjacent in the run) and border rows
  // (y neighbours, nsx apart) are resident on that XCD at the same time and
  // walk their planes together: the shared cache lines come from its L2
  const int per = (int)(gridDim.x / 8);
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= a.ntiles) return;  // the whole workgroup
  const int sx = t % a.nsx, r2 = t / a.nsx, ry = r2 % a.nry, cz = r2 / a.nry;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int xp = sx * kFvOutX - 1 + lane, yp = ry * OUT_ROWS - 1 + w;
  const int xs = wrap_index(xp, a.X, a.periodic), ys = wrap_index(yp, a.Y, a.periodic);
  const bool out_row = w >= 1 && w <= OUT_ROWS && yp < a.Y;  // uniform per wavefront
  const bool out = out_row && lane >= 1 && lane <= kFvOutX && xp < a.X;
  const bool face_above = out_row || w == 0;
  const size_t plane = (size_t)a.X * a.Y, vol = plane * (size_t)a.Z;
  const size_t col = (size_t)ys * a.X + xs;
  const int z0 = cz * a.zlen, z1 = min(a.Z, z0 + a.zlen);
  const double dt = a.state->dt;
  const double cdt[3] = {dt / a.h[0], dt / a.h[1], dt / a.h[2]};
  auto load = [&](int z, double* q) __device__ {
    const size_t o = (size_t)wrap_index(z, a.Z, a.periodic) * plane + col;
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = a.u[v * vol + o];
  };

  // every row walks with the next plane's state already loaded and the one
  // after in flight (the rows without output only supply their state)
  double q[5], qn[5], qnn[5], fzlo[5];
  CellBase b{};
  load(z0, q);
  load(z0 + 1, qn);
  if (out_row) {
    double qm[5], Fm[5], Fz[5];
    load(z0 - 1, qm);
    const CellBase bm = cell_base(qm);
    b = cell_base(q);
    cell_flux(qm, bm, 2, Fm);
    cell_flux(q, b, 2, Fz);
    rusanov_from_parts(qm, Fm, cell_lambda(qm, bm, 2), q, Fz, cell_lambda(q, b, 2), fzlo);
  }
  double lmx = 0.0, lmy = 0.0, lmz = 0.0;
  for (int z = z0; z < z1; ++z) {
    // phase A: this row's state, y flux and y eigenvalue into LDS
    load(z + 2, qnn);  // in flight during this plane
    if (!out_row) b = cell_base(q);
    {
      double Fy[5];
      cell_flux(q, b, 1, Fy);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        sQ[w][v][lane] = q[v];
        sQ[w][5 + v][lane] = Fy[v];
      }
      sQ[w][10][lane] = cell_lambda(q, b, 1);
    }
    __syncthreads();
    // phase B: the y face above the row into LDS; the x faces
    if (face_above) {
      double Fy[5], qa[5], Fa[5], fyhi[5];
      cell_flux(q, b, 1, Fy);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        qa[v] = sQ[w + 1][v][lane];
        Fa[v] = sQ[w + 1][5 + v][lane];
      }
      rusanov_from_parts(q, Fy, cell_lambda(q, b, 1), qa, Fa, sQ[w + 1][10][lane], fyhi);
#pragma unroll
      for (int v = 0; v < 5; ++v) sFy[w][v][lane] = fyhi[v];
    }
    double un[5];
    if (out_row) {
      double Fx[5], qr[5], Fr[5], fxhi[5];
      cell_flux(q, b, 0, Fx);
      const double lx = cell_lambda(q, b, 0);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        qr[v] = from_next_lane(q[v]);
        Fr[v] = from_next_lane(Fx[v]);
      }
      rusanov_from_parts(q, Fx, lx, qr, Fr, from_next_lane(lx), fxhi);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double fxlo = from_prev_lane(fxhi[v]);
        un[v] = q[v];
        un[v] += cdt[0] * fxlo;
        un[v] -= cdt[0] * fxhi[v];
      }
    }
    __syncthreads();
    // phase C: y and z terms, store, eigenvalue maxima of the new state
    if (out_row) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        un[v] += cdt[1] * sFy[w - 1][v][lane];
        un[v] -= cdt[1] * sFy[w][v][lane];
      }
      const CellBase bn = cell_base(qn);
      double Fz[5], Fzn[5], fzhi[5];
      cell_flux(q, b, 2, Fz);
      cell_flux(qn, bn, 2, Fzn);
      rusanov_from_parts(q, Fz, cell_lambda(q, b, 2), qn, Fzn, cell_lambda(qn, bn, 2), fzhi);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        un[v] += cdt[2] * fzlo[v];
        un[v] -= cdt[2] * fzhi[v];
      }
      if (out) {
        const size_t o = (size_t)z * plane + col;
#pragma unroll
        for (int v = 0; v < 5; ++v) a.un[v * vol + o] = un[v];
        const CellBase bu = cell_base(un);
        lmx = dmax(lmx, cell_lambda(un, bu, 0));
        lmy = dmax(lmy, cell_lambda(un, bu, 1));
        lmz = dmax(lmz, cell_lambda(un, bu, 2));
      }
      // walk one plane up
#pragma unroll
      for (int v = 0; v < 5; ++v) fzlo[v] = fzhi[v];
      b = bn;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      q[v] = qn[v];
      qn[v] = qnn[v];
    }
  }
  reduce_lambda(a.state, 0, lmx);
  reduce_lambda(a.state, 1, lmy);
  reduce_lambda(a.state, 2, lmz);
}

// dt = min over d of ch[d] / lam_max[d] (the minimum over subcells and
// directions of FVC:409-460's candidates, see FvState); 0 / non-finite -> bad.
__device__ inline double dt_from_lambda(FvState* st) {
  double dt = DBL_MAX;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // read and clear at L2 (not through the scalar cache)
    const double lm = __longlong_as_double((long long)atomic_take(&st->lam_max[d], 0ull));
    dt = dmin(dt, st->ch[d] / lm);
  }
  return dt;
}

__global__ void fv_finalize_kernel(FvState* st) {
  if (threadIdx.x != 0) return;
  const double nd = dt_from_lambda(st);
  st->t += st->dt;
  st->dt = nd;
  st->steps += 1;
  if (!(nd > 0.0) || !(nd < DBL_MAX)) st->bad = 1;
}

__global__ __launch_bounds__(256) void fv_lambda_kernel(FvGridArgs a) {
  const size_t vol = (size_t)a.X * a.Y * a.Z;
  double lmx = 0.0, lmy = 0.0, lmz = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < vol; i += (size_t)gridDim.x * blockDim.x) {
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = a.u[v * vol + i];
    const CellBase b = cell_base(q);
    lmx = dmax(lmx, cell_lambda(q, b, 0));
    lmy = dmax(lmy, cell_lambda(q, b, 1));
    lmz = dmax(lmz, cell_lambda(q, b, 2));
  }
  reduce_lambda(a.state, 0, lmx);
  reduce_lambda(a.state, 1, lmy);
  reduce_lambda(a.state, 2, lmz);
}

__global__ void fv_begin_kernel(FvState* st) {
  if (threadIdx.x != 0) return;
  const double nd = dt_from_lambda(st);
  st->t = 0.0;
  st->dt = nd;
  st->steps = 0;
  st->bad = (!(nd > 0.0) || !(nd < DBL_MAX)) ? 1 : 0;
}

__global__ void fv_reset_lambda_kernel(FvState* st) {
  if (threadIdx.x < 3) st->lam_max[threadIdx.x] = 0ull;
}

// planes[v][Z][Y][X] <-> cells[cell][z][y][x][v]
template <bool PACK>
__global__ __launch_bounds__(256) void fv_pack_kernel(const double* src, double* dst, int nx, int ny, int nz, int L) {
  const int X = nx * L, Y = ny * L;
  const size_t vol = (size_t)X * Y * (size_t)(nz * L);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < vol; i += (size_t)gridDim.x * blockDim.x) {
    const int xg = (int)(i % X), yg = (int)((i / X) % Y), zg = (int)(i / ((size_t)X * Y));
    const size_t cell = ((size_t)(zg / L) * ny + yg / L) * nx + xg / L;
    const size_t h = (((cell * L + zg % L) * L + yg % L) * L + xg % L) * 5;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if constexpr (PACK)
        dst[v * vol + i] = src[h + v];
      else
        dst[h + v] = src[v * vol + i];
    }
  }
}

__global__ __launch_bounds__(256) void fv_euler_ic_kernel(ICArgs a, double* planes, int X, int Y, int Z, double h0,
                                                         double h1, double h2) {
  const size_t vol = (size_t)X * Y * (size_t)Z;
  const double twopi = 6.283185307179586476925286766559;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < vol; i += (size_t)gridDim.x * blockDim.x) {
    const int xg = (int)(i % X), yg = (int)((i / X) % Y), zg = (int)(i / ((size_t)X * Y));
    const double P[3] = {a.lo[0] + (xg + 0.5) * h0, a.lo[1] + (yg + 0.5) * h1, a.lo[2] + (zg + 0.5) * h2};
    double rho = a.rho0 + a.amp * sin(twopi * (a.k[0] * P[0] + a.k[1] * P[1] + a.k[2] * P[2]));
    double u[3] = {a.vel[0], a.vel[1], a.vel[2]};
    for (int mo = 0; mo < a.nmodes; ++mo) {
      const double arg = twopi * (a.mode_k[mo][0] * P[0] + a.mode_k[mo][1] * P[1] + a.mode_k[mo][2] * P[2]);
      rho += a.mode_amp * sin(arg + a.mode_phase[mo][0]);
      u[0] += a.mode_amp * sin(arg + a.mode_phase[mo][1]);
      u[1] += a.mode_amp * sin(arg + a.mode_phase[mo][2]);
      u[2] += a.mode_amp * sin(arg + a.mode_phase[mo][3]);
    }
    planes[i] = rho;
    planes[vol + i] = rho * u[0];
    planes[2 * vol + i] = rho * u[1];
    planes[3 * vol + i] = rho * u[2];
    planes[4 * vol + i] = a.p / (kGamma - 1.0) + 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  }
}

// The MUSCL-Hancock grid keeps one patch per cell with its ghost layers,
// [cells][2 buffers][S^3][5] (S = L + 2G, the current buffer per cell in
// cur): the eigenvalue maxima over the current interiors, for the grid's
// stable dt as above.
__global__ __launch_bounds__(256) void fv_patch_lambda_kernel(const double* lim, const uint8_t* cur, long ncells,
                                                             int L, int G, FvState* st) {
  const int S = L + 2 * G;
  const long L3 = (long)L * L * L;
  const size_t PATCH = (size_t)S * S * S * 5;
  const size_t n = (size_t)ncells * L3;
  double lm[3] = {0.0, 0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t c = i / L3;
    const int j = (int)(i % L3);
    const int z = j / (L * L), y = (j / L) % L, x = j % L;
    const double* p = lim + (c * 2 + cur[c]) * PATCH + ((size_t)((z + G) * S + (y + G)) * S + (x + G)) * 5;
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = p[v];
    const CellBase b = cell_base(q);
#pragma unroll
    for (int d = 0; d < 3; ++d) lm[d] = dmax(lm[d], cell_lambda(q, b, d));
  }
  reduce_lambda(st, 0, lm[0]);
  reduce_lambda(st, 1, lm[1]);
  reduce_lambda(st, 2, lm[2]);
}

// cells [cells][L^3][5] (the reference's patch interiors, cell-major) <->
// the interiors of the current patch buffers
template <bool TO_PATCH>
__global__ __launch_bounds__(256) void fv_patch_io_kernel(double* cells, double* lim, const uint8_t* cur, long ncells,
                                                         int L, int G) {
  const int S = L + 2 * G;
  const long L3 = (long)L * L * L;
  const size_t PATCH = (size_t)S * S * S * 5;
  const size_t n = (size_t)ncells * L3;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t c = i / L3;
    const int j = (int)(i % L3);
    const int z = j / (L * L), y = (j / L) % L, x = j % L;
    double* p = lim + (c * 2 + cur[c]) * PATCH + ((size_t)((z + G) * S + (y + G)) * S + (x + G)) * 5;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if constexpr (TO_PATCH)
        p[v] = cells[i * 5 + v];
      else
        cells[i * 5 + v] = p[v];
    }
  }
}

// the dt of the next update into the step state the patch kernels read
__global__ void fv_dt_to_step_kernel(const FvState* fs, StepState* ss) {
  if (threadIdx.x == 0) ss->dt = fs->dt;
}

unsigned stride_grid(size_t n) {
  const size_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

}  // namespace

hipError_t launch_godunov_patch(const GodunovPatchArgs& a, hipStream_t s) {
  const int S = a.L + 2 * a.g;
  const int n = S * S * S;
  hipLaunchKernelGGL(godunov_patch_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fv_godunov_step(const FvGridArgs& a, hipStream_t s) {
  const int grid = (a.ntiles + 7) / 8 * 8;
  hipLaunchKernelGGL(fv_godunov_kernel, dim3(grid), dim3(kWave * kFvRows), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fv_finalize(FvState* st, hipStream_t s) {
  hipLaunchKernelGGL(fv_finalize_kernel, dim3(1), dim3(64), 0, s, st);
  return hipGetLastError();
}

hipError_t launch_fv_initial_dt(const FvGridArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(fv_reset_lambda_kernel, dim3(1), dim3(64), 0, s, a.state);
  hipLaunchKernelGGL(fv_lambda_kernel, dim3(stride_grid((size_t)a.X * a.Y * a.Z)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(fv_begin_kernel, dim3(1), dim3(64), 0, s, a.state);
  return hipGetLastError();
}

hipError_t launch_fv_patch_dt(const double* lim, const uint8_t* cur, long ncells, int L, int G, FvState* st,
                              bool begin, hipStream_t s) {
  hipLaunchKernelGGL(fv_patch_lambda_kernel, dim3(stride_grid((size_t)ncells * L * L * L)), dim3(256), 0, s, lim, cur,
                     ncells, L, G, st);
  if (begin)
    hipLaunchKernelGGL(fv_begin_kernel, dim3(1), dim3(64), 0, s, st);
  else
    hipLaunchKernelGGL(fv_finalize_kernel, dim3(1), dim3(64), 0, s, st);
  return hipGetLastError();
}

hipError_t launch_fv_patch_io(double* cells, double* lim, const uint8_t* cur, long ncells, int L, int G, bool to_patch,
                              hipStream_t s) {
  const size_t n = (size_t)ncells * L * L * L;
  if (to_patch)
    hipLaunchKernelGGL(fv_patch_io_kernel<true>, dim3(stride_grid(n)), dim3(256), 0, s, cells, lim, cur, ncells, L, G);
  else
    hipLaunchKernelGGL(fv_patch_io_kernel<false>, dim3(stride_grid(n)), dim3(256), 0, s, cells, lim, cur, ncells, L, G);
  return hipGetLastError();
}

hipError_t launch_fv_dt_to_step(const FvState* fs, StepState* ss, hipStream_t s) {
  hipLaunchKernelGGL(fv_dt_to_step_kernel, dim3(1), dim3(64), 0, s, fs, ss);
  return hipGetLastError();
}

hipError_t launch_fv_pack(const double* cells, double* planes, int nx, int ny, int nz, int L, hipStream_t s) {
  const size_t vol = (size_t)nx * ny * nz * L * L * L;
  hipLaunchKernelGGL(fv_pack_kernel<true>, dim3(stride_grid(vol)), dim3(256), 0, s, cells, planes, nx, ny, nz, L);
  return hipGetLastError();
}

hipError_t launch_fv_unpack(const double* planes, double* cells, int nx, int ny, int nz, int L, hipStream_t s) {
  const size_t vol = (size_t)nx * ny * nz * L * L * L;
  hipLaunchKernelGGL(fv_pack_kernel<false>, dim3(stride_grid(vol)), dim3(256), 0, s, planes, cells, nx, ny, nz, L);
  return hipGetLastError();
}

hipError_t launch_fv_euler_ic(const ICArgs& ic, double* planes, int X, int Y, int Z, const double* h,
                              hipStream_t s) {
  hipLaunchKernelGGL(fv_euler_ic_kernel, dim3(stride_grid((size_t)X * Y * Z)), dim3(256), 0, s, ic, planes, X, Y, Z,
                     h[0], h[1], h[2]);
  return hipGetLastError();
}

}  // namespace exa
