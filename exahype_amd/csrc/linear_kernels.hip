// linear_kernels.hip -- ADER-DG kernels for linear PDEs on gfx950, FP64.
//
// Reference semantics followed (K = ExaHyPE/kernels/aderdg/generic/c/3d):
//   linear_predictor_kernel  K/spaceTimePredictorLinear.cpph:34-443 (Cauchy-Kowalevski
//                            procedure, time averages, extrapolation)
//                            + K/volumeIntegralLinear.cpph:21-53 (fused epilogue)
//   linear_riemann_kernel    K/riemannSolverLinear.cpph:23-132
//   linear_volume_kernel     K/volumeIntegralLinear.cpph:21-53
//   linear_surface_kernel    K/surfaceIntegralLinear.cpph:21-90
//
// The predictor runs one workgroup per cell and three threads per node
// (z,y,x), one for each direction d.  The time derivatives q_0 .. q_B of the
// Taylor expansion are formed one after the other: q_i sits in LDS, thread
// (d, node) takes the node's gradient along d and the PDE's directional
// non-conservative product B_d grad_d q_i, the three products meet in LDS and
// the direction-0 thread forms q_{i+1} = -sum_d B_d grad_d q_i.  Thread (d,
// node) keeps the running time average of its own product in registers (the
// direction-0 thread also that of q); nothing of the space-time expansion is
// written to memory unless the caller asks for the reference's temporaries.
// The face extrapolation reads the time averages straight from LDS.
#include "aderdg.hpp"
#include "exa_device.hpp"

#include <type_traits>

namespace exa {

namespace {

// Node index of line position l through face node (a, b) of a face normal
// to direction dir: x faces (a, b) = (z, y), y faces (z, x), z faces (y, x).
__device__ inline int face_line_node(int dir, int a, int b, int l, int B) {
  return dir == 0 ? (a * B + b) * B + l : (dir == 1 ? (a * B + l) * B + b : (l * B + a) * B + b);
}

}  // namespace

// Direction groups of the predictor: three groups of node threads (one per
// direction) up to B = 5; at B = 6 the 768 threads of three groups leave too
// few registers per thread, and one group handles all three directions.
__host__ __device__ constexpr int linear_predictor_groups(int B) { return B <= 5 ? 3 : 1; }
__host__ __device__ constexpr int linear_predictor_threads(int B) {
  return linear_predictor_groups(B) * node_threads(B);
}

template <int B, class PDE, bool EXPORT, bool PS>
__global__ __launch_bounds__(linear_predictor_threads(B)) void linear_predictor_kernel(LinearPredictorArgs a) {
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B, B3 = B2 * B, B4 = B3 * B;
  constexpr int NT = node_threads(B);
  constexpr int G = linear_predictor_groups(B);
  constexpr int NT3 = G * NT;  // threads of the workgroup
  constexpr int ND = G == 3 ? 1 : 3;  // directions per thread
  __shared__ double sQ[D * B3];      // q_i [var][node] (rows V.. the parameters); lQhi at the end
  __shared__ double sF[3][V * B3];   // B_d grad_d q_i [d][var][node]; lFhi_d at the end
  __shared__ double sDu[B2];         // dudx[l][n]
  const size_t cell = (size_t)blockIdx.x + a.cell0;
  const int grp = threadIdx.x / NT;  // the direction this thread differentiates in (wave-uniform)
  const int tid = threadIdx.x - grp * NT;
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  for (int e = threadIdx.x; e < B2; e += NT3) sDu[e] = exa_dudx[N][e / B][e % B];
  const double* dxp = a.dx + cell * a.dx_stride;
  const double rdx0 = 1.0 / dxp[0], rdx1 = 1.0 / dxp[1], rdx2 = 1.0 / dxp[2];
  const double dt = *a.dt;
  double Q[D];  // the node's current derivative, then the (constant) parameters
  {
    const double* luh = a.luh + cell * (size_t)(B3 * D) + node * D;
#pragma unroll
    for (int m = 0; m < D; ++m) Q[m] = luh[m];
  }
  // sum_n w_n PSi[n+1] of the node's variable m (:79-98, 292-309)
  auto point_source_sum = [&](int m) {
    double s = 0.0;
#pragma unroll
    for (int n = 1; n < B + 1; ++n) s = s + exa_gl_weights[N][n - 1] * a.PSi[((size_t)n * B3 + node) * V + m];
    return s;
  };
  if constexpr (PS) {  // lQi[0] += dt * sum
#pragma unroll
    for (int m = 0; m < V; ++m) Q[m] += dt * point_source_sum(m);
  }
  if (grp == 0 && active) {
#pragma unroll
    for (int m = 0; m < D; ++m) sQ[m * B3 + node] = Q[m];
    if (EXPORT) {
#pragma unroll
      for (int m = 0; m < D; ++m) a.lQi[(size_t)node * D + m] = Q[m];
    }
  }
  // Time averages lQhi = q_0 + sum_{i>0} fac_i q_i (group 0) and lFhi_d alike
  // (group d), formed as 0 + 1 q_0 + ..., exact in the first term.
  double qh[V], fh[ND][V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    qh[m] = 0.0;
#pragma unroll
    for (int k = 0; k < ND; ++k) fh[k][m] = 0.0;
  }
  double fac = 0.5 * dt;  // Taylor factor dt^i / (i+1)! of the time average, :235-250
#pragma unroll 1
  for (int i = 0; i < B; ++i) {
    const double c = i == 0 ? 1.0 : fac;
    __syncthreads();  // q_i complete in sQ
#pragma unroll
    for (int m = 0; m < V; ++m) Q[m] = sQ[m * B3 + node];
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      const int d = G == 3 ? grp : k;
      // gradient of q_i in direction d, :118-186: sum_n (1/dx_d q_i) dudx[l][n],
      // formed as 1/dx_d (sum_n q_i dudx[l][n]) (one rounding apart per term)
      const double rdx = d == 0 ? rdx0 : (d == 1 ? rdx1 : rdx2);
      const int l = d == 0 ? x : (d == 1 ? y : z);
      const int base = d == 0 ? (z * B + y) * B : (d == 1 ? z * B2 + x : y * B + x);
      const int stride = d == 0 ? 1 : (d == 1 ? B : B2);
      double g[V], F[V];
#pragma unroll
      for (int m = 0; m < V; ++m) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n < B; ++n) acc += sQ[m * B3 + base + n * stride] * sDu[l * B + n];
        g[m] = rdx * acc;
      }
      if (d == 0)  // :188-203, B_d grad_d q_i
        PDE::template ncp_dir<0>(Q, g, F);
      else if (d == 1)
        PDE::template ncp_dir<1>(Q, g, F);
      else
        PDE::template ncp_dir<2>(Q, g, F);
      if (active) {
#pragma unroll
        for (int m = 0; m < V; ++m) {
          fh[k][m] += c * F[m];
          sF[d][m * B3 + node] = F[m];
        }
        if (EXPORT) {
          const size_t o = (((size_t)i * B3 + node) * 3 + d) * V;
#pragma unroll
          for (int m = 0; m < V; ++m) {
            a.gradQ[o + m] = g[m];
            a.lFi[o + m] = F[m];
          }
        }
      }
    }
    __syncthreads();  // the three products complete in sF; sQ no longer read
    if (grp == 0 && active) {
      // q_{i+1} = init - F_x - F_y - F_z (:99-105 point source, :205-219)
#pragma unroll
      for (int m = 0; m < V; ++m) {
        double qn = (PS && i == 0) ? a.PSi[(size_t)node * V + m] : 0.0;
        qn -= sF[0][m * B3 + node];
        qn -= sF[1][m * B3 + node];
        qn -= sF[2][m * B3 + node];
        qh[m] += c * Q[m];
        sQ[m * B3 + node] = qn;
        if (EXPORT) a.lQi[((size_t)(i + 1) * B3 + node) * D + m] = qn;
      }
      if (EXPORT) {  // parameters copied (:221-232)
#pragma unroll
        for (int m = V; m < D; ++m) a.lQi[((size_t)(i + 1) * B3 + node) * D + m] = Q[m];
      }
    }
    if (i > 0) fac *= dt / (i + 2);
  }
  if constexpr (PS) {  // :292-309
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const double s3 = point_source_sum(m) / 3;
#pragma unroll
      for (int k = 0; k < ND; ++k) fh[k][m] += s3;
    }
  }
  __syncthreads();  // the last q_B written; sF free
  if (active) {
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      const int d = G == 3 ? grp : k;
#pragma unroll
      for (int m = 0; m < V; ++m) {
        sF[d][m * B3 + node] = fh[k][m];
        if (EXPORT) a.lFhi[((size_t)d * B3 + node) * V + m] = fh[k][m];
      }
    }
    if (grp == 0) {
#pragma unroll
      for (int m = 0; m < V; ++m) sQ[m * B3 + node] = qh[m];  // parameter rows unchanged
    }
    if (EXPORT) {
      if (grp == 0) {
#pragma unroll
        for (int m = 0; m < V; ++m) {
          a.lQhi[(size_t)node * D + m] = qh[m];
          a.lFhi[((size_t)3 * B3 + node) * V + m] = 0.0;
        }
#pragma unroll
        for (int m = V; m < D; ++m) a.lQhi[(size_t)node * D + m] = Q[m];
      }
    }
  }
  if constexpr (EXPORT) {
    for (size_t e = threadIdx.x; e < (size_t)B4 * V; e += NT3) a.lFi[(size_t)B4 * 3 * V + e] = 0.0;
  }
  __syncthreads();  // lQhi in sQ, lFhi_d in sF[d]
  if (grp == 0 && active && a.lduh) {  // volumeIntegralLinear, lduh = -w^3 sum_d lFhi_d
    const double w = exa_gl_weights[N][z] * exa_gl_weights[N][y] * exa_gl_weights[N][x];
    double* du = a.lduh + cell * (size_t)(B3 * V) + node * V;
#pragma unroll
    for (int m = 0; m < V; ++m) {
      double v = 0.0;
      v -= w * sF[0][m * B3 + node];
      v -= w * sF[1][m * B3 + node];
      v -= w * sF[2][m * B3 + node];
      du[m] = v;
    }
  }
  // Extrapolation to the six faces, :312-436: lQhi (D values) onto every face,
  // lFhi_d onto faces 2d and 2d+1; one thread per face node, all variables.
  for (int e = threadIdx.x; e < 6 * B2; e += NT3) {
    const int face = e / B2, ab = e % B2, dir = face >> 1;
    const bool right = face & 1;
    int nd[B];
#pragma unroll
    for (int ll = 0; ll < B; ++ll) nd[ll] = face_line_node(dir, ab / B, ab % B, ll, B);
    double cf[B];
#pragma unroll
    for (int ll = 0; ll < B; ++ll) cf[ll] = right ? exa_fr_coeff[N][ll] : exa_fl_coeff[N][ll];
    double* const qb = a.lQhbnd + cell * (size_t)(6 * B2 * D) + (size_t)e * D;
#pragma unroll
    for (int m = 0; m < D; ++m) {
      double acc = 0.0;
#pragma unroll
      for (int ll = 0; ll < B; ++ll) acc += sQ[m * B3 + nd[ll]] * cf[ll];
      qb[m] = acc;
    }
    const double* const sf = sF[dir];
    double* const fb = a.lFhbnd + cell * (size_t)(6 * B2 * V) + (size_t)e * V;
#pragma unroll
    for (int m = 0; m < V; ++m) {
      double acc = 0.0;
#pragma unroll
      for (int ll = 0; ll < B; ++ll) acc += sf[m * B3 + nd[ll]] * cf[ll];
      fb[m] = acc;
    }
  }
}

// One face per wavefront, one face node per lane.  The face averages of both
// states give the dissipation speed smax and the averaged state Qav; the
// fluctuations are FL = 0.5 (Bn - smax I)^T (QR - QL) into the left cell's
// face 2d+1 and FR = 0.5 (Bn + smax I)^T (QR - QL) into the right cell's face
// 2d, Bn[l][k] = B_d(k, l) the coefficient matrix at Qav (:80-128).  Bn^T q is
// B_d q, which is what the PDE's directional non-conservative product
// computes, so each lane forms ncp_dir<d>(Qav, dq) -+ smax dq for its node
// with dq = 0.5 (QR - QL); the zero entries of Bn the reference multiplies
// through add nothing.
template <int B, class PDE>
__global__ __launch_bounds__(256) void linear_riemann_kernel(RiemannArgs a) {
  constexpr int V = PDE::V, P = PDE::P, D = V + P;
  constexpr int N = B - 1;
  constexpr int B2 = B * B;
  constexpr int WPB = 256 / kWave;
  static_assert(B2 <= kWave, "one face node per lane");
  __shared__ double sStage[WPB][2 * B2 * D];  // QL | QR, then FL | FR
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const long f = (long)blockIdx.x * WPB + w;
  if (f >= a.nfaces) return;  // whole wavefront exits together
  const FaceDesc fd = a.faces[f];
  const int d = fd.dir;
  double* FL = static_cast<double*>(a.lFhbnd) + ((size_t)fd.left * 6 + 2 * d + 1) * (B2 * V);
  double* FR = static_cast<double*>(a.lFhbnd) + ((size_t)fd.right * 6 + 2 * d) * (B2 * V);
  if (fd.pad != 0) {
    // A transmissive boundary face (pad 1: face 2d of cell `right`, pad 2: face
    // 2d+1 of cell `left`): boundaryValues returns the inside state as the
    // outside one, so QR - QL = 0 and both fluctuations vanish.
    double* F = fd.pad == 1 ? FR : FL;
    for (int e = lane; e < B2 * V; e += kWave) F[e] = 0.0;
    return;
  }
  const double* QL = static_cast<const double*>(a.lQhbnd) + ((size_t)fd.left * 6 + 2 * d + 1) * (B2 * D);
  const double* QR = static_cast<const double*>(a.lQhbnd) + ((size_t)fd.right * 6 + 2 * d) * (B2 * D);
  double* const sQL = sStage[w];
  double* const sQR = sQL + B2 * D;
  for (int e = lane; e < B2 * D; e += kWave) {  // contiguous, coalesced loads
    sQL[e] = QL[e];
    sQR[e] = QR[e];
  }
  wave_lds_sync();
  const bool active = lane < B2;
  const int node = active ? lane : 0;
  const double weight = active ? exa_gl_weights[N][node / B] * exa_gl_weights[N][node % B] : 0.0;
  double qL[D], qR[D], QavL[D], QavR[D], Qav[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {  // :40-66
    qL[k] = sQL[node * D + k];
    qR[k] = sQR[node * D + k];
    QavL[k] = wave_sum(weight * qL[k]);
    QavR[k] = wave_sum(weight * qR[k]);
    Qav[k] = 0.5 * (QavL[k] + QavR[k]);
  }
  const double sL = pde_max_abs_eigenvalue<PDE>(QavL, d);  // :68-78
  const double sR = pde_max_abs_eigenvalue<PDE>(QavR, d);
  const double smax = (sL < sR) ? sR : sL;
  double dq[V], F[V];
#pragma unroll
  for (int k = 0; k < V; ++k) dq[k] = 0.5 * (qR[k] - qL[k]);
  if (d == 0)
    PDE::template ncp_dir<0>(Qav, dq, F);
  else if (d == 1)
    PDE::template ncp_dir<1>(Qav, dq, F);
  else
    PDE::template ncp_dir<2>(Qav, dq, F);
  wave_lds_sync();  // every lane has read its states: the stage takes the fluctuations
  double* const sFL = sStage[w];
  double* const sFR = sFL + B2 * V;
  if (active) {
#pragma unroll
    for (int k = 0; k < V; ++k) {
      sFL[node * V + k] = F[k] - smax * dq[k];
      sFR[node * V + k] = F[k] + smax * dq[k];
    }
  }
  wave_lds_sync();
  for (int e = lane; e < B2 * V; e += kWave) {
    FL[e] = sFL[e];
    FR[e] = sFR[e];
  }
}

// The projection step of pointSource (K/pointSource.cpph:40-126): space-time
// node (n, z, y, x) per thread; its force values become phi_z(xRef_z)
// phi_y(xRef_y) phi_x(xRef_x) F / (w_z dz w_y dy w_x dx) for the source
// position x0 the user's pointSource returned at that node, xRef = (x0 -
// (center - dx/2)) / dx, or zero when x0 lies outside the cell.  phi: the
// Lagrange basis on the Gauss-Legendre nodes in product form.
template <int B>
__global__ __launch_bounds__(256) void point_source_kernel(double* PSi, const double* x0, const double* geo, int V) {
  constexpr int N = B - 1, B3 = B * B * B;
  const int node = blockIdx.x * 256 + threadIdx.x;
  if (node >= (B + 1) * B3) return;
  const int k = node % B, j = (node / B) % B, i = (node / (B * B)) % B;
  double xr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) xr[d] = (x0[node * 3 + d] - (geo[d] - geo[3 + d] / 2.)) / geo[3 + d];
  const bool inside = xr[0] >= 0. && xr[1] >= 0. && xr[2] >= 0. && xr[0] <= 1.0 && xr[1] <= 1.0 && xr[2] <= 1.0;
  auto phi = [](int n, double x) {
    double v = 1.0;
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q != n) v *= (x - exa_gl_nodes[N][q]) / (exa_gl_nodes[N][n] - exa_gl_nodes[N][q]);
    return v;
  };
  const double b = phi(i, xr[2]) * phi(j, xr[1]) * phi(k, xr[0]);
  const double den = exa_gl_weights[N][i] * geo[5] * exa_gl_weights[N][j] * geo[4] * exa_gl_weights[N][k] * geo[3];
  for (int m = 0; m < V; ++m) {
    double* p = PSi + (size_t)node * V + m;
    *p = inside ? b * (*p) / den : 0.0;
  }
}

hipError_t launch_point_source(const SolverDesc& sd, double* PSi, const double* x0, const double* geo,
                               hipStream_t s) {
  const int B = sd.order + 1;
  const unsigned blocks = (unsigned)(((B + 1) * B * B * B + 255) / 256);
  switch (B) {
#define EXA_PS(BB) \
  case BB: hipLaunchKernelGGL((point_source_kernel<BB>), dim3(blocks), dim3(256), 0, s, PSi, x0, geo, sd.nvar); break;
    EXA_PS(2) EXA_PS(3) EXA_PS(4) EXA_PS(5) EXA_PS(6) EXA_PS(7) EXA_PS(8)
#undef EXA_PS
    default: return hipErrorNotSupported;
  }
  return hipGetLastError();
}

template <int B, int V>
__global__ __launch_bounds__(node_threads(B)) void linear_volume_kernel(double* lduh, const double* lFhi) {
  constexpr int N = B - 1, B2 = B * B, B3 = B2 * B;
  const int node = threadIdx.x;
  if (node >= B3) return;
  const size_t cell = blockIdx.x;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const double w = exa_gl_weights[N][z] * exa_gl_weights[N][y] * exa_gl_weights[N][x];
  const double* F = lFhi + cell * (size_t)(4 * B3 * V);  // three directions + the source block
#pragma unroll
  for (int m = 0; m < V; ++m) {
    double v = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) v -= w * F[((size_t)d * B3 + node) * V + m];
    lduh[cell * (size_t)(B3 * V) + node * V + m] = v;
  }
}

// lduh -= (F_{2d+1} FRCoeff + F_{2d} FLCoeff) w_a w_b / dx_d per direction.
template <int B, int V>
__global__ __launch_bounds__(node_threads(B)) void linear_surface_kernel(double* lduh, const double* lFhbnd,
                                                                        const double* dx) {
  constexpr int N = B - 1, B2 = B * B, B3 = B2 * B;
  const int node = threadIdx.x;
  if (node >= B3) return;
  const size_t cell = blockIdx.x;
  const int z = node / B2, y = (node / B) % B, x = node % B;
  const double wx = exa_gl_weights[N][x], wy = exa_gl_weights[N][y], wz = exa_gl_weights[N][z];
  const double* lF = lFhbnd + cell * (size_t)(6 * B2 * V);
  double* du = lduh + cell * (size_t)(B3 * V) + node * V;
  const int fn[3] = {z * B + y, z * B + x, y * B + x};
  const int ln[3] = {x, y, z};
  const double us[3] = {(wz * wy) / dx[0], (wz * wx) / dx[1], (wy * wx) / dx[2]};
#pragma unroll
  for (int m = 0; m < V; ++m) {
    double v = du[m];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      v -= (lF[((2 * d + 1) * B2 + fn[d]) * V + m] * exa_fr_coeff[N][ln[d]] +
            lF[((2 * d) * B2 + fn[d]) * V + m] * exa_fl_coeff[N][ln[d]]) *
           us[d];
    du[m] = v;
  }
}

// ---------------------------------------------------------------------------
// Host-side dispatch.  Linear solvers: elasticity at orders 2..5, and the
// five-variable Euler test solver with one parameter at order 3 (the sizes of
// the reference's linear-kernel unit tests); the integrals at those sizes.
// ---------------------------------------------------------------------------
template <class F>
static hipError_t dispatch_linear(const SolverDesc& sd, F&& f) {
  if (sd.pde == kPdeEuler && sd.nvar == 5 && sd.npar == 1 && sd.order == 3) return f.template operator()<4, Euler<1>>();
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

// The integrals depend on (basis size, variables) only.
template <class F>
static hipError_t dispatch_linear_sizes(const SolverDesc& sd, F&& f) {
  if (sd.precision != kFP64) return hipErrorNotSupported;
  if (sd.nvar == 5 && sd.order == 3) return f.template operator()<4, 5>();
  if (sd.nvar == Elastic::V) {
    switch (sd.order) {
      case 2: return f.template operator()<3, 9>();
      case 3: return f.template operator()<4, 9>();
      case 4: return f.template operator()<5, 9>();
      case 5: return f.template operator()<6, 9>();
      default: break;
    }
  }
  return hipErrorNotSupported;
}

bool linear_sizes_supported(const SolverDesc& sd) {
  return dispatch_linear_sizes(sd, []<int B, int V>() -> hipError_t { return hipSuccess; }) == hipSuccess;
}

bool linear_supported(const SolverDesc& sd) {
  if (!sd.linear || sd.precision != kFP64) return false;
  return dispatch_linear(sd, []<int B, class PDE>() -> hipError_t { return hipSuccess; }) == hipSuccess;
}

hipError_t launch_linear_predictor(const SolverDesc& sd, bool export_temps, const LinearPredictorArgs& a,
                                   long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (!linear_supported(sd)) return hipErrorNotSupported;
  if ((export_temps || a.PSi) && ncells != 1) return hipErrorInvalidValue;  // one cell's temporaries / sources
  if (export_temps && (!a.lQi || !a.lFi || !a.gradQ || !a.lQhi || !a.lFhi)) return hipErrorInvalidValue;
  return dispatch_linear(sd, [&]<int B, class PDE>() -> hipError_t {
    const dim3 grid((unsigned)ncells), block(linear_predictor_threads(B));
    if (export_temps) {
      if (a.PSi)
        hipLaunchKernelGGL((linear_predictor_kernel<B, PDE, true, true>), grid, block, 0, s, a);
      else
        hipLaunchKernelGGL((linear_predictor_kernel<B, PDE, true, false>), grid, block, 0, s, a);
    } else if (a.PSi) {
      hipLaunchKernelGGL((linear_predictor_kernel<B, PDE, false, true>), grid, block, 0, s, a);
    } else {
      hipLaunchKernelGGL((linear_predictor_kernel<B, PDE, false, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
  });
}

hipError_t launch_linear_riemann(const SolverDesc& sd, const RiemannArgs& a, hipStream_t s) {
  if (a.nfaces <= 0) return hipSuccess;
  if (!linear_supported(sd) || a.status) return hipErrorNotSupported;
  return dispatch_linear(sd, [&]<int B, class PDE>() -> hipError_t {
    const dim3 grid((unsigned)((a.nfaces + 3) / 4)), block(256);
    hipLaunchKernelGGL((linear_riemann_kernel<B, PDE>), grid, block, 0, s, a);
    return hipGetLastError();
  });
}

hipError_t launch_linear_volume(const SolverDesc& sd, double* lduh, const double* lFhi, long ncells,
                                hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  return dispatch_linear_sizes(sd, [&]<int B, int V>() -> hipError_t {
    hipLaunchKernelGGL((linear_volume_kernel<B, V>), dim3((unsigned)ncells), dim3(node_threads(B)), 0, s, lduh, lFhi);
    return hipGetLastError();
  });
}

hipError_t launch_linear_surface(const SolverDesc& sd, double* lduh, const double* lFhbnd, const double* dx,
                                 long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  return dispatch_linear_sizes(sd, [&]<int B, int V>() -> hipError_t {
    hipLaunchKernelGGL((linear_surface_kernel<B, V>), dim3((unsigned)ncells), dim3(node_threads(B)), 0, s, lduh,
                       lFhbnd, dx);
    return hipGetLastError();
  });
}

}  // namespace exa
