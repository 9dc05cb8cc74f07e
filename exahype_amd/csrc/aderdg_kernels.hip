// aderdg_kernels.hip -- instantiation and dispatch of the ADER-DG time-step
// kernels for gfx950 (CDNA4).  The kernel templates (predictor_kernel,
// riemann_kernel, update_kernel, corrector_kernel) live in
// include/exahype_amd/device/{predictor,corrector}.hpp, templated on the PDE
// functor and on the arithmetic type T of the bulk work (double: the plain
// form; float: the FP32-bulk form); this file instantiates them for the
// library's PDEs and holds the PDE-independent kernels:
//   dt_finalize_kernel global min of the cell time steps (mappings/SolutionUpdate.cpp:269-327)
//   volume_integral_kernel, time_average_lfi_kernel  single-cell API forms
//   euler_ic_kernel, elastic_ic_kernel               the workloads' initial conditions
#include "corrector.hpp"
#include "predictor.hpp"
#include "aderdg.hpp"
#include "exa_device.hpp"

#include <cfloat>
#include <type_traits>

namespace exa {

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

// Fused time stepping (fuse-algorithmic-steps), the end of a sweep: stable =
// min(slots), the admissible dt of the updated solution; the stability check
// of the dt the sweep's prediction used and the next predictor dt
// (TimeStepSizeComputation.cpp:167-204), then startNewTimeStep
// (ADERDGSolver.cpp:560-592): t += dt, dt <- dtp, dtp <- next.  init: the
// start (updateTimeStepSizesFused, ADERDGSolver.cpp:594-616), dt = dtp =
// stable.  Contraction off: the oracle's arithmetic, so that the check takes
// the same branch.
__global__ __launch_bounds__(256) void dt_finalize_fused_kernel(StepState* st, double weight, int init) {
#pragma clang fp contract(off)
  __shared__ double sRed[256 / kWave];
  double m = 1.0e308;
  for (int i = threadIdx.x; i < kDtSlots; i += 256) {
    const double v = __longlong_as_double((long long)atomic_take(&st->dt_slots[i], kDtSlotEmpty));
    m = v < m ? v : m;
  }
  m = block_min<256>(m, sRed);
  if (threadIdx.x != 0) return;
  if (init) {
    st->dt = m;
    st->dtp = m;
    st->violated = 0;
  } else {
    const double stable = m;
    double next = stable;
    double used = st->dtp;
    if (used == 0.0) used = stable;
    const int violated = used > stable ? 1 : 0;
    if (violated) {
      next = fmin(next, weight * stable);
      st->dtp = weight * stable;
    } else {
      next = fmin(next, 0.5 * (used + weight * stable));
    }
    st->t += st->dt;
    st->steps += 1;
    st->dt = st->dtp;
    st->dtp = next;
    st->violated = violated;
    st->reruns += violated;
  }
  if (st->dt_hist && st->steps < st->dt_hist_len) st->dt_hist[st->steps] = st->dt;
}

hipError_t launch_dt_finalize_fused(StepState* st, double weight, bool init, hipStream_t s) {
  hipLaunchKernelGGL(dt_finalize_fused_kernel, dim3(1), dim3(256), 0, s, st, weight, init ? 1 : 0);
  return hipGetLastError();
}

// Fused time stepping of the limiting solver, in a one-process run.  Phase 1
// is the SolutionUpdate endIteration (SolutionUpdate.cpp:305-318): without an
// irregular limiter-domain change the check and startNewTimeStep of the plain
// solver; with one, the rollback (Reinitialisation.cpp:105-120) takes the
// time stamps back again, so only the checked predictor dt survives (the
// "used" dt of the repeated check).  Phase 2, after the local recomputation,
// repeats the check with the recomputed cells' candidates
// (LocalRecomputation.cpp:168-222) and then the adapter's
// TimeStepSizeComputation sets corrector and predictor dt to the new stable
// dt (updateTimeStepSizesFused, ADERDGSolver.cpp:594-612).
__global__ __launch_bounds__(256) void dt_finalize_fused_limiting_kernel(StepState* st, double weight,
                                                                          const int* irregular, int phase) {
#pragma clang fp contract(off)
  __shared__ double sRed[256 / kWave];
  const bool irr = *irregular != 0;
  if (phase == 2 && !irr) return;
  double m = 1.0e308;
  for (int i = threadIdx.x; i < kDtSlots; i += 256) {
    const double v = __longlong_as_double((long long)atomic_take(&st->dt_slots[i], kDtSlotEmpty));
    m = v < m ? v : m;
  }
  m = block_min<256>(m, sRed);
  if (threadIdx.x != 0) return;
  const double stable = m;
  if (phase == 1) {
    double next = stable;
    double used = st->dtp;
    if (used == 0.0) used = stable;
    const int violated = used > stable ? 1 : 0;
    double checked = st->dtp;
    if (violated) {
      next = fmin(next, weight * stable);
      checked = weight * stable;
    } else {
      next = fmin(next, 0.5 * (used + weight * stable));
    }
    if (irr) {
      st->dtp_checked = checked;
      return;
    }
    st->t += st->dt;
    st->steps += 1;
    st->dt = checked;
    st->dtp = next;
    st->violated = violated;
    st->reruns += violated;
  } else {
    double used = st->dtp_checked;
    if (used == 0.0) used = stable;
    const int violated = used > stable ? 1 : 0;
    st->t += st->dt;
    st->steps += 1;
    st->dt = stable;
    st->dtp = stable;
    st->violated = violated;
    st->reruns += violated;
  }
  if (st->dt_hist && st->steps < st->dt_hist_len) st->dt_hist[st->steps] = st->dt;
}

hipError_t launch_dt_finalize_fused_limiting(StepState* st, double weight, const int* irregular, int phase,
                                             hipStream_t s) {
  hipLaunchKernelGGL(dt_finalize_fused_limiting_kernel, dim3(1), dim3(256), 0, s, st, weight, irregular, phase);
  return hipGetLastError();
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
