// capi.hip -- the C-ABI (include/exahype_amd.h): single-cell and batched
// drop-ins for kernels::aderdg::generic::c, and the device-resident grid
// driver that replaces the reference's per-cell solver loop.
#include "../../include/exahype_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "aderdg.hpp"
#include "amr.hpp"

using exa::FaceDesc;
using exa::SolverDesc;

#include "capi_util.hpp"

// ---------------------------------------------------------------------------
extern "C" const char* exa_status_string(int status) {
  switch (status) {
    case EXA_OK: return "ok";
    case EXA_E_HIP: return "HIP runtime error";
    case EXA_E_UNSUPPORTED: return "no compiled instantiation for this solver";
    case EXA_E_INVALID: return "invalid argument";
    case EXA_E_NOMEM: return "device out of memory";
    case EXA_E_NONFINITE: return "non-finite time step size";
  }
  return "unknown status";
}

extern "C" int exa_init(int device) {
  EXA_TRY(hipSetDevice(device));
  return exa::noalloc_scratch_setup();
}

extern "C" int exa_solver_supported(const exa_solver* s) {
  SolverDesc sd;
  if (!to_desc(s, &sd)) return 0;
  return (sd.linear ? exa::linear_supported(sd) : exa::supported(sd)) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Single-cell and batched entry points
// ---------------------------------------------------------------------------
static int predictor_impl(const exa_solver* s, long ncells, double* lQhbnd, double* lFhbnd,
                          double* lduh, double* lQi, double* lFi, double* lQhi, double* lFhi,
                          const double* luh, const double* dx, double dt, int* iters) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || ncells < 1 || !lQhbnd || !lFhbnd || !luh || !dx) return EXA_E_INVALID;
  // a linear solver's glue calls the linear kernels (spaceTimePredictorLinear)
  if (sd.linear || !exa::supported(sd) || sd.precision != exa::kFP64) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  const bool exp = lQi || lFi || lQhi || lFhi;
  DevBuf d_luh, d_qb, d_fb, d_du, d_dx, d_dt, d_it, d_qi, d_fi, d_qhi, d_fhi;
  EXA_TRY(d_luh.alloc(sizeof(double) * ncells * z.B3 * z.D));
  EXA_TRY(d_qb.alloc(sizeof(double) * ncells * 6 * z.B2 * z.D));
  EXA_TRY(d_fb.alloc(sizeof(double) * ncells * 6 * z.B2 * z.V));
  EXA_TRY(d_du.alloc(sizeof(double) * ncells * z.B3 * z.V));
  EXA_TRY(d_dx.alloc(sizeof(double) * ncells * 3));
  EXA_TRY(d_dt.alloc(sizeof(double)));
  EXA_TRY(d_it.alloc(sizeof(int) * ncells));
  if (exp) {
    EXA_TRY(d_qi.alloc(sizeof(double) * ncells * z.B4 * z.D));
    EXA_TRY(d_fi.alloc(sizeof(double) * ncells * 4 * z.B4 * z.V));
    EXA_TRY(d_qhi.alloc(sizeof(double) * ncells * z.B3 * z.D));
    EXA_TRY(d_fhi.alloc(sizeof(double) * ncells * 4 * z.B3 * z.V));
    EXA_TRY(hipMemset(d_fi.p, 0, sizeof(double) * ncells * 4 * z.B4 * z.V));
    EXA_TRY(hipMemset(d_fhi.p, 0, sizeof(double) * ncells * 4 * z.B3 * z.V));
  }
  EXA_TRY(hipMemcpy(d_luh.p, luh, sizeof(double) * ncells * z.B3 * z.D, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dx.p, dx, sizeof(double) * ncells * 3, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dt.p, &dt, sizeof(double), hipMemcpyHostToDevice));
  exa::PredictorArgs a{};
  a.luh = d_luh.d();
  a.lduh = lduh ? d_du.d() : nullptr;
  a.lQhbnd = d_qb.d();
  a.lFhbnd = d_fb.d();
  a.dx = d_dx.d();
  a.dx_stride = 3;
  a.dt = d_dt.d();
  a.iters = static_cast<int*>(d_it.p);
  a.cell0 = 0;
  a.lQi = d_qi.d();
  a.lFi = d_fi.d();
  a.lQhi = d_qhi.d();
  a.lFhi = d_fhi.d();
  DevBuf d_ws;
  if (const size_t wsb = exa::predictor_ws_bytes_per_cell(sd)) {
    a.ws_cells = std::min(ncells, exa::kPredictorWsCells);
    EXA_TRY(d_ws.alloc(wsb * a.ws_cells));
    a.ws = d_ws.p;
  }
  EXA_TRY(exa::launch_predictor(sd, exp, a, ncells, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(lQhbnd, d_qb.p, sizeof(double) * ncells * 6 * z.B2 * z.D, hipMemcpyDeviceToHost));
  EXA_TRY(hipMemcpy(lFhbnd, d_fb.p, sizeof(double) * ncells * 6 * z.B2 * z.V, hipMemcpyDeviceToHost));
  if (lduh) EXA_TRY(hipMemcpy(lduh, d_du.p, sizeof(double) * ncells * z.B3 * z.V, hipMemcpyDeviceToHost));
  if (iters) EXA_TRY(hipMemcpy(iters, d_it.p, sizeof(int) * ncells, hipMemcpyDeviceToHost));
  if (lQi) EXA_TRY(hipMemcpy(lQi, d_qi.p, sizeof(double) * ncells * z.B4 * z.D, hipMemcpyDeviceToHost));
  if (lFi) EXA_TRY(hipMemcpy(lFi, d_fi.p, sizeof(double) * ncells * 4 * z.B4 * z.V, hipMemcpyDeviceToHost));
  if (lQhi) EXA_TRY(hipMemcpy(lQhi, d_qhi.p, sizeof(double) * ncells * z.B3 * z.D, hipMemcpyDeviceToHost));
  if (lFhi) EXA_TRY(hipMemcpy(lFhi, d_fhi.p, sizeof(double) * ncells * 4 * z.B3 * z.V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_space_time_predictor_nonlinear(const exa_solver* s, double* lQhbnd, double* lFhbnd,
                                                  double* lQi, double* lFi, double* lQhi, double* lFhi,
                                                  const double* luh, const double* dx, double dt,
                                                  int* picard_iterations) {
  return predictor_impl(s, 1, lQhbnd, lFhbnd, nullptr, lQi, lFi, lQhi, lFhi, luh, dx, dt,
                        picard_iterations);
}

extern "C" int exa_batch_predictor_volume(const exa_solver* s, long ncells, double* lQhbnd,
                                          double* lFhbnd, double* lduh, const double* luh,
                                          const double* dx, double dt, int* picard_iterations) {
  return predictor_impl(s, ncells, lQhbnd, lFhbnd, lduh, nullptr, nullptr, nullptr, nullptr, luh, dx,
                        dt, picard_iterations);
}

extern "C" int exa_volume_integral_nonlinear(const exa_solver* s, double* lduh, const double* lFhi,
                                             const double* dx) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || !lduh || !lFhi || !dx) return EXA_E_INVALID;
  if (sd.linear || !exa::supported_sizes(sd) || sd.precision != exa::kFP64) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  DevBuf d_du, d_fhi, d_dx, d_fi;
  EXA_TRY(d_du.alloc(sizeof(double) * z.B3 * z.V));
  EXA_TRY(d_fhi.alloc(sizeof(double) * 4 * z.B3 * z.V));
  EXA_TRY(d_dx.alloc(sizeof(double) * 3));
  if (sd.no_time_averaging) {  // the argument is lFi: average it over time first
    EXA_TRY(d_fi.alloc(sizeof(double) * 4 * z.B3 * z.B * z.V));
    EXA_TRY(hipMemcpy(d_fi.p, lFhi, sizeof(double) * 4 * z.B3 * z.B * z.V, hipMemcpyHostToDevice));
    EXA_TRY(exa::launch_time_average_lfi(sd.order, z.V, d_fi.d(), d_fhi.d(), nullptr));
  } else {
    EXA_TRY(hipMemcpy(d_fhi.p, lFhi, sizeof(double) * 4 * z.B3 * z.V, hipMemcpyHostToDevice));
  }
  EXA_TRY(hipMemcpy(d_dx.p, dx, sizeof(double) * 3, hipMemcpyHostToDevice));
  exa::VolumeArgs a{d_du.d(), d_fhi.d(), d_dx.d(), 0};
  EXA_TRY(exa::launch_volume_integral(sd, a, 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(lduh, d_du.p, sizeof(double) * z.B3 * z.V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

// Shared body of the single-cell surface / update / time-step entry points.
static int update_impl(const exa_solver* s, int mode, double* luh, double* lduh, const double* lFhbnd,
                       const double* dx, double dt, double* dt_out) {
  SolverDesc sd;
  if (!to_desc(s, &sd)) return EXA_E_INVALID;
  const bool surface_only = mode == (exa::kUpdSurface | exa::kUpdWriteLduh);
  // solutionUpdate and stableTimeStepSize serve linear and nonlinear solvers
  // alike; a linear solver's surface integral is surfaceIntegralLinear
  if (surface_only && sd.linear) return EXA_E_UNSUPPORTED;
  sd.linear = 0;
  if (sd.precision != exa::kFP64) return EXA_E_UNSUPPORTED;
  if (surface_only ? !exa::supported_sizes(sd) : !exa::supported(sd)) return EXA_E_UNSUPPORTED;
  if (surface_only) {  // PDE-independent: run the Euler instantiation of this order
    sd.pde = exa::kPdeEuler;
    sd.npar = 0;
    sd.terms = exa::kFlux;
  }
  const Sizes z(sd);
  const double one[3] = {1.0, 1.0, 1.0};
  DevBuf d_luh, d_du, d_fb, d_dx, d_dt, d_cell;
  EXA_TRY(d_luh.alloc(sizeof(double) * z.B3 * z.D));
  EXA_TRY(d_du.alloc(sizeof(double) * z.B3 * z.V));
  EXA_TRY(d_fb.alloc(sizeof(double) * 6 * z.B2 * z.V));
  EXA_TRY(d_dx.alloc(sizeof(double) * 3));
  EXA_TRY(d_dt.alloc(sizeof(double)));
  EXA_TRY(d_cell.alloc(sizeof(double)));
  EXA_TRY(hipMemset(d_luh.p, 0, sizeof(double) * z.B3 * z.D));
  EXA_TRY(hipMemset(d_du.p, 0, sizeof(double) * z.B3 * z.V));
  EXA_TRY(hipMemset(d_fb.p, 0, sizeof(double) * 6 * z.B2 * z.V));
  if (luh) EXA_TRY(hipMemcpy(d_luh.p, luh, sizeof(double) * z.B3 * z.D, hipMemcpyHostToDevice));
  if (lduh) EXA_TRY(hipMemcpy(d_du.p, lduh, sizeof(double) * z.B3 * z.V, hipMemcpyHostToDevice));
  if (lFhbnd) EXA_TRY(hipMemcpy(d_fb.p, lFhbnd, sizeof(double) * 6 * z.B2 * z.V, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dx.p, dx ? dx : one, sizeof(double) * 3, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dt.p, &dt, sizeof(double), hipMemcpyHostToDevice));
  exa::UpdateArgs a{};
  a.luh = d_luh.d();
  a.lduh = d_du.d();
  a.lFhbnd = d_fb.d();
  a.dx = d_dx.d();
  a.dx_stride = 0;
  a.dt = d_dt.d();
  a.cfl_pnpm = sd.cfl * kPNPM[sd.order];
  a.dt_slots = nullptr;
  a.dt_cell = dt_out ? d_cell.d() : nullptr;
  a.cell0 = 0;
  EXA_TRY(exa::launch_update(sd, mode, a, 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  if (mode & exa::kUpdUpdate) EXA_TRY(hipMemcpy(luh, d_luh.p, sizeof(double) * z.B3 * z.D, hipMemcpyDeviceToHost));
  if (mode & exa::kUpdWriteLduh) EXA_TRY(hipMemcpy(lduh, d_du.p, sizeof(double) * z.B3 * z.V, hipMemcpyDeviceToHost));
  if (dt_out) EXA_TRY(hipMemcpy(dt_out, d_cell.p, sizeof(double), hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_surface_integral_nonlinear(const exa_solver* s, double* lduh, const double* lFhbnd,
                                              const double* dx) {
  if (!lduh || !lFhbnd || !dx) return EXA_E_INVALID;
  return update_impl(s, exa::kUpdSurface | exa::kUpdWriteLduh, nullptr, lduh, lFhbnd, dx, 0.0, nullptr);
}

extern "C" int exa_solution_update(const exa_solver* s, double* luh, const double* lduh, double dt) {
  if (!luh || !lduh) return EXA_E_INVALID;
  return update_impl(s, exa::kUpdUpdate, luh, const_cast<double*>(lduh), nullptr, nullptr, dt, nullptr);
}

extern "C" int exa_stable_time_step_size(const exa_solver* s, const double* luh, const double* dx,
                                         double* dt_out) {
  if (!luh || !dx || !dt_out) return EXA_E_INVALID;
  return update_impl(s, 0, const_cast<double*>(luh), nullptr, nullptr, dx, 0.0, dt_out);
}

extern "C" int exa_riemann_solver_nonlinear(const exa_solver* s, double* FL, double* FR, const double* QL,
                                            const double* QR, double dt, int direction) {
  (void)dt;  // unused by the nonlinear Rusanov solver, as in the reference
  SolverDesc sd;
  if (!to_desc(s, &sd) || !FL || !FR || !QL || !QR || direction < 0 || direction > 2) return EXA_E_INVALID;
  if (sd.linear || !exa::supported(sd) || sd.precision != exa::kFP64) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  // Two cells: cell 0 owns face 2d+1 (left state), cell 1 owns face 2d (right state).
  DevBuf d_qb, d_fb, d_face;
  const long qf = z.B2 * z.D, ff = z.B2 * z.V;
  EXA_TRY(d_qb.alloc(sizeof(double) * 2 * 6 * qf));
  EXA_TRY(d_fb.alloc(sizeof(double) * 2 * 6 * ff));
  EXA_TRY(d_face.alloc(sizeof(FaceDesc)));
  const long iqL = (0 * 6 + 2 * direction + 1) * qf, iqR = (1 * 6 + 2 * direction) * qf;
  const long ifL = (0 * 6 + 2 * direction + 1) * ff, ifR = (1 * 6 + 2 * direction) * ff;
  EXA_TRY(hipMemcpy(d_qb.d() + iqL, QL, sizeof(double) * qf, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_qb.d() + iqR, QR, sizeof(double) * qf, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_fb.d() + ifL, FL, sizeof(double) * ff, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_fb.d() + ifR, FR, sizeof(double) * ff, hipMemcpyHostToDevice));
  const FaceDesc f{0, 1, direction, 0};
  EXA_TRY(hipMemcpy(d_face.p, &f, sizeof(FaceDesc), hipMemcpyHostToDevice));
  exa::RiemannArgs a{static_cast<const FaceDesc*>(d_face.p), 1, d_qb.d(), d_fb.d()};
  EXA_TRY(exa::launch_riemann(sd, a, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(FL, d_fb.d() + ifL, sizeof(double) * ff, hipMemcpyDeviceToHost));
  EXA_TRY(hipMemcpy(FR, d_fb.d() + ifR, sizeof(double) * ff, hipMemcpyDeviceToHost));
  return EXA_OK;
}

// ---- linear kernels (the entry point names the kernel: the solver's linear flag is implied) ----

extern "C" int exa_space_time_predictor_linear(const exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi,
                                               double* lFi, double* gradQ, double* lQhi, double* lFhi,
                                               const double* luh, const double* dx, double dt, const double* PSi) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || !lQhbnd || !lFhbnd || !luh || !dx) return EXA_E_INVALID;
  sd.linear = 1;
  const int ntemp = (lQi != nullptr) + (lFi != nullptr) + (gradQ != nullptr) + (lQhi != nullptr) + (lFhi != nullptr);
  if (ntemp != 0 && ntemp != 5) return EXA_E_INVALID;  // the temporaries come all together or not at all
  if (!exa::linear_supported(sd)) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  const bool exp = ntemp == 5;
  const size_t nqi = (z.B + 1) * z.B3 * z.D, nfi = 4 * z.B4 * z.V, ngq = 3 * z.B4 * z.V, nqhi = z.B3 * z.D,
               nfhi = 4 * z.B3 * z.V, nps = (z.B + 1) * z.B3 * z.V;
  DevBuf d_luh, d_qb, d_fb, d_dx, d_dt, d_ps, d_qi, d_fi, d_gq, d_qhi, d_fhi;
  EXA_TRY(d_luh.alloc(sizeof(double) * z.B3 * z.D));
  EXA_TRY(d_qb.alloc(sizeof(double) * 6 * z.B2 * z.D));
  EXA_TRY(d_fb.alloc(sizeof(double) * 6 * z.B2 * z.V));
  EXA_TRY(d_dx.alloc(sizeof(double) * 3));
  EXA_TRY(d_dt.alloc(sizeof(double)));
  if (PSi) {
    EXA_TRY(d_ps.alloc(sizeof(double) * nps));
    EXA_TRY(hipMemcpy(d_ps.p, PSi, sizeof(double) * nps, hipMemcpyHostToDevice));
  }
  if (exp) {
    EXA_TRY(d_qi.alloc(sizeof(double) * nqi));
    EXA_TRY(d_fi.alloc(sizeof(double) * nfi));
    EXA_TRY(d_gq.alloc(sizeof(double) * ngq));
    EXA_TRY(d_qhi.alloc(sizeof(double) * nqhi));
    EXA_TRY(d_fhi.alloc(sizeof(double) * nfhi));
  }
  EXA_TRY(hipMemcpy(d_luh.p, luh, sizeof(double) * z.B3 * z.D, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dx.p, dx, sizeof(double) * 3, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dt.p, &dt, sizeof(double), hipMemcpyHostToDevice));
  exa::LinearPredictorArgs a{};
  a.luh = d_luh.d();
  a.lQhbnd = d_qb.d();
  a.lFhbnd = d_fb.d();
  a.dx = d_dx.d();
  a.dt = d_dt.d();
  a.PSi = PSi ? d_ps.d() : nullptr;
  if (exp) {
    a.lQi = d_qi.d();
    a.lFi = d_fi.d();
    a.gradQ = d_gq.d();
    a.lQhi = d_qhi.d();
    a.lFhi = d_fhi.d();
  }
  EXA_TRY(exa::launch_linear_predictor(sd, exp, a, 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(lQhbnd, d_qb.p, sizeof(double) * 6 * z.B2 * z.D, hipMemcpyDeviceToHost));
  EXA_TRY(hipMemcpy(lFhbnd, d_fb.p, sizeof(double) * 6 * z.B2 * z.V, hipMemcpyDeviceToHost));
  if (exp) {
    EXA_TRY(hipMemcpy(lQi, d_qi.p, sizeof(double) * nqi, hipMemcpyDeviceToHost));
    EXA_TRY(hipMemcpy(lFi, d_fi.p, sizeof(double) * nfi, hipMemcpyDeviceToHost));
    EXA_TRY(hipMemcpy(gradQ, d_gq.p, sizeof(double) * ngq, hipMemcpyDeviceToHost));
    EXA_TRY(hipMemcpy(lQhi, d_qhi.p, sizeof(double) * nqhi, hipMemcpyDeviceToHost));
    EXA_TRY(hipMemcpy(lFhi, d_fhi.p, sizeof(double) * nfhi, hipMemcpyDeviceToHost));
  }
  return EXA_OK;
}

extern "C" int exa_point_source_projection(const exa_solver* s, double* pointForceSources, const double* x0,
                                           const double* center, const double* dx) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || !pointForceSources || !x0 || !center || !dx) return EXA_E_INVALID;
  if (sd.order < 1 || sd.order > 7) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  const size_t nodes = (z.B + 1) * z.B3;
  DevBuf d_ps, d_x0, d_geo;
  EXA_TRY(d_ps.alloc(sizeof(double) * nodes * z.V));
  EXA_TRY(d_x0.alloc(sizeof(double) * nodes * 3));
  EXA_TRY(d_geo.alloc(sizeof(double) * 6));
  const double geo[6] = {center[0], center[1], center[2], dx[0], dx[1], dx[2]};
  EXA_TRY(hipMemcpy(d_ps.p, pointForceSources, sizeof(double) * nodes * z.V, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_x0.p, x0, sizeof(double) * nodes * 3, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_geo.p, geo, sizeof(geo), hipMemcpyHostToDevice));
  EXA_TRY(exa::launch_point_source(sd, d_ps.d(), d_x0.d(), d_geo.d(), nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(pointForceSources, d_ps.p, sizeof(double) * nodes * z.V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" double exa_quadrature_node(int order, int i) {
  // the tables as constant-initialised locals (host code; only the nodes are used)
#define EXA_TABLE static const
#include "dg_tables.inc"
#undef EXA_TABLE
  if (order < 0 || order > 9 || i < 0 || i > order) return 0.0;
  return exa_gl_nodes[order][i];
}

extern "C" int exa_volume_integral_linear(const exa_solver* s, double* lduh, const double* lFhi, const double* dx) {
  (void)dx;  // unused by the linear volume integral, as in the reference
  SolverDesc sd;
  if (!to_desc(s, &sd) || !lduh || !lFhi) return EXA_E_INVALID;
  sd.linear = 1;
  if (!exa::linear_sizes_supported(sd)) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  DevBuf d_du, d_fhi;
  EXA_TRY(d_du.alloc(sizeof(double) * z.B3 * z.V));
  EXA_TRY(d_fhi.alloc(sizeof(double) * 4 * z.B3 * z.V));
  EXA_TRY(hipMemcpy(d_fhi.p, lFhi, sizeof(double) * 4 * z.B3 * z.V, hipMemcpyHostToDevice));
  EXA_TRY(exa::launch_linear_volume(sd, d_du.d(), d_fhi.d(), 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(lduh, d_du.p, sizeof(double) * z.B3 * z.V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_surface_integral_linear(const exa_solver* s, double* lduh, const double* lFhbnd,
                                           const double* dx) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || !lduh || !lFhbnd || !dx) return EXA_E_INVALID;
  sd.linear = 1;
  if (!exa::linear_sizes_supported(sd)) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  DevBuf d_du, d_fb, d_dx;
  EXA_TRY(d_du.alloc(sizeof(double) * z.B3 * z.V));
  EXA_TRY(d_fb.alloc(sizeof(double) * 6 * z.B2 * z.V));
  EXA_TRY(d_dx.alloc(sizeof(double) * 3));
  EXA_TRY(hipMemcpy(d_du.p, lduh, sizeof(double) * z.B3 * z.V, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_fb.p, lFhbnd, sizeof(double) * 6 * z.B2 * z.V, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_dx.p, dx, sizeof(double) * 3, hipMemcpyHostToDevice));
  EXA_TRY(exa::launch_linear_surface(sd, d_du.d(), d_fb.d(), d_dx.d(), 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(lduh, d_du.p, sizeof(double) * z.B3 * z.V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_riemann_solver_linear(const exa_solver* s, double* FL, double* FR, const double* QL,
                                         const double* QR, double dt, int direction) {
  (void)dt;  // unused by the linear Riemann solver, as in the reference
  SolverDesc sd;
  if (!to_desc(s, &sd) || !FL || !FR || !QL || !QR || direction < 0 || direction > 2) return EXA_E_INVALID;
  sd.linear = 1;
  if (!exa::linear_supported(sd)) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  // Two cells: cell 0 owns face 2d+1 (left state), cell 1 owns face 2d (right state).
  DevBuf d_qb, d_fb, d_face;
  const long qf = z.B2 * z.D, ff = z.B2 * z.V;
  EXA_TRY(d_qb.alloc(sizeof(double) * 2 * 6 * qf));
  EXA_TRY(d_fb.alloc(sizeof(double) * 2 * 6 * ff));
  EXA_TRY(d_face.alloc(sizeof(FaceDesc)));
  EXA_TRY(hipMemcpy(d_qb.d() + (2 * direction + 1) * qf, QL, sizeof(double) * qf, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_qb.d() + (6 + 2 * direction) * qf, QR, sizeof(double) * qf, hipMemcpyHostToDevice));
  const FaceDesc f{0, 1, direction, 0};
  EXA_TRY(hipMemcpy(d_face.p, &f, sizeof(FaceDesc), hipMemcpyHostToDevice));
  exa::RiemannArgs a{static_cast<const FaceDesc*>(d_face.p), 1, d_qb.d(), d_fb.d()};
  EXA_TRY(exa::launch_linear_riemann(sd, a, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(FL, d_fb.d() + (2 * direction + 1) * ff, sizeof(double) * ff, hipMemcpyDeviceToHost));
  EXA_TRY(hipMemcpy(FR, d_fb.d() + (6 + 2 * direction) * ff, sizeof(double) * ff, hipMemcpyDeviceToHost));
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------
struct ProfRec {
  int stage;
  hipEvent_t b, e;
};

struct exa_grid {
  SolverDesc sd;
  int nx, ny, nz;
  long ncells;
  double lo[3], hi[3], dx[3];
  bool periodic = true;
  int boundary = exa::kBoundaryPeriodic;
  bool fused_corrector = true;  // Riemann + surface + update + dt in one kernel (nonlinear solvers)
  bool adjust = false;          // the PDE's adjust_point_solution after every update
  bool fused_ts = false;        // fused time stepping (fuse-algorithmic-steps)
  double fused_weight = 0.9;    // fuse-algorithmic-steps-factor
  long nfaces;
  hipStream_t stream = nullptr;
  double* luh = nullptr;
  void* lduh = nullptr;    // bulk arrays: double (FP64 form) or float (FP32-bulk form)
  void* lQhbnd = nullptr;
  void* lFhbnd = nullptr;
  size_t bulk = 8;         // bytes per bulk element
  double* d_dx = nullptr;
  int* iters = nullptr;
  void* ws = nullptr;      // predictor workspace (large orders only)
  long ws_cells = 0;
  size_t ws_bytes = 0;
  int predictor_variant = exa::kPredictorAuto;
  const exa_user_solver* user = nullptr;  // a user solver's kernels (exa_grid_create_user)
  FaceDesc* faces = nullptr;
  exa::StepState* st = nullptr;
  double* dt_hist = nullptr;
  long dt_hist_len = 0;
  bool predicted = false;
  bool profiling = false;
  std::vector<ProfRec> prof;
  double prof_ms[EXA_NUM_STAGES] = {0, 0, 0, 0};
  long prof_n[EXA_NUM_STAGES] = {0, 0, 0, 0};
};

namespace {

template <class F>
hipError_t staged(exa_grid* g, int stage, F&& launch) {
  if (!g->profiling) return launch();
  ProfRec r{stage, nullptr, nullptr};
  hipError_t e = hipEventCreate(&r.b);
  if (e == hipSuccess) e = hipEventCreate(&r.e);
  if (e == hipSuccess) e = hipEventRecord(r.b, g->stream);
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(r.e, g->stream);
  g->prof.push_back(r);
  return e;
}

void free_grid(exa_grid* g) {
  if (!g) return;
  for (auto& r : g->prof) {
    (void)hipEventDestroy(r.b);
    (void)hipEventDestroy(r.e);
  }
  (void)hipFree(g->luh);
  (void)hipFree(g->lduh);
  (void)hipFree(g->lQhbnd);
  (void)hipFree(g->lFhbnd);
  (void)hipFree(g->d_dx);
  (void)hipFree(g->iters);
  (void)hipFree(g->ws);
  (void)hipFree(g->faces);
  (void)hipFree(g->st);
  (void)hipFree(g->dt_hist);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

// initial dt: dt of every cell into the slots, then the global minimum.
hipError_t launch_initial_dt(exa_grid* g) {
  exa::UpdateArgs a{};
  a.luh = g->luh;
  a.lduh = g->lduh;
  a.lFhbnd = g->lFhbnd;
  a.dx = g->d_dx;
  a.dx_stride = 0;
  a.dt = &g->st->dt;
  a.cfl_pnpm = g->sd.cfl * kPNPM[g->sd.order];
  a.dt_slots = g->st->dt_slots;
  a.dt_cell = nullptr;
  a.cell0 = 0;
  hipError_t e = staged(g, EXA_STAGE_UPDATE, [&] {
    return g->user ? (hipError_t)g->user->stable_dt(&a, g->ncells, g->stream)
                   : exa::launch_update(g->sd, 0, a, g->ncells, g->stream);
  });
  if (e != hipSuccess) return e;
  return staged(g, EXA_STAGE_DT, [&] {
    return g->fused_ts ? exa::launch_dt_finalize_fused(g->st, g->fused_weight, true, g->stream)
                       : exa::launch_dt_finalize(g->st, false, g->stream);
  });
}

// The prediction with the time step size at dt (device scalar); run_if (a
// device flag, may be null) gates the whole launch.
hipError_t launch_predict(exa_grid* g, const double* dt, const int* run_if = nullptr) {
  if (g->sd.linear) {  // Cauchy-Kowalevski predictor + volumeIntegralLinear
    exa::LinearPredictorArgs la{};
    la.luh = g->luh;
    la.lduh = static_cast<double*>(g->lduh);
    la.lQhbnd = static_cast<double*>(g->lQhbnd);
    la.lFhbnd = static_cast<double*>(g->lFhbnd);
    la.dx = g->d_dx;
    la.dx_stride = 0;
    la.dt = dt;
    la.cell0 = 0;
    if (run_if) return hipErrorNotSupported;  // fused time stepping is for nonlinear solvers
    return staged(g, EXA_STAGE_PREDICTOR,
                  [&] { return exa::launch_linear_predictor(g->sd, false, la, g->ncells, g->stream); });
  }
  exa::PredictorArgs a{};
  a.luh = g->luh;
  a.lduh = g->lduh;
  a.lQhbnd = g->lQhbnd;
  a.lFhbnd = g->lFhbnd;
  a.dx = g->d_dx;
  a.dx_stride = 0;
  a.dt = dt;
  a.iters = g->iters;
  a.cell0 = 0;
  a.ws = g->ws;
  a.ws_cells = g->ws_cells;
  a.variant = g->predictor_variant;
  a.run_if = run_if;
  hipError_t e = staged(g, EXA_STAGE_PREDICTOR, [&] {
    return g->user ? (hipError_t)g->user->predictor(&a, g->ncells, g->sd.precision, g->stream)
                   : exa::launch_predictor(g->sd, false, a, g->ncells, g->stream);
  });
  if (e == hipSuccess && g->profiling) e = exa::launch_count_iterations(g->iters, g->ncells, g->st, g->stream);
  return e;
}

// The corrector of one step: Riemann problems, surface integral, update with
// the corrector dt, the admissible dt of every cell into the dt slots.
hipError_t launch_correct(exa_grid* g) {
  if (!g->sd.linear && (g->fused_corrector || g->user)) {
    exa::CorrectorArgs ca{};
    ca.luh = g->luh;
    ca.lduh = g->lduh;
    ca.lQhbnd = g->lQhbnd;
    ca.lFhbnd = g->lFhbnd;
    ca.dx = g->d_dx;
    ca.dt = &g->st->dt;
    ca.cfl_pnpm = g->sd.cfl * kPNPM[g->sd.order];
    ca.dt_slots = g->st->dt_slots;
    ca.t = &g->st->t;
    for (int d = 0; d < 3; ++d) {
      ca.lo[d] = g->lo[d];
      ca.hdx[d] = g->dx[d];
    }
    ca.nx = g->nx;
    ca.ny = g->ny;
    ca.nz = g->nz;
    ca.boundary = g->boundary;
    ca.adjust = g->adjust ? 1 : 0;
    return staged(g, EXA_STAGE_UPDATE, [&] {
      return g->user ? (hipError_t)g->user->corrector(&ca, g->sd.precision, g->stream)
                     : exa::launch_corrector(g->sd, ca, g->stream);
    });
  }
  exa::RiemannArgs ra{g->faces, g->nfaces, g->lQhbnd, g->lFhbnd};
  hipError_t e = staged(g, EXA_STAGE_RIEMANN, [&] {
    return g->sd.linear ? exa::launch_linear_riemann(g->sd, ra, g->stream) : exa::launch_riemann(g->sd, ra, g->stream);
  });
  if (e != hipSuccess) return e;
  exa::UpdateArgs ua{};
  ua.luh = g->luh;
  ua.lduh = g->lduh;
  ua.lFhbnd = g->lFhbnd;
  ua.dx = g->d_dx;
  ua.dx_stride = 0;
  ua.dt = &g->st->dt;
  ua.cfl_pnpm = g->sd.cfl * kPNPM[g->sd.order];
  ua.dt_slots = g->st->dt_slots;
  ua.dt_cell = nullptr;
  ua.cell0 = 0;
  return staged(g, EXA_STAGE_UPDATE, [&] {
    return exa::launch_update(g->sd, exa::kUpdSurface | exa::kUpdUpdate, ua, g->ncells, g->stream);
  });
}

// One step.  Standard: correct, the new global dt, predict with it.  Fused
// time stepping (Runner::runOneTimeStepWithFusedAlgorithmicSteps, Runner.cpp:
// 1053-1110; ADERDGSolver::fusedTimeStep, ADERDGSolver.cpp:1853-1879):
// correct, predict with the predictor dt chosen one step earlier (no global
// reduction in between), the stability check, and the prediction again,
// with the new predictor dt, when the check failed (a launch that exits at
// once otherwise: no host round trip).
hipError_t launch_one_step(exa_grid* g) {
  hipError_t e = launch_correct(g);
  if (e != hipSuccess) return e;
  if (!g->fused_ts) {
    e = staged(g, EXA_STAGE_DT, [&] { return exa::launch_dt_finalize(g->st, true, g->stream); });
    if (e != hipSuccess) return e;
    return launch_predict(g, &g->st->dt);
  }
  e = launch_predict(g, &g->st->dtp);
  if (e != hipSuccess) return e;
  e = staged(g, EXA_STAGE_DT,
             [&] { return exa::launch_dt_finalize_fused(g->st, g->fused_weight, false, g->stream); });
  if (e != hipSuccess) return e;
  return launch_predict(g, &g->st->dtp, &g->st->violated);
}

}  // namespace

// Grid allocation and face lists, shared by the library's solvers and user
// solvers (validated by the callers).
static int grid_create_impl(const SolverDesc& sd, const exa_user_solver* user, int nx, int ny, int nz,
                            const double* lo, const double* hi, int boundary, exa_grid** out) {
  const bool periodic = boundary == EXA_BC_PERIODIC;
  // Linear solvers with transmissive boundaries: the linear Riemann solver's
  // fluctuations vanish at a boundary face (outside = inside), so those faces
  // join the face list to be zeroed.
  const bool linear_bc = sd.linear && !periodic;
  if ((double)nx * ny * nz > 2.0e9) return EXA_E_INVALID;
  exa_grid* g = new exa_grid();
  g->sd = sd;
  g->user = user;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->ncells = (long)nx * ny * nz;
  for (int d = 0; d < 3; ++d) {
    g->lo[d] = lo[d];
    g->hi[d] = hi[d];
  }
  g->periodic = periodic;
  g->boundary = boundary == EXA_BC_PERIODIC ? exa::kBoundaryPeriodic
                                            : (boundary == EXA_BC_USER ? exa::kBoundaryUser : exa::kBoundaryTransmissive);
  g->dx[0] = (hi[0] - lo[0]) / nx;
  g->dx[1] = (hi[1] - lo[1]) / ny;
  g->dx[2] = (hi[2] - lo[2]) / nz;
  const Sizes z(sd);
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  g->bulk = sd.precision == exa::kFP32Bulk ? sizeof(float) : sizeof(double);
  A((void**)&g->luh, sizeof(double) * g->ncells * z.B3 * z.D);
  A((void**)&g->lduh, g->bulk * g->ncells * z.B3 * z.V);
  A((void**)&g->lQhbnd, g->bulk * g->ncells * 6 * z.B2 * z.D);
  A((void**)&g->lFhbnd, g->bulk * g->ncells * 6 * z.B2 * z.V);
  A((void**)&g->d_dx, sizeof(double) * 3);
  A((void**)&g->iters, sizeof(int) * g->ncells);
  if (const size_t wsb = sd.linear ? 0 : exa::predictor_ws_bytes_per_cell(sd)) {
    g->ws_cells = std::min(g->ncells, exa::kPredictorWsCells);
    g->ws_bytes = wsb * g->ws_cells;
    A(&g->ws, g->ws_bytes);
  }
  // Periodic: every cell owns its +x, +y, +z face.  Transmissive: interior
  // faces only -- at a boundary face the reference solves a Riemann problem
  // against boundaryValues' outside state (ADERDGSolver.cpp:2893-2966), and
  // with transmissive values (outside = inside) Rusanov returns the inner
  // flux unchanged, so those faces need no work.
  g->nfaces = periodic ? 3 * g->ncells : (long)(nx - 1) * ny * nz + (long)nx * (ny - 1) * nz + (long)nx * ny * (nz - 1);
  if (linear_bc) g->nfaces += 2 * ((long)ny * nz + (long)nx * nz + (long)nx * ny);
  A((void**)&g->faces, sizeof(FaceDesc) * g->nfaces);
  A((void**)&g->st, sizeof(exa::StepState));
  g->dt_hist_len = 1 << 16;
  A((void**)&g->dt_hist, sizeof(double) * g->dt_hist_len);
  if (e != hipSuccess) {
    free_grid(g);
    return from_hip(e);
  }
  std::vector<FaceDesc> faces(g->nfaces);
  auto cid = [&](int i, int j, int k) {
    i = (i + nx) % nx;
    j = (j + ny) % ny;
    k = (k + nz) % nz;
    return (int)(((long)k * ny + j) * nx + i);
  };
  long f = 0;
  for (int dir = 0; dir < 3; ++dir)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const int up = dir == 0 ? i + 1 : (dir == 1 ? j + 1 : k + 1);
          const int nmax = dir == 0 ? nx : (dir == 1 ? ny : nz);
          if (!periodic && up >= nmax) continue;
          const int c = cid(i, j, k);
          const int r = dir == 0 ? cid(i + 1, j, k) : (dir == 1 ? cid(i, j + 1, k) : cid(i, j, k + 1));
          faces[f++] = FaceDesc{c, r, dir, 0};
        }
  if (linear_bc)
    for (int dir = 0; dir < 3; ++dir)
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            const int pos = dir == 0 ? i : (dir == 1 ? j : k);
            const int nmax = dir == 0 ? nx : (dir == 1 ? ny : nz);
            const int c = cid(i, j, k);
            if (pos == 0) faces[f++] = FaceDesc{c, c, dir, 1};
            if (pos == nmax - 1) faces[f++] = FaceDesc{c, c, dir, 2};
          }
  exa::StepState st0{};
  st0.dt = 0.0;
  st0.t = 0.0;
  st0.steps = 0;
  st0.dt_hist = g->dt_hist;
  st0.dt_hist_len = g->dt_hist_len;
  for (int i = 0; i < exa::kDtSlots; ++i) st0.dt_slots[i] = exa::kDtSlotEmpty;
  e = hipMemcpy(g->faces, faces.data(), sizeof(FaceDesc) * g->nfaces, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->st, &st0, sizeof(st0), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_dx, g->dx, sizeof(double) * 3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(g->lduh, 0, g->bulk * g->ncells * z.B3 * z.V);
  if (e == hipSuccess) e = hipMemset(g->iters, 0, sizeof(int) * g->ncells);
  if (e != hipSuccess) {
    free_grid(g);
    return from_hip(e);
  }
  *out = g;
  return EXA_OK;
}

extern "C" int exa_grid_create(const exa_solver* s, int nx, int ny, int nz, const double* lo,
                               const double* hi, int boundary, exa_grid** out) {
  if (!out) return EXA_E_INVALID;
  *out = nullptr;
  SolverDesc sd;
  if (!to_desc(s, &sd) || nx < 1 || ny < 1 || nz < 1 || !lo || !hi) return EXA_E_INVALID;
  if (boundary != EXA_BC_PERIODIC && boundary != EXA_BC_TRANSMISSIVE && boundary != EXA_BC_USER)
    return EXA_E_UNSUPPORTED;
  if (boundary == EXA_BC_USER && (sd.linear || !exa::corrector_supports(sd, exa::kBoundaryUser, false)))
    return EXA_E_UNSUPPORTED;
  if (sd.linear ? !exa::linear_supported(sd) : !exa::supported(sd)) return EXA_E_UNSUPPORTED;
  return grid_create_impl(sd, nullptr, nx, ny, nz, lo, hi, boundary, out);
}

extern "C" int exa_grid_create_user(const exa_user_solver* us, int precision, int nx, int ny, int nz,
                                    const double* lo, const double* hi, int boundary, exa_grid** out) {
  if (!out) return EXA_E_INVALID;
  *out = nullptr;
  if (!us || nx < 1 || ny < 1 || nz < 1 || !lo || !hi) return EXA_E_INVALID;
  if (us->abi_version != EXA_USER_SOLVER_ABI || !us->predictor || !us->corrector || !us->stable_dt ||
      !us->adjust)
    return EXA_E_INVALID;
  if (us->order < 1 || us->order > 7 || us->num_variables < 1 || us->num_parameters < 0 ||
      us->num_variables + us->num_parameters > 64 || !(us->cfl > 0))
    return EXA_E_INVALID;
  if (precision != EXA_PRECISION_FP64 && precision != EXA_PRECISION_FP32_BULK) return EXA_E_INVALID;
  if (boundary != EXA_BC_PERIODIC && boundary != EXA_BC_TRANSMISSIVE && boundary != EXA_BC_USER)
    return EXA_E_UNSUPPORTED;
  if (boundary == EXA_BC_USER && !us->has_boundary_values) return EXA_E_UNSUPPORTED;
  SolverDesc sd{};
  sd.pde = exa::kPdeUser;
  sd.order = us->order;
  sd.nvar = us->num_variables;
  sd.npar = us->num_parameters;
  sd.terms = us->terms;
  sd.cfl = us->cfl;
  sd.precision = precision;
  return grid_create_impl(sd, us, nx, ny, nz, lo, hi, boundary, out);
}

extern "C" int exa_grid_set_adjust_solution(exa_grid* g, int on) {
  if (!g) return EXA_E_INVALID;
  if (on && !(g->user && g->user->has_adjust)) return EXA_E_UNSUPPORTED;
  g->adjust = on != 0;
  if (on) g->fused_corrector = true;
  return EXA_OK;
}

extern "C" int exa_grid_adjust_solution(exa_grid* g) {
  if (!g) return EXA_E_INVALID;
  if (!(g->user && g->user->has_adjust)) return EXA_E_UNSUPPORTED;
  exa::AdjustArgs a{};
  a.luh = g->luh;
  a.t = &g->st->t;
  a.dt = 0.0;
  for (int d = 0; d < 3; ++d) {
    a.lo[d] = g->lo[d];
    a.dx[d] = g->dx[d];
  }
  a.nx = g->nx;
  a.ny = g->ny;
  a.nz = g->nz;
  EXA_TRY((hipError_t)g->user->adjust(&a, g->ncells, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  g->predicted = false;  // the solution changed: the next step starts with a new dt and prediction
  return EXA_OK;
}

extern "C" int exa_grid_destroy(exa_grid* g) {
  if (g && g->stream) (void)hipStreamSynchronize(g->stream);
  free_grid(g);
  return EXA_OK;
}

extern "C" long exa_grid_num_cells(const exa_grid* g) { return g ? g->ncells : 0; }

static int reset_time(exa_grid* g) {
  exa::StepState st0{};
  st0.dt_hist = g->dt_hist;
  st0.dt_hist_len = g->dt_hist_len;
  for (int i = 0; i < exa::kDtSlots; ++i) st0.dt_slots[i] = exa::kDtSlotEmpty;
  EXA_TRY(hipMemcpyAsync(g->st, &st0, sizeof(st0), hipMemcpyHostToDevice, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  g->predicted = false;
  return EXA_OK;
}

extern "C" int exa_grid_upload_solution(exa_grid* g, const double* luh) {
  if (!g || !luh) return EXA_E_INVALID;
  const Sizes z(g->sd);
  EXA_TRY(hipMemcpyAsync(g->luh, luh, sizeof(double) * g->ncells * z.B3 * z.D, hipMemcpyHostToDevice, g->stream));
  return reset_time(g);
}

extern "C" int exa_grid_download_solution(exa_grid* g, double* luh) {
  if (!g || !luh) return EXA_E_INVALID;
  const Sizes z(g->sd);
  EXA_TRY(hipMemcpyAsync(luh, g->luh, sizeof(double) * g->ncells * z.B3 * z.D, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_init_elastic(exa_grid* g, unsigned long long seed, double width, double amp) {
  if (!g || !(width > 0)) return EXA_E_INVALID;
  if (g->sd.pde != EXA_PDE_ELASTIC) return EXA_E_UNSUPPORTED;
  exa::ElasticICArgs a{};
  a.luh = g->luh;
  a.ncells = g->ncells;
  a.nx = g->nx;
  a.ny = g->ny;
  a.nz = g->nz;
  a.order = g->sd.order;
  for (int d = 0; d < 3; ++d) {
    a.lo[d] = g->lo[d];
    a.dx[d] = g->dx[d];
  }
  for (int c = 0; c < 6; ++c) {  // counter-based: state = seed + (c+1) * golden, one mixing round
    uint64_t st = seed + (uint64_t)c * 0x9E3779B97F4A7C15ull;
    a.r[c] = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
  }
  a.width = width;
  a.amp = amp;
  EXA_TRY(exa::launch_elastic_ic(a, g->stream));
  return reset_time(g);
}

extern "C" int exa_grid_init_euler(exa_grid* g, const exa_euler_ic* ic) {
  if (!g || !ic || ic->nmodes < 0 || ic->nmodes > exa::kMaxModes) return EXA_E_INVALID;
  if (g->sd.pde != EXA_PDE_EULER) return EXA_E_UNSUPPORTED;
  exa::ICArgs a{};
  a.luh = g->luh;
  a.ncells = g->ncells;
  a.nx = g->nx;
  a.ny = g->ny;
  a.nz = g->nz;
  a.order = g->sd.order;
  a.nvar = g->sd.nvar;
  a.npar = g->sd.npar;
  for (int d = 0; d < 3; ++d) {
    a.lo[d] = g->lo[d];
    a.dx[d] = g->dx[d];
  }
  fill_euler_ic(ic, &a);
  EXA_TRY(exa::launch_euler_ic(a, g->stream));
  return reset_time(g);
}

extern "C" int exa_grid_step(exa_grid* g, int nsteps) {
  if (!g || nsteps < 0) return EXA_E_INVALID;
  if (!g->predicted) {
    EXA_TRY(launch_initial_dt(g));
    EXA_TRY(launch_predict(g, g->fused_ts ? &g->st->dtp : &g->st->dt));
    g->predicted = true;
  }
  for (int n = 0; n < nsteps; ++n) EXA_TRY(launch_one_step(g));
  return EXA_OK;
}

extern "C" int exa_grid_synchronize(exa_grid* g) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_time(exa_grid* g, double* t, double* dt, long* steps) {
  if (!g) return EXA_E_INVALID;
  exa::StepState st;
  EXA_TRY(hipMemcpyAsync(&st, g->st, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  if (t) *t = st.t;
  if (dt) *dt = st.dt;
  if (steps) *steps = st.steps;
  if (!std::isfinite(st.dt)) return EXA_E_NONFINITE;
  return EXA_OK;
}

extern "C" int exa_grid_dt_history(exa_grid* g, double* out, long n) {
  if (!g || !out || n < 0 || n > g->dt_hist_len) return EXA_E_INVALID;
  EXA_TRY(hipMemcpyAsync(out, g->dt_hist, sizeof(double) * n, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_download_cells(exa_grid* g, long first, long count, double* luh) {
  if (!g || !luh || first < 0 || count < 0 || first + count > g->ncells) return EXA_E_INVALID;
  const Sizes z(g->sd);
  const size_t per = (size_t)z.B3 * z.D;
  EXA_TRY(hipMemcpyAsync(luh, g->luh + first * per, sizeof(double) * count * per, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_picard_totals(exa_grid* g, unsigned long long* iterations, unsigned long long* cells) {
  if (!g) return EXA_E_INVALID;
  exa::StepState st;
  EXA_TRY(hipMemcpyAsync(&st, g->st, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  if (iterations) *iterations = st.picard_total;
  if (cells) *cells = st.predictor_cells;
  return EXA_OK;
}

extern "C" int exa_grid_picard_iterations(exa_grid* g, int* out) {
  if (!g || !out) return EXA_E_INVALID;
  EXA_TRY(hipMemcpyAsync(out, g->iters, sizeof(int) * g->ncells, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_set_profiling(exa_grid* g, int on) {
  if (!g) return EXA_E_INVALID;
  g->profiling = on != 0;
  return EXA_OK;
}

extern "C" int exa_grid_set_fused_corrector(exa_grid* g, int on) {
  if (!g) return EXA_E_INVALID;
  if (on && g->sd.linear) return EXA_E_UNSUPPORTED;
  // boundary values and the per-step adjustment run in the one-kernel corrector
  if (!on && (g->boundary == exa::kBoundaryUser || g->adjust || g->user)) return EXA_E_UNSUPPORTED;
  g->fused_corrector = on != 0;
  return EXA_OK;
}

extern "C" int exa_grid_set_fused_time_stepping(exa_grid* g, int on, double weight) {
  if (!g) return EXA_E_INVALID;
  if (on && !(weight > 0.0 && weight <= 1.0)) return EXA_E_INVALID;
  if (on && g->sd.linear) return EXA_E_UNSUPPORTED;
  EXA_TRY(hipStreamSynchronize(g->stream));
  g->fused_ts = on != 0;
  if (on) g->fused_weight = weight;
  g->predicted = false;  // the next step starts with a new dt and prediction
  return EXA_OK;
}

extern "C" int exa_grid_predictor_reruns(exa_grid* g, long* reruns) {
  if (!g || !reruns) return EXA_E_INVALID;
  exa::StepState st;
  EXA_TRY(hipMemcpyAsync(&st, g->st, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  *reruns = (long)st.reruns;
  return EXA_OK;
}

extern "C" int exa_grid_set_predictor_variant(exa_grid* g, int variant) {
  if (!g || variant < EXA_PREDICTOR_AUTO || variant > EXA_PREDICTOR_MFMA) return EXA_E_INVALID;
  if (!exa::predictor_variant_available(g->sd, variant)) return EXA_E_UNSUPPORTED;
  // the variant's Picard workspace (vector-ALU kernels at large orders)
  const long cells = std::min(g->ncells, exa::kPredictorWsCells);
  const size_t need = exa::predictor_ws_bytes_per_cell(g->sd, variant) * cells;
  if (need > g->ws_bytes) {
    EXA_TRY(hipStreamSynchronize(g->stream));
    void* ws = nullptr;
    EXA_TRY(hipMalloc(&ws, need));
    (void)hipFree(g->ws);
    g->ws = ws;
    g->ws_bytes = need;
  }
  if (need > 0) g->ws_cells = cells;
  g->predictor_variant = variant;
  return EXA_OK;
}

extern "C" int exa_grid_stage_times(exa_grid* g, double* ms, long* launches) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  for (auto& r : g->prof) {
    float t = 0.0f;
    EXA_TRY(hipEventElapsedTime(&t, r.b, r.e));
    g->prof_ms[r.stage] += t;
    g->prof_n[r.stage] += 1;
    (void)hipEventDestroy(r.b);
    (void)hipEventDestroy(r.e);
  }
  g->prof.clear();
  for (int i = 0; i < EXA_NUM_STAGES; ++i) {
    if (ms) ms[i] = g->prof_ms[i];
    if (launches) launches[i] = g->prof_n[i];
    g->prof_ms[i] = 0.0;
    g->prof_n[i] = 0;
  }
  return EXA_OK;
}

// ---- output: the Cartesian VTK plotter's vertex values ----

extern "C" int exa_plot_vertices(const exa_solver* s, long ncells, const double* luh, double* vertices) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || ncells < 0 || (ncells > 0 && (!luh || !vertices))) return EXA_E_INVALID;
  if (sd.order < 1 || sd.order > 6) return EXA_E_UNSUPPORTED;
  const Sizes z(sd);
  const size_t n = (size_t)ncells * z.B3 * z.D;
  DevBuf d_in, d_out;
  EXA_TRY(d_in.alloc(sizeof(double) * n));
  EXA_TRY(d_out.alloc(sizeof(double) * n));
  EXA_TRY(hipMemcpy(d_in.p, luh, sizeof(double) * n, hipMemcpyHostToDevice));
  EXA_TRY(exa::launch_plot_vertices(sd.order, (int)z.D, d_in.d(), d_out.d(), ncells, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(vertices, d_out.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_grid_plot_vertices(exa_grid* g, double* vertices, float* elapsed_ms) {
  if (!g || !vertices) return EXA_E_INVALID;
  if (g->sd.order < 1 || g->sd.order > 6) return EXA_E_UNSUPPORTED;
  const Sizes z(g->sd);
  const size_t n = (size_t)g->ncells * z.B3 * z.D;
  DevBuf d_out;
  EXA_TRY(d_out.alloc(sizeof(double) * n));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    EXA_TRY(hipEventCreate(&e0));
    EXA_TRY(hipEventCreate(&e1));
    EXA_TRY(hipEventRecord(e0, g->stream));
  }
  const hipError_t le = exa::launch_plot_vertices(g->sd.order, (int)z.D, g->luh, d_out.d(), g->ncells, g->stream);
  if (elapsed_ms && le == hipSuccess) (void)hipEventRecord(e1, g->stream);
  const hipError_t se = hipStreamSynchronize(g->stream);
  if (elapsed_ms) {
    if (le == hipSuccess && se == hipSuccess) (void)hipEventElapsedTime(elapsed_ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  EXA_TRY(le);
  EXA_TRY(se);
  EXA_TRY(hipMemcpy(vertices, d_out.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return EXA_OK;
}
