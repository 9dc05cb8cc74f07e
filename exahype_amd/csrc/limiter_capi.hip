// limiter_capi.hip -- C-ABI of the subcell limiter (include/exahype_amd.h):
// single-cell drop-ins for kernels::limiter::generic::c and the MUSCL-Hancock
// update, and the device-resident limiting grid.
#include <hip/hip_runtime.h>

#include <cfloat>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "capi_util.hpp"
#include "exa_device.hpp"
#include "limiter.hpp"

namespace {

struct LProf {
  int stage;
  hipEvent_t b, e;
};

// Spherical blast at GL nodes (subcells = false) or FV subcell centres; noise
// from splitmix64 keyed by the global node / subcell index.  Evaluated without
// contraction so the host-side generator of the tests reproduces it.
#pragma clang fp contract(off)
__global__ __launch_bounds__(256) void blast_kernel(double* out, int nx, int ny, int nz, int P, int subcells,
                                                   double lo0, double lo1, double lo2, double dx0, double dx1,
                                                   double dx2, unsigned long long seed, double radius, double p_in,
                                                   double p_out, double noise, int order) {
  const long n = (long)nx * ny * nz * P * P * P;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % P), y = (int)((i / P) % P), z = (int)((i / (P * P)) % P);
  const long c = i / ((long)P * P * P);
  const int ci = (int)(c % nx), cj = (int)((c / nx) % ny), ck = (int)(c / ((long)nx * ny));
  double xi[3];
  const int loc[3] = {x, y, z};
  for (int d = 0; d < 3; ++d)
    xi[d] = subcells ? (2.0 * loc[d] + 1.0) / (2.0 * P) : exa::exa_gl_nodes[order][loc[d]];
  const double X = (lo0 + (ci + 0.5) * dx0) + dx0 * (xi[0] - 0.5);
  const double Y = (lo1 + (cj + 0.5) * dx1) + dx1 * (xi[1] - 0.5);
  const double Z = (lo2 + (ck + 0.5) * dx2) + dx2 * (xi[2] - 0.5);
  const unsigned long long gx = (unsigned long long)ci * P + x, gy = (unsigned long long)cj * P + y,
                           gz = (unsigned long long)ck * P + z;
  const unsigned long long gid = (gz * ((unsigned long long)ny * P) + gy) * ((unsigned long long)nx * P) + gx;
  unsigned long long h = seed + (gid + 1ull) * 0x9E3779B97F4A7C15ull;
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h = h ^ (h >> 31);
  const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
  const double rho = 1.0 + noise * (2.0 * u - 1.0);
  const double r = sqrt(((X - 0.5) * (X - 0.5) + (Y - 0.5) * (Y - 0.5)) + (Z - 0.5) * (Z - 0.5));
  const double p = r < radius ? p_in : p_out;
  double* q = out + i * 5;
  q[0] = rho;
  q[1] = 0.0;
  q[2] = 0.0;
  q[3] = 0.0;
  q[4] = p / (1.4 - 1.0);
}
#pragma clang fp contract(on)

__global__ void lim_count_irregular_kernel(int* flags) {
  if (threadIdx.x == 0) flags[3] += flags[exa::kLimFlagIrregular];
}

}  // namespace

struct exa_lgrid {
  SolverDesc sd;
  int nx, ny, nz;
  long ncells;
  double lo[3], hi[3], dx[3];
  int periodic;
  int h, cure, fv_godunov;
  double relax, scaling;
  hipStream_t stream = nullptr;
  double *luh = nullptr, *luh_prev = nullptr, *lduh = nullptr, *lQhbnd = nullptr, *lFhbnd = nullptr;
  double *lim = nullptr, *mn = nullptr, *mx = nullptr, *limmm = nullptr, *d_dx = nullptr;
  uint8_t *cur = nullptr, *has_patch = nullptr;
  int *st = nullptr, *pst = nullptr, *st2 = nullptr, *itc = nullptr, *nbr = nullptr, *flags = nullptr;
  int *fv_list = nullptr, *iters = nullptr, *rc_list = nullptr;
  double *loc_mm = nullptr, *dt_cell = nullptr;
  FaceDesc* faces = nullptr;
  long nfaces = 0;
  exa::StepState* state = nullptr;
  void* ws = nullptr;
  long ws_cells = 0;
  bool profiling = false;
  bool fused_ts = false;      // fused time stepping (fuse-algorithmic-steps)
  double fused_weight = 0.9;  // fuse-algorithmic-steps-factor
  std::vector<LProf> prof;
  double prof_ms[EXA_NUM_LSTAGES] = {0, 0, 0, 0, 0, 0};
  long prof_n[EXA_NUM_LSTAGES] = {0, 0, 0, 0, 0, 0};
};

namespace {

exa::LimArgs largs(exa_lgrid* g) {
  exa::LimArgs a{};
  a.luh = g->luh;
  a.luh_prev = g->luh_prev;
  a.lduh = g->lduh;
  a.lFhbnd = g->lFhbnd;
  a.lim = g->lim;
  a.cur = g->cur;
  a.has_patch = g->has_patch;
  a.st = g->st;
  a.pst = g->pst;
  a.st2 = g->st2;
  a.itc = g->itc;
  a.nbr = g->nbr;
  a.mn = g->mn;
  a.mx = g->mx;
  a.limmm = g->limmm;
  a.flags = g->flags;
  a.fv_list = g->fv_list;
  a.rc_list = g->rc_list;
  a.loc_mm = g->loc_mm;
  a.dt_cell = g->dt_cell;
  a.dx = g->d_dx;
  a.st_state = g->state;
  a.cfl_pnpm = g->sd.cfl * kPNPM[g->sd.order];
  a.relax = g->relax;
  a.scaling = g->scaling;
  a.fv_godunov = g->fv_godunov;
  a.T = 2 * g->h + 1;
  a.A = g->h + 1;
  a.cure = g->cure;
  a.ncells = (int)g->ncells;
  return a;
}

template <class F>
hipError_t lstaged(exa_lgrid* g, int stage, F&& launch) {
  if (!g->profiling) return launch();
  LProf r{stage, nullptr, nullptr};
  hipError_t e = hipEventCreate(&r.b);
  if (e == hipSuccess) e = hipEventCreate(&r.e);
  if (e == hipSuccess) e = hipEventRecord(r.b, g->stream);
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(r.e, g->stream);
  g->prof.push_back(r);
  return e;
}

void lfree(exa_lgrid* g) {
  if (!g) return;
  for (auto& r : g->prof) {
    (void)hipEventDestroy(r.b);
    (void)hipEventDestroy(r.e);
  }
  void* ptrs[] = {g->luh,  g->luh_prev, g->lduh, g->lQhbnd, g->lFhbnd, g->lim,     g->mn,      g->mx,
                  g->limmm, g->d_dx,    g->cur,  g->has_patch, g->st,  g->pst,     g->st2,     g->itc,
                  g->nbr,  g->flags,    g->fv_list, g->iters, g->faces, g->state, g->ws,
                  g->rc_list, g->loc_mm, g->dt_cell};
  for (void* p : ptrs) (void)hipFree(p);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

// Prediction of the cells with status[c] < troubled (the limiter status, or
// the re-prediction mask) with the device dt; with run_if, only if *run_if.
hipError_t lpredict(exa_lgrid* g, const double* dt, const int* status, const int* run_if = nullptr) {
  exa::PredictorArgs a{};
  a.luh = g->luh;
  a.lduh = g->lduh;
  a.lQhbnd = g->lQhbnd;
  a.lFhbnd = g->lFhbnd;
  a.dx = g->d_dx;
  a.dx_stride = 0;
  a.dt = dt;
  // the gated launches skip cells without writing their iteration count
  const bool gated = run_if != nullptr;
  a.iters = gated ? nullptr : g->iters;
  a.cell0 = 0;
  a.ws = g->ws;
  a.ws_cells = g->ws_cells;
  a.status = status;  // Prediction.cpp:177: cells with status < troubled
  a.status_limit = 2 * g->h + 1;
  a.run_if = run_if;
  hipError_t e =
      lstaged(g, EXA_STAGE_PREDICTOR, [&] { return exa::launch_predictor(g->sd, false, a, g->ncells, g->stream); });
  if (e == hipSuccess && g->profiling && !gated)
    e = exa::launch_count_iterations(g->iters, g->ncells, g->state, g->stream);
  return e;
}

// First time step size and prediction of an initial state.
hipError_t lstart(exa_lgrid* g) {
  exa::StepState st0{};
  for (int i = 0; i < exa::kDtSlots; ++i) st0.dt_slots[i] = exa::kDtSlotEmpty;
  hipError_t e = hipMemcpyAsync(g->state, &st0, sizeof(st0), hipMemcpyHostToDevice, g->stream);
  if (e != hipSuccess) return e;
  exa::UpdateArgs u{};
  u.luh = g->luh;
  u.lduh = g->lduh;
  u.lFhbnd = g->lFhbnd;
  u.dx = g->d_dx;
  u.dx_stride = 0;
  u.dt = &g->state->dt;
  u.cfl_pnpm = g->sd.cfl * kPNPM[g->sd.order];
  u.dt_slots = g->state->dt_slots;
  e = exa::launch_update(g->sd, 0, u, g->ncells, g->stream);
  if (e == hipSuccess)
    e = g->fused_ts ? exa::launch_dt_finalize_fused(g->state, g->fused_weight, true, g->stream)
                    : exa::launch_dt_finalize(g->state, false, g->stream);
  if (e == hipSuccess) e = lpredict(g, &g->state->dt, g->st);
  return e;
}

// One time step of the limiting solver: non-fused, or fused (fusedTimeStep,
// LS:806-836, in the SolutionUpdate traversal; updateMeshAndSubdomains,
// Runner.cpp:866-946, after an irregular change; the re-run of the
// prediction after a failed stability check, Runner.cpp:1101-1107).
hipError_t lstep(exa_lgrid* g) {
  const exa::LimArgs a = largs(g);
  const SolverDesc& sd = g->sd;
  hipStream_t s = g->stream;
  hipError_t e = lstaged(g, EXA_STAGE_LIMITER, [&] { return exa::launch_lim_begin_step(sd, a, s); });
  // Merging: DG Riemann problems unless both cells are FV-active
  // (LS:1575-1610); transmissive DG boundary faces need no work: Rusanov with
  // equal states returns the inner flux unchanged.
  if (e == hipSuccess)
    e = lstaged(g, EXA_STAGE_RIEMANN, [&] {
      exa::RiemannArgs ra{g->faces, g->nfaces, g->lQhbnd, g->lFhbnd, g->st, a.A};
      return exa::launch_riemann(sd, ra, s);
    });
  if (e == hipSuccess)
    e = lstaged(g, EXA_STAGE_LIMITER, [&] {
      hipError_t r = exa::launch_lim_minmax_merge(sd, a, g->faces, g->nfaces, s);
      if (r == hipSuccess) r = exa::launch_lim_ghosts(sd, a, false, s);
      return r;
    });
  if (e == hipSuccess) e = lstaged(g, EXA_STAGE_FV, [&] { return exa::launch_lim_fv_update(sd, a, false, s); });
  if (e == hipSuccess) e = lstaged(g, EXA_STAGE_UPDATE, [&] { return exa::launch_lim_update_status(sd, a, s); });
  const int* irregular = g->flags + exa::kLimFlagIrregular;
  if (g->fused_ts) {
    // the cells' prediction with the predictor dt right after their update
    // and status (the status before spreading; spreading keeps the set of
    // cells below troubled), then the stability check
    if (e == hipSuccess) e = lpredict(g, &g->state->dtp, g->st);
    if (e == hipSuccess)
      e = lstaged(g, EXA_STAGE_DT, [&] {
        return exa::launch_dt_finalize_fused_limiting(g->state, g->fused_weight, irregular, 1, s);
      });
  }
  if (e == hipSuccess)
    e = lstaged(g, EXA_STAGE_LIMITER, [&] {
      hipError_t r = exa::launch_lim_spread_rollback(sd, a, s);
      if (r == hipSuccess) r = exa::launch_lim_ghosts(sd, a, true, s);
      if (r == hipSuccess) r = exa::launch_lim_fv_update(sd, a, true, s);
      if (r == hipSuccess) r = exa::launch_lim_recompute_finish(sd, a, s);
      if (r == hipSuccess) {
        hipLaunchKernelGGL(lim_count_irregular_kernel, dim3(1), dim3(64), 0, s, g->flags);
        r = hipGetLastError();
      }
      if (r == hipSuccess && g->fused_ts) r = exa::launch_lim_repredict_mask(a, g->st2, s);
      return r;
    });
  if (!g->fused_ts) {
    if (e == hipSuccess)
      e = lstaged(g, EXA_STAGE_DT, [&] { return exa::launch_dt_finalize(g->state, true, s); });
    if (e == hipSuccess) e = lpredict(g, &g->state->dt, g->st);
    return e;
  }
  // irregular change: the recomputed cells' prediction with the predictor dt
  // they were rolled back to (recomputePredictorLocally), the repeated check
  if (e == hipSuccess) e = lpredict(g, &g->state->dtp, g->st2, irregular);
  if (e == hipSuccess)
    e = lstaged(g, EXA_STAGE_DT, [&] {
      return exa::launch_dt_finalize_fused_limiting(g->state, g->fused_weight, irregular, 2, s);
    });
  // a failed check: the prediction again with the new dt
  if (e == hipSuccess) e = lpredict(g, &g->state->dt, g->st, &g->state->violated);
  return e;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int exa_lgrid_create(const exa_solver* s, int nx, int ny, int nz, const double* lo, const double* hi,
                                int boundary, int helper_layers, int steps_till_cured, double dmp_relaxation,
                                double dmp_difference_scaling, int fv_scheme, exa_lgrid** out) {
  if (!out) return EXA_E_INVALID;
  *out = nullptr;
  SolverDesc sd;
  if (!to_desc(s, &sd) || nx < 1 || ny < 1 || nz < 1 || !lo || !hi || helper_layers < 1 || steps_till_cured < 0)
    return EXA_E_INVALID;
  if (boundary != EXA_BC_PERIODIC && boundary != EXA_BC_TRANSMISSIVE) return EXA_E_INVALID;
  if (fv_scheme != EXA_FV_GODUNOV && fv_scheme != EXA_FV_MUSCLHANCOCK) return EXA_E_INVALID;
  // the limiting ADER-DG solver runs the nonlinear kernels only
  if (sd.linear || !exa::limiter_supported(sd) || !exa::supported(sd)) return EXA_E_UNSUPPORTED;
  const long ncells = (long)nx * ny * nz;
  if (ncells > 0x7fffffffL) return EXA_E_INVALID;
  EXA_TRY(exa::ensure_limiter_tables());
  exa_lgrid* g = new exa_lgrid();
  g->sd = sd;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->ncells = ncells;
  g->periodic = boundary == EXA_BC_PERIODIC;
  g->h = helper_layers;
  g->cure = steps_till_cured;
  g->relax = dmp_relaxation;
  g->scaling = dmp_difference_scaling;
  g->fv_godunov = fv_scheme == EXA_FV_GODUNOV;
  for (int d = 0; d < 3; ++d) {
    g->lo[d] = lo[d];
    g->hi[d] = hi[d];
  }
  g->dx[0] = (hi[0] - lo[0]) / nx;
  g->dx[1] = (hi[1] - lo[1]) / ny;
  g->dx[2] = (hi[2] - lo[2]) / nz;
  const Sizes z(sd);
  const long S = exa::lim_S(sd.order);
  const long patch = S * S * S * z.V;
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  A((void**)&g->luh, sizeof(double) * ncells * z.B3 * z.V);
  A((void**)&g->luh_prev, sizeof(double) * ncells * z.B3 * z.V);
  A((void**)&g->lduh, sizeof(double) * ncells * z.B3 * z.V);
  A((void**)&g->lQhbnd, sizeof(double) * ncells * 6 * z.B2 * z.V);
  A((void**)&g->lFhbnd, sizeof(double) * ncells * 6 * z.B2 * z.V);
  A((void**)&g->lim, sizeof(double) * ncells * 2 * patch);
  A((void**)&g->mn, sizeof(double) * ncells * 6 * z.V);
  A((void**)&g->mx, sizeof(double) * ncells * 6 * z.V);
  A((void**)&g->limmm, sizeof(double) * ncells * 2 * z.V);
  A((void**)&g->d_dx, sizeof(double) * 3);
  A((void**)&g->cur, ncells);
  A((void**)&g->has_patch, ncells);
  A((void**)&g->st, sizeof(int) * ncells);
  A((void**)&g->pst, sizeof(int) * ncells);
  A((void**)&g->st2, sizeof(int) * ncells);
  A((void**)&g->itc, sizeof(int) * ncells);
  A((void**)&g->nbr, sizeof(int) * ncells * 6);
  A((void**)&g->flags, sizeof(int) * 4);
  A((void**)&g->fv_list, sizeof(int) * ncells);
  A((void**)&g->rc_list, sizeof(int) * ncells);
  A((void**)&g->loc_mm, sizeof(double) * ncells * 10);
  A((void**)&g->dt_cell, sizeof(double) * ncells);
  A((void**)&g->iters, sizeof(int) * ncells);
  A((void**)&g->state, sizeof(exa::StepState));
  if (const size_t wsb = exa::predictor_ws_bytes_per_cell(sd)) {
    g->ws_cells = std::min(ncells, exa::kPredictorWsCells);
    A(&g->ws, wsb * g->ws_cells);
  }
  // patch edges and corners must read zero forever (FiniteVolumesSolver.cpp:477-481)
  if (e == hipSuccess) e = hipMemsetAsync(g->lim, 0, sizeof(double) * ncells * 2 * patch, g->stream);
  if (e == hipSuccess) e = hipMemsetAsync(g->flags, 0, sizeof(int) * 4, g->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(g->d_dx, g->dx, sizeof(double) * 3, hipMemcpyHostToDevice, g->stream);
  // neighbours and the interior face list (each face once, owned by its left cell)
  std::vector<int> nbr(ncells * 6);
  std::vector<FaceDesc> faces;
  faces.reserve(ncells * 3);
  auto cid = [&](int i, int j, int k) -> int {
    if (g->periodic) {
      i = (i + nx) % nx;
      j = (j + ny) % ny;
      k = (k + nz) % nz;
    } else if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) {
      return -1;
    }
    return (int)(((long)k * ny + j) * nx + i);
  };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int c = cid(i, j, k);
        const int nb[6] = {cid(i - 1, j, k), cid(i + 1, j, k), cid(i, j - 1, k),
                           cid(i, j + 1, k), cid(i, j, k - 1), cid(i, j, k + 1)};
        for (int f = 0; f < 6; ++f) nbr[(size_t)c * 6 + f] = nb[f];
        for (int d = 0; d < 3; ++d)
          if (nb[2 * d + 1] >= 0) faces.push_back(FaceDesc{c, nb[2 * d + 1], d, 0});
      }
  g->nfaces = (long)faces.size();
  A((void**)&g->faces, sizeof(FaceDesc) * std::max<long>(g->nfaces, 1));
  if (e == hipSuccess) e = hipMemcpy(g->nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && g->nfaces)
    e = hipMemcpy(g->faces, faces.data(), sizeof(FaceDesc) * faces.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) {
    lfree(g);
    return from_hip(e);
  }
  *out = g;
  return EXA_OK;
}

extern "C" int exa_lgrid_destroy(exa_lgrid* g) {
  if (g && g->stream) (void)hipStreamSynchronize(g->stream);
  lfree(g);
  return EXA_OK;
}

extern "C" long exa_lgrid_num_cells(const exa_lgrid* g) { return g ? g->ncells : 0; }

static int lgrid_init_from_device(exa_lgrid* g, const double* d_lim_ic) {
  const exa::LimArgs a = largs(g);
  EXA_TRY(hipMemsetAsync(g->flags, 0, sizeof(int) * 4, g->stream));  // new run: counters from zero
  EXA_TRY(exa::launch_lim_init(g->sd, a, d_lim_ic, g->stream));
  EXA_TRY(lstart(g));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_lgrid_upload(exa_lgrid* g, const double* luh, const double* lim_ic) {
  if (!g || !luh) return EXA_E_INVALID;
  const Sizes z(g->sd);
  const long L = exa::lim_L(g->sd.order);
  EXA_TRY(hipMemcpyAsync(g->luh, luh, sizeof(double) * g->ncells * z.B3 * z.V, hipMemcpyHostToDevice, g->stream));
  DevBuf ic;
  if (lim_ic) {
    EXA_TRY(ic.alloc(sizeof(double) * g->ncells * L * L * L * z.V));
    EXA_TRY(hipMemcpyAsync(ic.p, lim_ic, sizeof(double) * g->ncells * L * L * L * z.V, hipMemcpyHostToDevice,
                           g->stream));
  }
  return lgrid_init_from_device(g, lim_ic ? ic.d() : nullptr);
}

extern "C" int exa_lgrid_init_blast(exa_lgrid* g, unsigned long long seed, double radius, double p_in, double p_out,
                                    double noise) {
  if (!g) return EXA_E_INVALID;
  const int B = g->sd.order + 1, L = exa::lim_L(g->sd.order);
  DevBuf ic;
  EXA_TRY(ic.alloc(sizeof(double) * g->ncells * L * L * L * 5));
  for (int sub = 0; sub < 2; ++sub) {
    const int P = sub ? L : B;
    const long n = g->ncells * P * P * P;
    double* out = sub ? ic.d() : g->luh;
    hipLaunchKernelGGL(blast_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->stream, out, g->nx, g->ny,
                       g->nz, P, sub, g->lo[0], g->lo[1], g->lo[2], g->dx[0], g->dx[1], g->dx[2],
                       seed + (unsigned long long)sub, radius, p_in, p_out, noise, g->sd.order);
    EXA_TRY(hipGetLastError());
  }
  return lgrid_init_from_device(g, ic.d());
}

extern "C" int exa_lgrid_step(exa_lgrid* g, int nsteps) {
  if (!g || nsteps < 0) return EXA_E_INVALID;
  for (int n = 0; n < nsteps; ++n) EXA_TRY(lstep(g));
  return EXA_OK;
}

extern "C" int exa_lgrid_set_fused_time_stepping(exa_lgrid* g, int on, double weight) {
  if (!g) return EXA_E_INVALID;
  if (on && !(weight > 0.0 && weight <= 1.0)) return EXA_E_INVALID;
  g->fused_ts = on != 0;
  if (on) g->fused_weight = weight;
  // the predictor dt starts as the current dt: set before the first step,
  // this is the fused start (first dt unweighted, TimeStepSizeComputation.cpp:226-237)
  EXA_TRY(hipMemcpyAsync(&g->state->dtp, &g->state->dt, sizeof(double), hipMemcpyDeviceToDevice, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_lgrid_predictor_reruns(exa_lgrid* g, long* reruns) {
  if (!g || !reruns) return EXA_E_INVALID;
  long long r = 0;
  EXA_TRY(hipMemcpyAsync(&r, &g->state->reruns, sizeof(long long), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  *reruns = (long)r;
  return EXA_OK;
}

extern "C" int exa_lgrid_synchronize(exa_lgrid* g) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_lgrid_time(exa_lgrid* g, double* t, double* dt, long* steps) {
  if (!g) return EXA_E_INVALID;
  exa::StepState st;
  EXA_TRY(hipMemcpyAsync(&st, g->state, offsetof(exa::StepState, dt_slots), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  if (t) *t = st.t;
  if (dt) *dt = st.dt;
  if (steps) *steps = st.steps;
  return std::isfinite(st.dt) ? EXA_OK : EXA_E_NONFINITE;
}

extern "C" int exa_lgrid_download(exa_lgrid* g, double* luh) {
  if (!g || !luh) return EXA_E_INVALID;
  const Sizes z(g->sd);
  EXA_TRY(hipMemcpyAsync(luh, g->luh, sizeof(double) * g->ncells * z.B3 * z.V, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_lgrid_status(exa_lgrid* g, int* status, long* irregular_steps) {
  if (!g) return EXA_E_INVALID;
  int flags[4];
  if (status) EXA_TRY(hipMemcpyAsync(status, g->st, sizeof(int) * g->ncells, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipMemcpyAsync(flags, g->flags, sizeof(flags), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  if (irregular_steps) *irregular_steps = flags[3];
  return EXA_OK;
}

extern "C" int exa_lgrid_limiter_solution(exa_lgrid* g, long cell, double* lim) {
  if (!g || !lim || cell < 0 || cell >= g->ncells) return EXA_E_INVALID;
  const long S = exa::lim_S(g->sd.order), patch = S * S * S * 5;
  uint8_t cur = 0;
  EXA_TRY(hipMemcpyAsync(&cur, g->cur + cell, 1, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  EXA_TRY(hipMemcpyAsync(lim, g->lim + ((size_t)cell * 2 + cur) * patch, sizeof(double) * patch,
                         hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_lgrid_min_max(exa_lgrid* g, double* mn, double* mx) {
  if (!g || !mn || !mx) return EXA_E_INVALID;
  EXA_TRY(hipMemcpyAsync(mn, g->mn, sizeof(double) * g->ncells * 30, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipMemcpyAsync(mx, g->mx, sizeof(double) * g->ncells * 30, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_lgrid_set_profiling(exa_lgrid* g, int on) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  g->profiling = on != 0;
  if (on) {  // Picard totals count from here
    const unsigned long long zero[2] = {0, 0};
    EXA_TRY(hipMemcpy(&g->state->picard_total, zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  return EXA_OK;
}

extern "C" int exa_lgrid_picard_totals(exa_lgrid* g, unsigned long long* iterations, unsigned long long* cells) {
  if (!g) return EXA_E_INVALID;
  exa::StepState st;
  EXA_TRY(hipMemcpyAsync(&st, g->state, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  if (iterations) *iterations = st.picard_total;
  if (cells) *cells = st.predictor_cells;
  return EXA_OK;
}

extern "C" int exa_lgrid_stage_times(exa_lgrid* g, double* ms, long* launches) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  for (auto& r : g->prof) {
    float t = 0.f;
    EXA_TRY(hipEventElapsedTime(&t, r.b, r.e));
    g->prof_ms[r.stage] += t;
    g->prof_n[r.stage] += 1;
    (void)hipEventDestroy(r.b);
    (void)hipEventDestroy(r.e);
  }
  g->prof.clear();
  for (int i = 0; i < EXA_NUM_LSTAGES; ++i) {
    if (ms) ms[i] = g->prof_ms[i];
    if (launches) launches[i] = g->prof_n[i];
    g->prof_ms[i] = 0;
    g->prof_n[i] = 0;
  }
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// single-cell limiter kernels
// ---------------------------------------------------------------------------
static bool lim_sizes_ok(int V, int basis_size) {
  return V == 5 && basis_size >= 3 && basis_size <= exa::kLimMaxOrder + 1;
}

extern "C" int exa_project_on_fv_limiter_space(const double* luh, int num_variables, int basis_size,
                                               int ghost_layer_width, double* lim) {
  if (!luh || !lim || ghost_layer_width < 0) return EXA_E_INVALID;
  if (!lim_sizes_ok(num_variables, basis_size)) return EXA_E_UNSUPPORTED;
  EXA_TRY(exa::ensure_limiter_tables());
  const long B = basis_size, L = 2 * B - 1, S = L + 2 * ghost_layer_width, V = num_variables;
  DevBuf d_luh, d_lim;
  EXA_TRY(d_luh.alloc(sizeof(double) * B * B * B * V));
  EXA_TRY(d_lim.alloc(sizeof(double) * S * S * S * V));
  EXA_TRY(hipMemcpy(d_luh.p, luh, sizeof(double) * B * B * B * V, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_lim.p, lim, sizeof(double) * S * S * S * V, hipMemcpyHostToDevice));
  EXA_TRY(exa::launch_project_on_fv(basis_size - 1, num_variables, ghost_layer_width, d_luh.d(), d_lim.d(), 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(lim, d_lim.p, sizeof(double) * S * S * S * V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_project_on_dg_space(const double* lim, int num_variables, int basis_size, int ghost_layer_width,
                                       double* luh) {
  if (!luh || !lim || ghost_layer_width < 0) return EXA_E_INVALID;
  if (!lim_sizes_ok(num_variables, basis_size)) return EXA_E_UNSUPPORTED;
  EXA_TRY(exa::ensure_limiter_tables());
  const long B = basis_size, L = 2 * B - 1, S = L + 2 * ghost_layer_width, V = num_variables;
  DevBuf d_luh, d_lim;
  EXA_TRY(d_luh.alloc(sizeof(double) * B * B * B * V));
  EXA_TRY(d_lim.alloc(sizeof(double) * S * S * S * V));
  EXA_TRY(hipMemcpy(d_lim.p, lim, sizeof(double) * S * S * S * V, hipMemcpyHostToDevice));
  EXA_TRY(exa::launch_project_on_dg(basis_size - 1, num_variables, ghost_layer_width, d_lim.d(), d_luh.d(), 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(luh, d_luh.p, sizeof(double) * B * B * B * V, hipMemcpyDeviceToHost));
  return EXA_OK;
}

static int dmp_impl(const double* luh, int V, int basis_size, double relax, double scaling, double* mn, double* mx,
                    double* bmin, double* bmax, int* ok) {
  if (!lim_sizes_ok(V, basis_size)) return EXA_E_UNSUPPORTED;
  EXA_TRY(exa::ensure_limiter_tables());
  const long B = basis_size;
  DevBuf d_luh, d_mm, d_b, d_ok;
  EXA_TRY(d_luh.alloc(sizeof(double) * B * B * B * V));
  EXA_TRY(d_mm.alloc(sizeof(double) * 2 * V));
  EXA_TRY(d_b.alloc(sizeof(double) * 12 * V));
  EXA_TRY(d_ok.alloc(sizeof(int)));
  EXA_TRY(hipMemcpy(d_luh.p, luh, sizeof(double) * B * B * B * V, hipMemcpyHostToDevice));
  if (bmin) {
    EXA_TRY(hipMemcpy(d_b.p, bmin, sizeof(double) * 6 * V, hipMemcpyHostToDevice));
    EXA_TRY(hipMemcpy(d_b.d() + 6 * V, bmax, sizeof(double) * 6 * V, hipMemcpyHostToDevice));
  }
  EXA_TRY(exa::launch_dmp(basis_size - 1, d_luh.d(), d_mm.d(), d_mm.d() + V, bmin ? d_b.d() : nullptr,
                          bmin ? d_b.d() + 6 * V : nullptr, static_cast<int*>(d_ok.p), relax, scaling, 1, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  if (mn) EXA_TRY(hipMemcpy(mn, d_mm.p, sizeof(double) * V, hipMemcpyDeviceToHost));
  if (mx) EXA_TRY(hipMemcpy(mx, d_mm.d() + V, sizeof(double) * V, hipMemcpyDeviceToHost));
  if (bmin) {
    EXA_TRY(hipMemcpy(bmin, d_b.p, sizeof(double) * 6 * V, hipMemcpyDeviceToHost));
    EXA_TRY(hipMemcpy(bmax, d_b.d() + 6 * V, sizeof(double) * 6 * V, hipMemcpyDeviceToHost));
    EXA_TRY(hipMemcpy(ok, d_ok.p, sizeof(int), hipMemcpyDeviceToHost));
  }
  return EXA_OK;
}

extern "C" int exa_find_cell_local_min_and_max(const double* luh, int num_variables, int basis_size,
                                               double* local_min, double* local_max) {
  if (!luh || !local_min || !local_max) return EXA_E_INVALID;
  return dmp_impl(luh, num_variables, basis_size, 0, 0, local_min, local_max, nullptr, nullptr, nullptr);
}

extern "C" int exa_discrete_maximum_principle_and_min_and_max_search(const double* luh, int num_variables,
                                                                     int basis_size, double relaxation,
                                                                     double difference_scaling, double* bmin,
                                                                     double* bmax, int* dmp_ok) {
  if (!luh || !bmin || !bmax || !dmp_ok) return EXA_E_INVALID;
  return dmp_impl(luh, num_variables, basis_size, relaxation, difference_scaling, nullptr, nullptr, bmin, bmax,
                  dmp_ok);
}

extern "C" int exa_musclhancock_solution_update(const exa_solver* s, double* luh_new, const double* luh,
                                                const double* dx, double dt, int patch_size,
                                                int ghost_layer_width, double* dt_max_allowed) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || !luh_new || !luh || !dx) return EXA_E_INVALID;
  if (!exa::limiter_supported(sd)) return EXA_E_UNSUPPORTED;
  if (patch_size != exa::lim_L(sd.order) || ghost_layer_width != exa::kLimGhost) return EXA_E_UNSUPPORTED;
  EXA_TRY(exa::ensure_limiter_tables());
  const long S = exa::lim_S(sd.order), patch = S * S * S * 5, B3 = (long)(sd.order + 1) * (sd.order + 1) * (sd.order + 1);
  DevBuf d_lim, d_luh, d_mm, d_cur, d_list, d_flags, d_dx, d_state, d_dtmax;
  EXA_TRY(d_lim.alloc(sizeof(double) * 2 * patch));
  EXA_TRY(d_luh.alloc(sizeof(double) * B3 * 5));
  EXA_TRY(d_mm.alloc(sizeof(double) * 10));
  EXA_TRY(d_cur.alloc(1));
  EXA_TRY(d_list.alloc(sizeof(int)));
  EXA_TRY(d_flags.alloc(sizeof(int) * 4));
  EXA_TRY(d_dx.alloc(sizeof(double) * 3));
  EXA_TRY(d_state.alloc(sizeof(exa::StepState)));
  EXA_TRY(d_dtmax.alloc(sizeof(unsigned long long)));
  const double dmax0 = DBL_MAX;
  EXA_TRY(hipMemcpy(d_dtmax.p, &dmax0, sizeof(double), hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_lim.p, luh, sizeof(double) * patch, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_lim.d() + patch, luh_new, sizeof(double) * patch, hipMemcpyHostToDevice));
  EXA_TRY(hipMemset(d_cur.p, 0, 1));
  EXA_TRY(hipMemset(d_flags.p, 0, sizeof(int) * 4));
  EXA_TRY(hipMemcpy(d_dx.p, dx, sizeof(double) * 3, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_state.p, &dt, sizeof(double), hipMemcpyHostToDevice));  // StepState::dt is first
  exa::LimArgs a{};
  a.lim = d_lim.d();
  a.cur = static_cast<uint8_t*>(d_cur.p);
  a.luh = d_luh.d();
  a.limmm = d_mm.d();
  a.flags = static_cast<int*>(d_flags.p);
  a.fv_list = static_cast<int*>(d_list.p);
  a.dx = d_dx.d();
  a.st_state = static_cast<exa::StepState*>(d_state.p);
  a.ncells = 1;
  a.fv_dt_max = static_cast<unsigned long long*>(d_dtmax.p);
  EXA_TRY(exa::launch_lim_iota(a.fv_list, 1, a.flags, nullptr));
  EXA_TRY(exa::launch_lim_fv_update(sd, a, false, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(luh_new, d_lim.d() + patch, sizeof(double) * patch, hipMemcpyDeviceToHost));
  if (dt_max_allowed) EXA_TRY(hipMemcpy(dt_max_allowed, d_dtmax.p, sizeof(double), hipMemcpyDeviceToHost));
  return EXA_OK;
}
