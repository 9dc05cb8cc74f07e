// fv_capi.hip -- C-ABI of the finite-volume kernels (include/exahype_amd.h):
// the single-patch Godunov update and the device-resident finite-volumes
// grid (FiniteVolumesSolver's time step on one uniform level): Godunov on
// dense subcell planes, MUSCL-Hancock on per-cell patches with ghost layers
// (the limiter's FV kernels, every cell FV-active).
//   GOD = ExaHyPE/kernels/finitevolumes/godunov/c/3d/godunov.cpph
//   FVS = ExaHyPE/exahype/solvers/FiniteVolumesSolver.cpp
#include <hip/hip_runtime.h>

#include <cfloat>
#include <utility>
#include <vector>

#include "capi_util.hpp"
#include "fv.hpp"
#include "limiter.hpp"

namespace {

bool euler_fv_solver(const SolverDesc& sd) {
  return sd.pde == EXA_PDE_EULER && sd.nvar == 5 && sd.npar == 0 && sd.cfl > 0.0;
}

}  // namespace

extern "C" int exa_godunov_solution_update(const exa_solver* s, double* luh_new, const double* luh,
                                           const double* dx, double dt, int patch_size, int ghost_layer_width,
                                           double* dt_max_allowed) {
  SolverDesc sd;
  if (!to_desc(s, &sd) || !luh_new || !luh || !dx) return EXA_E_INVALID;
  if (patch_size < 1 || ghost_layer_width < 1 || patch_size > 512 || ghost_layer_width > 8) return EXA_E_INVALID;
  if (!euler_fv_solver(sd)) return EXA_E_UNSUPPORTED;
  const long S = patch_size + 2L * ghost_layer_width;
  const size_t n = (size_t)S * S * S * 5;
  DevBuf d_old, d_new, d_dtmax;
  EXA_TRY(d_old.alloc(sizeof(double) * n));
  EXA_TRY(d_new.alloc(sizeof(double) * n));
  EXA_TRY(d_dtmax.alloc(sizeof(unsigned long long)));
  const double dmax0 = DBL_MAX;
  EXA_TRY(hipMemcpy(d_dtmax.p, &dmax0, sizeof(double), hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_old.p, luh, sizeof(double) * n, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_new.p, luh_new, sizeof(double) * n, hipMemcpyHostToDevice));
  exa::GodunovPatchArgs a{};
  a.luh = d_old.d();
  a.luh_new = d_new.d();
  a.L = patch_size;
  a.g = ghost_layer_width;
  for (int d = 0; d < 3; ++d) a.h[d] = dx[d] / patch_size;  // cellSize, GOD:51
  a.dt = dt;
  a.cfl = sd.cfl;
  a.dt_max = static_cast<unsigned long long*>(d_dtmax.p);
  EXA_TRY(exa::launch_godunov_patch(a, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  EXA_TRY(hipMemcpy(luh_new, d_new.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (dt_max_allowed) EXA_TRY(hipMemcpy(dt_max_allowed, d_dtmax.p, sizeof(double), hipMemcpyDeviceToHost));
  return EXA_OK;
}

struct exa_fvgrid {
  SolverDesc sd;
  int scheme = EXA_FV_GODUNOV;
  int L = 0, nx = 0, ny = 0, nz = 0, X = 0, Y = 0, Z = 0;
  int periodic = 1;
  double lo[3] = {0, 0, 0}, dx[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  size_t vol = 0;
  hipStream_t stream = nullptr;
  double* u = nullptr;   // current solution planes [5][Z][Y][X]
  double* un = nullptr;  // next solution planes (also the upload / download staging buffer)
  exa::FvState* state = nullptr;
  exa::FvGridArgs args{};
  bool initialised = false;
  bool profiling = false;
  // MUSCL-Hancock: per-cell patches [cells][2][S^3][5] with kLimGhost ghost
  // layers, stepped by the limiter's ghost and FV kernels with every cell in
  // the FV list; luh_dg receives their (unused) projection to DG of order N,
  // L = 2N + 1
  SolverDesc lim_sd{};
  exa::LimArgs la{};
  double *lim = nullptr, *luh_dg = nullptr, *limmm = nullptr, *d_dx = nullptr;
  uint8_t* cur = nullptr;
  int *nbr = nullptr, *fv_list = nullptr, *flags = nullptr;
  exa::StepState* step = nullptr;
  std::vector<hipEvent_t> ev;  // per profiled step: before update, after update, after finalize
  double prof_ms[EXA_FV_NUM_STAGES] = {0, 0};
  long prof_n[EXA_FV_NUM_STAGES] = {0, 0};
};

namespace {

void fv_free(exa_fvgrid* g) {
  if (!g) return;
  for (hipEvent_t e : g->ev) (void)hipEventDestroy(e);
  if (g->u) (void)hipFree(g->u);
  if (g->un) (void)hipFree(g->un);
  if (g->state) (void)hipFree(g->state);
  void* muscl[] = {g->lim, g->luh_dg, g->limmm, g->d_dx, g->cur, g->nbr, g->fv_list, g->flags, g->step};
  for (void* p : muscl)
    if (p) (void)hipFree(p);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

// Tiling of the update kernel: 62 x 6 columns per workgroup, the z range cut
// into chunks so that the launch has about 8 workgroups per CU in total,
// chunks of at least 16 planes.
void fv_tiling(exa_fvgrid* g) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  exa::FvGridArgs& a = g->args;
  constexpr int out_rows = exa::kFvRows - 2;
  a.nsx = (g->X + exa::kFvOutX - 1) / exa::kFvOutX;
  a.nry = (g->Y + out_rows - 1) / out_rows;
  const long per_chunk = (long)a.nsx * a.nry;
  long nzc = (8L * cus + per_chunk - 1) / per_chunk;
  const long max_chunks = g->Z / 16 > 1 ? g->Z / 16 : 1;
  if (nzc > max_chunks) nzc = max_chunks;
  if (nzc < 1) nzc = 1;
  a.zlen = (int)((g->Z + nzc - 1) / nzc);
  a.nzc = (g->Z + a.zlen - 1) / a.zlen;
  a.ntiles = (int)(per_chunk * a.nzc);
}

int fv_initial_dt(exa_fvgrid* g) {
  if (g->scheme == EXA_FV_MUSCLHANCOCK) {
    EXA_TRY(exa::launch_fv_patch_dt(g->lim, g->cur, (long)g->nx * g->ny * g->nz, g->L, exa::kLimGhost, g->state, true,
                                    g->stream));
  } else {
    exa::FvGridArgs a = g->args;
    a.u = g->u;
    a.un = g->un;
    EXA_TRY(exa::launch_fv_initial_dt(a, g->stream));
  }
  g->initialised = true;
  return EXA_OK;
}

// The MUSCL-Hancock grid's patch state (zeroed patches: the edge and corner
// ghosts stay zero, as the reference's freshly allocated patches, FVS:477-481).
int muscl_setup(exa_fvgrid* g) {
  const long nc = (long)g->nx * g->ny * g->nz;
  const int L = g->L, S = L + 2 * exa::kLimGhost, N = (L - 1) / 2, B = N + 1;
  const size_t patch = (size_t)S * S * S * 5;
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 8);
  };
  A((void**)&g->lim, sizeof(double) * patch * 2 * nc);
  A((void**)&g->luh_dg, sizeof(double) * (size_t)B * B * B * 5 * nc);
  A((void**)&g->limmm, sizeof(double) * 10 * nc);
  A((void**)&g->d_dx, sizeof(double) * 3);
  A((void**)&g->cur, nc);
  A((void**)&g->nbr, sizeof(int) * 6 * nc);
  A((void**)&g->fv_list, sizeof(int) * nc);
  A((void**)&g->flags, sizeof(int) * 3);
  A((void**)&g->step, sizeof(exa::StepState));
  if (e != hipSuccess) return from_hip(e);
  EXA_TRY(hipMemsetAsync(g->lim, 0, sizeof(double) * patch * 2 * nc, g->stream));
  EXA_TRY(hipMemsetAsync(g->cur, 0, nc, g->stream));
  EXA_TRY(hipMemsetAsync(g->step, 0, sizeof(exa::StepState), g->stream));
  std::vector<int> nbr(6 * nc), list(nc);
  for (long c = 0; c < nc; ++c) {
    list[c] = (int)c;
    const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
    for (int f = 0; f < 6; ++f) {
      int ii = i, jj = j, kk = k;
      const int s = (f & 1) ? 1 : -1;
      if (f < 2) ii += s; else if (f < 4) jj += s; else kk += s;
      const bool out = ii < 0 || ii >= g->nx || jj < 0 || jj >= g->ny || kk < 0 || kk >= g->nz;
      if (out && !g->periodic) {
        nbr[6 * c + f] = -1;
        continue;
      }
      ii = (ii + g->nx) % g->nx;
      jj = (jj + g->ny) % g->ny;
      kk = (kk + g->nz) % g->nz;
      nbr[6 * c + f] = (int)(((long)kk * g->ny + jj) * g->nx + ii);
    }
  }
  const int flags[3] = {0, (int)nc, 0};
  EXA_TRY(hipMemcpyAsync(g->nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice, g->stream));
  EXA_TRY(hipMemcpyAsync(g->fv_list, list.data(), sizeof(int) * nc, hipMemcpyHostToDevice, g->stream));
  EXA_TRY(hipMemcpyAsync(g->flags, flags, sizeof(flags), hipMemcpyHostToDevice, g->stream));
  EXA_TRY(hipMemcpyAsync(g->d_dx, g->dx, sizeof(double) * 3, hipMemcpyHostToDevice, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  g->lim_sd = g->sd;
  g->lim_sd.order = N;
  g->lim_sd.terms = exa::kFlux;
  g->lim_sd.precision = exa::kFP64;
  exa::LimArgs& a = g->la;
  a.luh = g->luh_dg;
  a.lim = g->lim;
  a.cur = g->cur;
  a.nbr = g->nbr;
  a.flags = g->flags;
  a.fv_list = g->fv_list;
  a.limmm = g->limmm;
  a.dx = g->d_dx;
  a.st_state = g->step;
  a.ncells = (int)nc;
  a.fv_godunov = 0;
  a.fv_dt_max = nullptr;
  return from_hip(exa::ensure_limiter_tables());
}

// One MUSCL-Hancock step of every patch (FVS:662-846): ghost layers from the
// face neighbours or the transmissive boundary, the patch update with the
// grid's dt, the stable dt of the new interiors.
hipError_t muscl_step(exa_fvgrid* g) {
  hipError_t e = exa::launch_fv_dt_to_step(g->state, g->step, g->stream);
  if (e == hipSuccess) e = exa::launch_lim_ghosts(g->lim_sd, g->la, false, g->stream);
  if (e == hipSuccess) e = exa::launch_lim_fv_update(g->lim_sd, g->la, false, g->stream);
  return e;
}

}  // namespace

extern "C" int exa_fvgrid_create(const exa_solver* s, int scheme, int patch_size, int nx, int ny, int nz,
                                 const double* lo, const double* hi, int boundary, exa_fvgrid** out) {
  if (!out) return EXA_E_INVALID;
  *out = nullptr;
  SolverDesc sd;
  if (!to_desc(s, &sd) || !lo || !hi || patch_size < 1 || nx < 1 || ny < 1 || nz < 1) return EXA_E_INVALID;
  if (boundary != EXA_BC_PERIODIC && boundary != EXA_BC_TRANSMISSIVE) return EXA_E_INVALID;
  if (scheme != EXA_FV_GODUNOV && scheme != EXA_FV_MUSCLHANCOCK) return EXA_E_INVALID;
  if (!euler_fv_solver(sd)) return EXA_E_UNSUPPORTED;
  // MUSCL-Hancock: the limiter's patch kernels, L = 2N + 1 for N = 2, 3
  if (scheme == EXA_FV_MUSCLHANCOCK && patch_size != 5 && patch_size != 7) return EXA_E_UNSUPPORTED;
  if (scheme == EXA_FV_MUSCLHANCOCK && (long)nx * ny * nz > (1L << 30)) return EXA_E_INVALID;
  const long X = (long)nx * patch_size, Y = (long)ny * patch_size, Z = (long)nz * patch_size;
  if (X > (1L << 20) || Y > (1L << 20) || Z > (1L << 20)) return EXA_E_INVALID;
  exa_fvgrid* g = new exa_fvgrid();
  g->sd = sd;
  g->scheme = scheme;
  g->L = patch_size;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->X = (int)X;
  g->Y = (int)Y;
  g->Z = (int)Z;
  g->periodic = boundary == EXA_BC_PERIODIC;
  const int n[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    g->lo[d] = lo[d];
    g->dx[d] = (hi[d] - lo[d]) / n[d];
    g->h[d] = g->dx[d] / patch_size;  // cellSize = dx / patchSize (GOD:51, FVC:409-460)
  }
  g->vol = (size_t)X * Y * Z;
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&g->u, sizeof(double) * 5 * g->vol);
  if (e == hipSuccess) e = hipMalloc((void**)&g->un, sizeof(double) * 5 * g->vol);
  if (e == hipSuccess) e = hipMalloc((void**)&g->state, sizeof(exa::FvState));
  exa::FvState st0{};
  for (int d = 0; d < 3; ++d) st0.ch[d] = sd.cfl / 3 * g->h[d];  // FVC:409-460's cfl/3 * h_d
  if (e == hipSuccess) e = hipMemcpyAsync(g->state, &st0, sizeof(st0), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) {
    fv_free(g);
    return from_hip(e);
  }
  if (scheme == EXA_FV_MUSCLHANCOCK) {
    const int r = muscl_setup(g);
    if (r != EXA_OK) {
      fv_free(g);
      return r;
    }
  }
  exa::FvGridArgs& a = g->args;
  a.X = g->X;
  a.Y = g->Y;
  a.Z = g->Z;
  a.periodic = g->periodic;
  for (int d = 0; d < 3; ++d) a.h[d] = g->h[d];
  a.state = g->state;
  fv_tiling(g);
  *out = g;
  return EXA_OK;
}

extern "C" int exa_fvgrid_destroy(exa_fvgrid* g) {
  if (!g) return EXA_E_INVALID;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  fv_free(g);
  return EXA_OK;
}

extern "C" long exa_fvgrid_num_cells(const exa_fvgrid* g) { return g ? (long)g->nx * g->ny * g->nz : 0; }

extern "C" int exa_fvgrid_upload(exa_fvgrid* g, const double* interior) {
  if (!g || !interior) return EXA_E_INVALID;
  EXA_TRY(hipMemcpyAsync(g->un, interior, sizeof(double) * 5 * g->vol, hipMemcpyHostToDevice, g->stream));
  if (g->scheme == EXA_FV_MUSCLHANCOCK)
    EXA_TRY(exa::launch_fv_patch_io(g->un, g->lim, g->cur, (long)g->nx * g->ny * g->nz, g->L, exa::kLimGhost, true,
                                    g->stream));
  else
    EXA_TRY(exa::launch_fv_pack(g->un, g->u, g->nx, g->ny, g->nz, g->L, g->stream));
  const int r = fv_initial_dt(g);
  if (r != EXA_OK) return r;
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_fvgrid_download(exa_fvgrid* g, double* interior) {
  if (!g || !interior) return EXA_E_INVALID;
  if (g->scheme == EXA_FV_MUSCLHANCOCK)
    EXA_TRY(exa::launch_fv_patch_io(g->un, g->lim, g->cur, (long)g->nx * g->ny * g->nz, g->L, exa::kLimGhost, false,
                                    g->stream));
  else
    EXA_TRY(exa::launch_fv_unpack(g->u, g->un, g->nx, g->ny, g->nz, g->L, g->stream));
  EXA_TRY(hipMemcpyAsync(interior, g->un, sizeof(double) * 5 * g->vol, hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_fvgrid_init_euler(exa_fvgrid* g, const exa_euler_ic* ic) {
  if (!g || !ic || ic->nmodes < 0 || ic->nmodes > exa::kMaxModes) return EXA_E_INVALID;
  exa::ICArgs a{};
  for (int d = 0; d < 3; ++d) a.lo[d] = g->lo[d];
  fill_euler_ic(ic, &a);
  EXA_TRY(exa::launch_fv_euler_ic(a, g->u, g->X, g->Y, g->Z, g->h, g->stream));
  if (g->scheme == EXA_FV_MUSCLHANCOCK) {  // planes -> cell-major -> patch interiors
    EXA_TRY(exa::launch_fv_unpack(g->u, g->un, g->nx, g->ny, g->nz, g->L, g->stream));
    EXA_TRY(exa::launch_fv_patch_io(g->un, g->lim, g->cur, (long)g->nx * g->ny * g->nz, g->L, exa::kLimGhost, true,
                                    g->stream));
  }
  const int r = fv_initial_dt(g);
  if (r != EXA_OK) return r;
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_fvgrid_step(exa_fvgrid* g, int nsteps) {
  if (!g || nsteps < 0) return EXA_E_INVALID;
  if (!g->initialised) return EXA_E_INVALID;
  for (int n = 0; n < nsteps; ++n) {
    exa::FvGridArgs a = g->args;
    a.u = g->u;
    a.un = g->un;
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    if (g->profiling) {
      for (int i = 0; i < 3; ++i) {
        EXA_TRY(hipEventCreate(&e[i]));
        g->ev.push_back(e[i]);
      }
      EXA_TRY(hipEventRecord(e[0], g->stream));
    }
    if (g->scheme == EXA_FV_MUSCLHANCOCK) {
      EXA_TRY(muscl_step(g));
      if (g->profiling) EXA_TRY(hipEventRecord(e[1], g->stream));
      EXA_TRY(exa::launch_fv_patch_dt(g->lim, g->cur, (long)g->nx * g->ny * g->nz, g->L, exa::kLimGhost, g->state,
                                      false, g->stream));
      if (g->profiling) EXA_TRY(hipEventRecord(e[2], g->stream));
      continue;
    }
    EXA_TRY(exa::launch_fv_godunov_step(a, g->stream));
    if (g->profiling) EXA_TRY(hipEventRecord(e[1], g->stream));
    EXA_TRY(exa::launch_fv_finalize(g->state, g->stream));
    if (g->profiling) EXA_TRY(hipEventRecord(e[2], g->stream));
    std::swap(g->u, g->un);
  }
  return EXA_OK;
}

extern "C" int exa_fvgrid_synchronize(exa_fvgrid* g) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_fvgrid_time(exa_fvgrid* g, double* t, double* dt, long* steps) {
  if (!g) return EXA_E_INVALID;
  exa::FvState st;
  EXA_TRY(hipMemcpyAsync(&st, g->state, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_TRY(hipStreamSynchronize(g->stream));
  if (t) *t = st.t;
  if (dt) *dt = st.dt;
  if (steps) *steps = (long)st.steps;
  return st.bad ? EXA_E_NONFINITE : EXA_OK;
}

extern "C" int exa_fvgrid_set_profiling(exa_fvgrid* g, int on) {
  if (!g) return EXA_E_INVALID;
  g->profiling = on != 0;
  return EXA_OK;
}

extern "C" int exa_fvgrid_stage_times(exa_fvgrid* g, double* ms, long* launches) {
  if (!g) return EXA_E_INVALID;
  EXA_TRY(hipStreamSynchronize(g->stream));
  for (size_t i = 0; i + 2 < g->ev.size(); i += 3) {
    float a = 0.f, b = 0.f;
    EXA_TRY(hipEventElapsedTime(&a, g->ev[i], g->ev[i + 1]));
    EXA_TRY(hipEventElapsedTime(&b, g->ev[i + 1], g->ev[i + 2]));
    g->prof_ms[EXA_FV_STAGE_UPDATE] += a;
    g->prof_ms[EXA_FV_STAGE_FINALIZE] += b;
    g->prof_n[EXA_FV_STAGE_UPDATE] += 1;
    g->prof_n[EXA_FV_STAGE_FINALIZE] += 1;
  }
  for (hipEvent_t e : g->ev) (void)hipEventDestroy(e);
  g->ev.clear();
  for (int s = 0; s < EXA_FV_NUM_STAGES; ++s) {
    if (ms) ms[s] = g->prof_ms[s];
    if (launches) launches[s] = g->prof_n[s];
    g->prof_ms[s] = 0.0;
    g->prof_n[s] = 0;
  }
  return EXA_OK;
}
