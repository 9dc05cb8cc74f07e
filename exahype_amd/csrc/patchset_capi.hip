// patchset_capi.hip -- many independent periodic patches of uneven size,
// order and PDE advanced together (the batching workload).  Patches of one
// PDE and order form a group: their cells are concatenated, so every stage is
// one launch per group over all of its patches, with per-patch dt, dx and t
// through a cell -> patch map.  Patches that do not fit the HBM budget
// together run in waves (patches are independent, so this is exact).  The
// time step of a wave is captured once in a hipGraph and replayed; an
// autotuner picks, per group, the predictor variant (matrix-core or
// vector-ALU contractions, LDS budget) by timing each on the wave's first
// state.  The corrector is the one-kernel form (each cell solves its six
// faces' Riemann problems inside its patch) or the two passes (face list,
// then update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "capi_util.hpp"

namespace {

constexpr int kSlotsPerPatch = 16;

constexpr int kPdes = 2;  // EXA_PDE_EULER, EXA_PDE_ELASTIC

struct Group {
  SolverDesc sd{};
  int pde = 0;
  int order = 0;
  std::vector<int> patches;  // indices into the wave's patch list
  long ncells = 0;
  double* luh = nullptr;
  void *lduh = nullptr, *lQhbnd = nullptr, *lFhbnd = nullptr;
  int* cell_patch = nullptr;
  int* iters = nullptr;
  FaceDesc* faces = nullptr;
  long nfaces = 0;
  void* ws = nullptr;
  long ws_cells = 0;
  int variant = exa::kPredictorAuto;
  exa::StepState* counters = nullptr;  // Picard totals of the group (profiling)
};

size_t bytes_per_cell(const SolverDesc& sd) {
  const Sizes z(sd);
  const size_t bulk = sd.precision == exa::kFP32Bulk ? 4 : 8;
  return 8 * z.B3 * z.D + bulk * (z.B3 * z.V + 6 * z.B2 * (z.D + z.V)) + sizeof(int) * 2 + sizeof(FaceDesc) * 3;
}

}  // namespace

struct exa_patchset {
  std::vector<exa_patch_spec> patches;
  int precision = 0;
  double budget = 0;
  long keep_max_cells = 0;
  hipStream_t stream = nullptr;
  std::vector<std::vector<int>> waves;
  std::vector<std::vector<double>> solutions;  // per patch with at most keep_max_cells cells
  std::vector<double> t, dt;                   // per patch, after the run
  double seconds = 0, dof_updates = 0;
  int variant_choice[kPdes][10];
  bool fused_corrector = true;
  bool profiling = false;
  // per group (pde, order), summed over waves and timed steps (profiling)
  double pred_ms[kPdes][10] = {}, corr_ms[kPdes][10] = {};
  long launches[kPdes][10] = {};
  unsigned long long picard[kPdes][10] = {}, pred_cells[kPdes][10] = {};
  exa_patchset() {
    for (auto& row : variant_choice)
      for (int& v : row) v = -1;
  }
};

namespace {

SolverDesc patch_desc(int pde, int order, int precision) {
  SolverDesc sd{};
  sd.pde = pde == EXA_PDE_ELASTIC ? exa::kPdeElastic : exa::kPdeEuler;
  sd.order = order;
  sd.nvar = pde == EXA_PDE_ELASTIC ? 9 : 5;  // elasticity: 9 variables, 3 material parameters
  sd.npar = pde == EXA_PDE_ELASTIC ? 3 : 0;
  sd.terms = pde == EXA_PDE_ELASTIC ? exa::kNCP : exa::kFlux;
  sd.no_time_averaging = 0;
  sd.cfl = 0.9;
  sd.precision = precision;
  return sd;
}

bool patch_supported(const exa_patch_spec& p) {
  if (p.n < 1) return false;
  if (p.pde == EXA_PDE_EULER) return p.order >= 2 && p.order <= 7;
  if (p.pde == EXA_PDE_ELASTIC) return p.order >= 2 && p.order <= 5 && p.width > 0;
  return false;
}

struct Wave {
  std::vector<int> ids;  // global patch ids
  std::vector<Group> groups;
  double* dt = nullptr;
  double* t = nullptr;
  double* dx = nullptr;
  unsigned long long* slots = nullptr;
  long* patch_first = nullptr;  // per wave patch: its first cell in its group's arrays
  int* patch_n = nullptr;       // per wave patch: cells per side
};

void free_wave(Wave& w) {
  for (auto& g : w.groups) {
    void* ptrs[] = {g.luh, g.lduh, g.lQhbnd, g.lFhbnd, g.cell_patch, g.iters, g.faces, g.ws, g.counters};
    for (void* p : ptrs) (void)hipFree(p);
  }
  (void)hipFree(w.dt);
  (void)hipFree(w.t);
  (void)hipFree(w.dx);
  (void)hipFree(w.slots);
  (void)hipFree(w.patch_first);
  (void)hipFree(w.patch_n);
  w.groups.clear();
}

exa::PredictorArgs pred_args(const Wave& w, const Group& g) {
  exa::PredictorArgs a{};
  a.luh = g.luh;
  a.lduh = g.lduh;
  a.lQhbnd = g.lQhbnd;
  a.lFhbnd = g.lFhbnd;
  a.dx = w.dx;
  a.dt = w.dt;
  a.iters = g.iters;
  a.ws = g.ws;
  a.ws_cells = g.ws_cells;
  a.cell_patch = g.cell_patch;
  a.variant = g.variant;
  return a;
}

exa::UpdateArgs upd_args(const Wave& w, const Group& g) {
  exa::UpdateArgs u{};
  u.luh = g.luh;
  u.lduh = g.lduh;
  u.lFhbnd = g.lFhbnd;
  u.dx = w.dx;
  u.dt = w.dt;
  u.cfl_pnpm = 0.9 * kPNPM[g.order];
  u.dt_slots = w.slots;
  u.cell_patch = g.cell_patch;
  u.slots_per_patch = kSlotsPerPatch;
  return u;
}

exa::CorrectorArgs corr_args(const Wave& w, const Group& g) {
  exa::CorrectorArgs c{};
  c.luh = g.luh;
  c.lduh = g.lduh;
  c.lQhbnd = g.lQhbnd;
  c.lFhbnd = g.lFhbnd;
  c.dx = w.dx;
  c.dt = w.dt;
  c.t = w.t;
  c.cfl_pnpm = 0.9 * kPNPM[g.order];
  c.dt_slots = w.slots;
  c.nx = (int)g.ncells;
  c.ny = c.nz = 1;
  c.boundary = exa::kBoundaryPeriodic;
  c.cell_patch = g.cell_patch;
  c.patch_first = w.patch_first;
  c.patch_n = w.patch_n;
  c.slots_per_patch = kSlotsPerPatch;
  return c;
}

// Per-launch timing of one step (profiling): events around every launch.
struct StepTimer {
  exa_patchset* b = nullptr;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<std::pair<int, int>> where;  // (kind 0 predictor / 1 corrector, group index)
  template <class F>
  hipError_t run(hipStream_t s, int kind, int gi, F&& launch) {
    if (!b) return launch();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    ev.emplace_back(e0, e1);
    where.emplace_back(kind, gi);
    return e;
  }
};

hipError_t wave_step(const Wave& w, bool fused_corrector, hipStream_t s, StepTimer* timer = nullptr) {
  StepTimer none;
  StepTimer& tm = timer ? *timer : none;
  hipError_t e = hipSuccess;
  for (size_t gi = 0; gi < w.groups.size(); ++gi) {
    const auto& g = w.groups[gi];
    if (e != hipSuccess) break;
    if (fused_corrector) {
      const exa::CorrectorArgs c = corr_args(w, g);
      e = tm.run(s, 1, (int)gi, [&] { return exa::launch_corrector(g.sd, c, s); });
    } else {
      e = tm.run(s, 1, (int)gi, [&] {
        exa::RiemannArgs ra{g.faces, g.nfaces, g.lQhbnd, g.lFhbnd, nullptr, 0};
        hipError_t r = exa::launch_riemann(g.sd, ra, s);
        if (r == hipSuccess) r = exa::launch_update(g.sd, exa::kUpdSurface | exa::kUpdUpdate, upd_args(w, g), g.ncells, s);
        return r;
      });
    }
  }
  if (e == hipSuccess)
    e = exa::launch_patch_dt_finalize(w.dt, w.t, w.slots, (int)w.ids.size(), kSlotsPerPatch, true, s);
  for (size_t gi = 0; gi < w.groups.size(); ++gi) {
    const auto& g = w.groups[gi];
    if (e != hipSuccess) break;
    e = tm.run(s, 0, (int)gi, [&] { return exa::launch_predictor(g.sd, false, pred_args(w, g), g.ncells, s); });
    if (e == hipSuccess && timer) e = exa::launch_count_iterations(g.iters, g.ncells, g.counters, s);
  }
  return e;
}

int build_wave(exa_patchset* b, const std::vector<int>& ids, Wave& w) {
  w.ids = ids;
  const int np = (int)ids.size();
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 8);
  };
  A((void**)&w.dt, sizeof(double) * np);
  A((void**)&w.t, sizeof(double) * np);
  A((void**)&w.dx, sizeof(double) * 3 * np);
  A((void**)&w.slots, sizeof(unsigned long long) * np * kSlotsPerPatch);
  A((void**)&w.patch_first, sizeof(long) * np);
  A((void**)&w.patch_n, sizeof(int) * np);
  if (e != hipSuccess) return from_hip(e);
  std::vector<long> hfirst(np, 0);
  std::vector<int> hn(np);
  for (int i = 0; i < np; ++i) hn[i] = b->patches[ids[i]].n;
  std::vector<double> hdx(3 * np), zeros(np, 0.0);
  std::vector<unsigned long long> hslots((size_t)np * kSlotsPerPatch, exa::kDtSlotEmpty);
  for (int i = 0; i < np; ++i)
    for (int d = 0; d < 3; ++d) hdx[3 * i + d] = 1.0 / b->patches[ids[i]].n;
  EXA_TRY(hipMemcpy(w.dx, hdx.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(w.t, zeros.data(), sizeof(double) * np, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(w.dt, zeros.data(), sizeof(double) * np, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(w.slots, hslots.data(), sizeof(unsigned long long) * hslots.size(), hipMemcpyHostToDevice));
  // groups of one PDE and order
  for (int pde = 0; pde < kPdes; ++pde)
  for (int order = 2; order <= 7; ++order) {
    Group g;
    g.pde = pde;
    g.order = order;
    g.sd = patch_desc(pde, order, b->precision);
    for (int i = 0; i < np; ++i)
      if (b->patches[ids[i]].order == order && b->patches[ids[i]].pde == pde) g.patches.push_back(i);
    if (g.patches.empty()) continue;
    for (int i : g.patches) g.ncells += (long)b->patches[ids[i]].n * b->patches[ids[i]].n * b->patches[ids[i]].n;
    const Sizes z(g.sd);
    const size_t bulk = b->precision == exa::kFP32Bulk ? 4 : 8;
    A((void**)&g.luh, sizeof(double) * g.ncells * z.B3 * z.D);
    A(&g.lduh, bulk * g.ncells * z.B3 * z.V);
    A(&g.lQhbnd, bulk * g.ncells * 6 * z.B2 * z.D);
    A(&g.lFhbnd, bulk * g.ncells * 6 * z.B2 * z.V);
    A((void**)&g.cell_patch, sizeof(int) * g.ncells);
    A((void**)&g.iters, sizeof(int) * g.ncells);
    g.nfaces = 3 * g.ncells;
    A((void**)&g.faces, sizeof(FaceDesc) * g.nfaces);
    A((void**)&g.counters, sizeof(exa::StepState));
    if (e == hipSuccess) e = hipMemsetAsync(g.counters, 0, sizeof(exa::StepState), b->stream);
    const size_t wsb = std::max(exa::predictor_ws_bytes_per_cell(g.sd, 0), exa::predictor_ws_bytes_per_cell(g.sd, 1));
    if (wsb) {
      g.ws_cells = std::min(g.ncells, 32768L);
      A(&g.ws, wsb * g.ws_cells);
    }
    if (e != hipSuccess) {
      w.groups.push_back(g);
      return from_hip(e);
    }
    std::vector<int> cp(g.ncells);
    std::vector<FaceDesc> faces(g.nfaces);
    long base = 0, f = 0;
    for (int i : g.patches) {
      const int n = b->patches[ids[i]].n;
      const long nc = (long)n * n * n;
      hfirst[i] = base;
      for (long c = 0; c < nc; ++c) cp[base + c] = i;
      auto cid = [&](int x, int y, int zz) {
        x = (x + n) % n;
        y = (y + n) % n;
        zz = (zz + n) % n;
        return (int)(base + ((long)zz * n + y) * n + x);
      };
      for (int dir = 0; dir < 3; ++dir)
        for (int k = 0; k < n; ++k)
          for (int j = 0; j < n; ++j)
            for (int x = 0; x < n; ++x)
              faces[f++] = FaceDesc{cid(x, j, k), dir == 0 ? cid(x + 1, j, k) : (dir == 1 ? cid(x, j + 1, k) : cid(x, j, k + 1)),
                                    dir, 0};
      const exa_patch_spec& ps = b->patches[ids[i]];
      if (pde == EXA_PDE_EULER) {  // entropy wave
        exa::ICArgs ic{};
        ic.luh = g.luh + (size_t)base * z.B3 * z.D;
        ic.ncells = nc;
        ic.nx = ic.ny = ic.nz = n;
        ic.order = order;
        ic.nvar = 5;
        ic.npar = 0;
        for (int d = 0; d < 3; ++d) {
          ic.lo[d] = 0.0;
          ic.dx[d] = 1.0 / n;
          ic.k[d] = ps.k[d];
          ic.vel[d] = 1.0;
        }
        ic.rho0 = 1.0;
        ic.amp = ps.amp;
        ic.p = 1.0;
        EXA_TRY(exa::launch_euler_ic(ic, b->stream));
      } else {  // seeded material and Gaussian velocity pulse, as exa_grid_init_elastic
        exa::ElasticICArgs ic{};
        ic.luh = g.luh + (size_t)base * z.B3 * z.D;
        ic.ncells = nc;
        ic.nx = ic.ny = ic.nz = n;
        ic.order = order;
        for (int d = 0; d < 3; ++d) {
          ic.lo[d] = 0.0;
          ic.dx[d] = 1.0 / n;
        }
        for (int c = 0; c < 6; ++c) {
          uint64_t st = ps.seed + (uint64_t)c * 0x9E3779B97F4A7C15ull;
          ic.r[c] = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
        }
        ic.width = ps.width;
        ic.amp = ps.amp;
        EXA_TRY(exa::launch_elastic_ic(ic, b->stream));
      }
      base += nc;
    }
    EXA_TRY(hipMemcpyAsync(g.cell_patch, cp.data(), sizeof(int) * g.ncells, hipMemcpyHostToDevice, b->stream));
    EXA_TRY(hipMemcpyAsync(g.faces, faces.data(), sizeof(FaceDesc) * g.nfaces, hipMemcpyHostToDevice, b->stream));
    EXA_TRY(hipStreamSynchronize(b->stream));
    w.groups.push_back(g);
  }
  EXA_TRY(hipMemcpy(w.patch_first, hfirst.data(), sizeof(long) * np, hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(w.patch_n, hn.data(), sizeof(int) * np, hipMemcpyHostToDevice));
  // first time step sizes and predictions
  for (const auto& g : w.groups) EXA_TRY(exa::launch_update(g.sd, 0, upd_args(w, g), g.ncells, b->stream));
  EXA_TRY(exa::launch_patch_dt_finalize(w.dt, w.t, w.slots, np, kSlotsPerPatch, false, b->stream));
  for (const auto& g : w.groups) EXA_TRY(exa::launch_predictor(g.sd, false, pred_args(w, g), g.ncells, b->stream));
  EXA_TRY(hipStreamSynchronize(b->stream));
  return EXA_OK;
}

// Time the predictor of each order group in each of its variants on the
// wave's current state (the predictor is a pure function of luh and dt, so
// repeated launches leave the state unchanged) and keep the fastest.
int autotune(exa_patchset* b, Wave& w) {
  hipEvent_t e0, e1;
  EXA_TRY(hipEventCreate(&e0));
  EXA_TRY(hipEventCreate(&e1));
  for (auto& g : w.groups) {
    const unsigned variants = exa::predictor_variants(g.sd);
    if (variants == 1u) continue;
    float best = 1e30f;
    int choice = 0;
    for (int v = 0; v < exa::kPredictorVariants; ++v) {
      if (!(variants & (1u << v))) continue;
      g.variant = v;
      EXA_TRY(exa::launch_predictor(g.sd, false, pred_args(w, g), g.ncells, b->stream));  // warm-up
      EXA_TRY(hipEventRecord(e0, b->stream));
      for (int r = 0; r < 3; ++r) EXA_TRY(exa::launch_predictor(g.sd, false, pred_args(w, g), g.ncells, b->stream));
      EXA_TRY(hipEventRecord(e1, b->stream));
      EXA_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      EXA_TRY(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        choice = v;
      }
    }
    g.variant = choice;
    b->variant_choice[g.pde][g.order] = choice;
    // leave the state as the chosen variant computes it
    EXA_TRY(exa::launch_predictor(g.sd, false, pred_args(w, g), g.ncells, b->stream));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  EXA_TRY(hipStreamSynchronize(b->stream));
  return EXA_OK;
}

}  // namespace

extern "C" int exa_patchset_create(const exa_patch_spec* patches, int npatches, int precision, double hbm_budget_gb,
                                   long keep_max_cells, exa_patchset** out) {
  if (!out || !patches || npatches < 1 || !(hbm_budget_gb > 0)) return EXA_E_INVALID;
  if (precision != EXA_PRECISION_FP64 && precision != EXA_PRECISION_FP32_BULK) return EXA_E_INVALID;
  *out = nullptr;
  exa_patchset* b = new exa_patchset();
  b->patches.assign(patches, patches + npatches);
  b->precision = precision;
  b->budget = hbm_budget_gb * 1e9;
  b->keep_max_cells = keep_max_cells;
  for (const auto& p : b->patches)
    if (!patch_supported(p)) {
      delete b;
      return EXA_E_UNSUPPORTED;
    }
  // waves: largest patches first, first fit under the budget
  std::vector<int> order(npatches);
  std::iota(order.begin(), order.end(), 0);
  auto bytes = [&](int i) {
    const auto& p = b->patches[i];
    return (double)p.n * p.n * p.n * bytes_per_cell(patch_desc(p.pde, p.order, precision));
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return bytes(x) > bytes(y); });
  std::vector<double> load;
  for (int i : order) {
    if (bytes(i) > b->budget) {
      delete b;
      return EXA_E_NOMEM;
    }
    size_t w = 0;
    while (w < load.size() && load[w] + bytes(i) > b->budget) ++w;
    if (w == load.size()) {
      load.push_back(0);
      b->waves.emplace_back();
    }
    load[w] += bytes(i);
    b->waves[w].push_back(i);
  }
  for (auto& w : b->waves) std::sort(w.begin(), w.end());
  b->t.assign(npatches, 0.0);
  b->dt.assign(npatches, 0.0);
  b->solutions.resize(npatches);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return EXA_E_HIP;
  }
  *out = b;
  return EXA_OK;
}

extern "C" int exa_patchset_run(exa_patchset* b, int warmup_steps, int nsteps, int use_graph, int autotune_variants) {
  if (!b || nsteps < 0 || warmup_steps < 0) return EXA_E_INVALID;
  b->seconds = 0;
  b->dof_updates = 0;
  for (int p = 0; p < kPdes; ++p)
    for (int o = 0; o < 10; ++o) {
      b->pred_ms[p][o] = b->corr_ms[p][o] = 0.0;
      b->launches[p][o] = 0;
      b->picard[p][o] = b->pred_cells[p][o] = 0;
    }
  hipEvent_t e0, e1;
  EXA_TRY(hipEventCreate(&e0));
  EXA_TRY(hipEventCreate(&e1));
  for (const auto& ids : b->waves) {
    Wave w;
    int st = build_wave(b, ids, w);
    if (st == EXA_OK && autotune_variants) st = autotune(b, w);
    if (st == EXA_OK && !autotune_variants)
      for (auto& g : w.groups)
        if (b->variant_choice[g.pde][g.order] >= 0) g.variant = b->variant_choice[g.pde][g.order];
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (st == EXA_OK && use_graph) {
      hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t es = wave_step(w, b->fused_corrector, b->stream);
        e = hipStreamEndCapture(b->stream, &graph);
        if (e == hipSuccess) e = es;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) st = from_hip(e);
    }
    if (st == EXA_OK) {
      hipError_t e = hipSuccess;
      for (int n = 0; n < warmup_steps && e == hipSuccess; ++n)
        e = use_graph ? hipGraphLaunch(exec, b->stream) : wave_step(w, b->fused_corrector, b->stream);
      // profiling: the timed steps eagerly, with events around every corrector
      // and predictor launch and the Picard iterations counted per group
      StepTimer timer;
      if (b->profiling) timer.b = b;
      if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
      for (int n = 0; n < nsteps && e == hipSuccess; ++n)
        e = use_graph && !b->profiling ? hipGraphLaunch(exec, b->stream)
                                       : wave_step(w, b->fused_corrector, b->stream, b->profiling ? &timer : nullptr);
      if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      for (size_t k = 0; k < timer.ev.size(); ++k) {
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, timer.ev[k].first, timer.ev[k].second);
        const Group& g = w.groups[timer.where[k].second];
        if (timer.where[k].first == 0) {
          b->pred_ms[g.pde][g.order] += t;
          b->launches[g.pde][g.order] += 1;
        } else {
          b->corr_ms[g.pde][g.order] += t;
        }
      }
      for (auto& p : timer.ev) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
      }
      if (b->profiling)
        for (const auto& g : w.groups) {
          exa::StepState c;
          if (e == hipSuccess) e = hipMemcpy(&c, g.counters, sizeof(c), hipMemcpyDeviceToHost);
          b->picard[g.pde][g.order] += c.picard_total;
          b->pred_cells[g.pde][g.order] += c.predictor_cells;
        }
      if (e != hipSuccess) st = from_hip(e);
      b->seconds += ms * 1e-3;
      for (int id : ids) {
        const auto& p = b->patches[id];
        b->dof_updates += (double)p.n * p.n * p.n * (p.order + 1) * (p.order + 1) * (p.order + 1) * nsteps;
      }
    }
    if (st == EXA_OK) {  // per-patch time, dt and (optionally) solutions
      const int np = (int)ids.size();
      std::vector<double> t(np), dt(np);
      hipError_t e = hipMemcpy(t.data(), w.t, sizeof(double) * np, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(dt.data(), w.dt, sizeof(double) * np, hipMemcpyDeviceToHost);
      for (int i = 0; i < np && e == hipSuccess; ++i) {
        b->t[ids[i]] = t[i];
        b->dt[ids[i]] = dt[i];
      }
      if (b->keep_max_cells > 0)
        for (const auto& g : w.groups) {
          const Sizes z(g.sd);
          long base = 0;
          for (int i : g.patches) {
            const auto& p = b->patches[ids[i]];
            const long nc = (long)p.n * p.n * p.n;
            auto& out = b->solutions[ids[i]];
            if (nc > b->keep_max_cells) {
              base += nc;
              continue;
            }
            out.resize((size_t)nc * z.B3 * z.D);
            if (e == hipSuccess)
              e = hipMemcpy(out.data(), g.luh + (size_t)base * z.B3 * z.D, sizeof(double) * out.size(),
                            hipMemcpyDeviceToHost);
            base += nc;
          }
        }
      if (e != hipSuccess) st = from_hip(e);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamSynchronize(b->stream);
    free_wave(w);
    if (st != EXA_OK) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return st;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return EXA_OK;
}

extern "C" int exa_patchset_stats(exa_patchset* b, double* seconds, double* dof_updates, int* nwaves) {
  if (!b) return EXA_E_INVALID;
  if (seconds) *seconds = b->seconds;
  if (dof_updates) *dof_updates = b->dof_updates;
  if (nwaves) *nwaves = (int)b->waves.size();
  return EXA_OK;
}

extern "C" int exa_patchset_predictor_variant(exa_patchset* b, int order) {
  return exa_patchset_group_variant(b, EXA_PDE_EULER, order);
}

extern "C" int exa_patchset_group_variant(exa_patchset* b, int pde, int order) {
  if (!b || order < 0 || order > 9 || pde < 0 || pde >= kPdes) return -1;
  return b->variant_choice[pde][order];
}

extern "C" int exa_patchset_set_fused_corrector(exa_patchset* b, int on) {
  if (!b) return EXA_E_INVALID;
  b->fused_corrector = on != 0;
  return EXA_OK;
}

extern "C" int exa_patchset_set_profiling(exa_patchset* b, int on) {
  if (!b) return EXA_E_INVALID;
  b->profiling = on != 0;
  return EXA_OK;
}

extern "C" int exa_patchset_group_times(exa_patchset* b, int pde, int order, double* predictor_ms,
                                        double* corrector_ms, long* launches, unsigned long long* picard_iterations,
                                        unsigned long long* predictor_cells) {
  if (!b || order < 0 || order > 9 || pde < 0 || pde >= kPdes) return EXA_E_INVALID;
  if (predictor_ms) *predictor_ms = b->pred_ms[pde][order];
  if (corrector_ms) *corrector_ms = b->corr_ms[pde][order];
  if (launches) *launches = b->launches[pde][order];
  if (picard_iterations) *picard_iterations = b->picard[pde][order];
  if (predictor_cells) *predictor_cells = b->pred_cells[pde][order];
  return EXA_OK;
}

extern "C" int exa_patchset_solution(exa_patchset* b, int patch, double* luh) {
  if (!b || !luh || patch < 0 || patch >= (int)b->patches.size()) return EXA_E_INVALID;
  const auto& s = b->solutions[patch];
  if (s.empty()) return EXA_E_INVALID;
  std::memcpy(luh, s.data(), sizeof(double) * s.size());
  return EXA_OK;
}

extern "C" int exa_patchset_time(exa_patchset* b, int patch, double* t, double* dt) {
  if (!b || patch < 0 || patch >= (int)b->patches.size()) return EXA_E_INVALID;
  if (t) *t = b->t[patch];
  if (dt) *dt = b->dt[patch];
  return EXA_OK;
}

extern "C" int exa_patchset_destroy(exa_patchset* b) {
  if (!b) return EXA_OK;
  if (b->stream) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipStreamDestroy(b->stream);
  }
  delete b;
  return EXA_OK;
}
