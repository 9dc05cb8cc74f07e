// aderdg.hpp -- internal interface between the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace exa {

// kPdeUser: a PDE compiled by the user (exahype_amd/user_solver.hpp); the
// grid then launches the user's instantiations through exa_user_solver.
enum : int { kPdeEuler = 0, kPdeElastic = 1, kPdeUser = 100 };
// Terms of the PDE (the spec file's `terms`): flux, algebraic source, NCP.
enum : int { kFlux = 1, kSource = 2, kNCP = 4 };

// Runtime description of a solver instantiation (the reference's SolverType
// constants + the spec file's `terms`).
struct SolverDesc {
  int pde;
  int order;
  int nvar;
  int npar;
  int terms;  // kFlux | kSource | kNCP
  int no_time_averaging;
  double cfl;
  int precision;  // kFP64 or kFP32Bulk
  int linear;     // linear ADER-DG: the *Linear kernels (Cauchy-Kowalevski predictor)
};

// Precision forms.  kFP64: everything in double (the reference's arithmetic).
// kFP32Bulk: predictor, volume integral, Riemann solver and the face /
// update arrays in float; surface integral, solution update and the time
// step size in double on the double solution.
enum : int { kFP64 = 0, kFP32Bulk = 1 };

// A face between two cells; `left` owns face 2*dir+1, `right` owns face 2*dir.
// pad: 0 for such a face; in linear grids with transmissive boundaries, 1 marks
// the boundary face 2*dir of cell `right` and 2 the boundary face 2*dir+1 of
// cell `left` (the other index is unused).
struct FaceDesc {
  int left;
  int right;
  int dir;
  int pad;
};

constexpr int kDtSlots = 4096;
constexpr unsigned long long kDtSlotEmpty = 0x7FF0000000000000ull;  // +inf

// Device-resident time-stepping state of one grid.
struct StepState {
  double dt;  // current (predictor = corrector) time step size
  double t;   // time stamp of the current solution
  long steps;
  unsigned long long picard_total;  // sum of per-cell Picard counts (profiling only)
  unsigned long long predictor_cells;
  double* dt_hist;  // dt of every step (device), may be null
  long dt_hist_len;
  unsigned long long dt_slots[kDtSlots];  // per-slot minima of the next dt
  // fused time stepping: dt above is the corrector time step size, dtp the
  // predictor's; violated = the last stability check failed (the prediction
  // is re-run with the new dtp); reruns counts those
  double dtp;
  int violated;
  long long reruns;
  double dtp_checked;  // limiting solver: the predictor dt after the first check
};

// Arrays marked (bulk) hold double in the kFP64 form and float in kFP32Bulk.
struct PredictorArgs {
  const double* luh;  // [cells][B^3][V+P]
  void* lduh;         // (bulk) [cells][B^3][V]  (volume integral; may be null)
  void* lQhbnd;       // (bulk) [cells][6][B^2][V+P]
  void* lFhbnd;       // (bulk) [cells][6][B^2][V]
  const double* dx;   // dx[cell * dx_stride + d]
  long dx_stride;     // 0: one dx for all cells, 3: per cell
  const double* dt;   // device scalar
  int* iters;         // [cells] Picard iterations, may be null
  long cell0;         // first cell of this launch
  // optional exports of the reference's temporaries (single-cell API only)
  double* lQi;   // (z,y,x,t,var) B^4 (V+P)
  double* lFi;   // (t,z,y,x,{x,y,z,src},var) 4 B^4 V
  double* lQhi;  // (z,y,x,var)
  double* lFhi;  // 4 blocks of B^3 V
  // Device workspace for the time slices of the Picard iterate that fit neither
  // LDS nor registers (large orders): ws_cells cells of predictor_ws_bytes_per_cell
  // each; launches are split into chunks of ws_cells cells.  Unused (null, 0)
  // where predictor_ws_bytes_per_cell is 0.
  void* ws;
  long ws_cells;
  // optional gate: cells with status[cell] >= status_limit are skipped (the
  // limiter's troubled cells); null status = every cell
  const int* status;
  int status_limit;
  // optional batch of independent patches: cell -> patch; dt and dx are then
  // per patch (dt[patch], dx[3 * patch])
  const int* cell_patch;
  // Predictor variant (kPredictorAuto, kPredictorValu64, kPredictorValu144);
  // the batch autotuner picks one per order
  int variant;
  // optional gate for the whole launch: it does nothing unless *run_if != 0
  // (the fused time stepping's re-run after a violated stability condition)
  const int* run_if;
};

struct RiemannArgs {
  const FaceDesc* faces;
  long nfaces;
  const void* lQhbnd;  // (bulk)
  void* lFhbnd;        // (bulk)
  // optional gate: faces whose two cells both have status >= status_limit are
  // skipped (faces between FV-active limiter cells); null status = every face
  const int* status;
  int status_limit;
};

struct UpdateArgs {
  double* luh;
  void* lduh;           // (bulk)
  const void* lFhbnd;   // (bulk)
  const double* dx;
  long dx_stride;
  const double* dt;
  double cfl_pnpm;                // CFL * PNPM[order]
  unsigned long long* dt_slots;   // null: no time step computation
  double* dt_cell;                // [cells] per-cell stable dt, may be null
  long cell0;
  // optional batch of independent patches: dt[patch], dx[3 * patch] and
  // slots_per_patch dt slots per patch
  const int* cell_patch;
  int slots_per_patch;
};

// Boundary treatment of a grid's outer faces.  kBoundaryTransmissive: the
// outside state and flux are the inside ones; kBoundaryUser: the PDE's
// boundary_values (UserSolverInterface.h:171) at every face node.
enum : int { kBoundaryPeriodic = 0, kBoundaryTransmissive = 1, kBoundaryUser = 2 };

// The fused corrector of a uniform Cartesian grid (corrector.hpp): Riemann
// solves on each cell's six faces from the cell's and its neighbours' face
// blocks (or the boundary values), surface integral, solution update,
// optionally the PDE's adjust_point_solution at t + dt, and the cell's stable
// dt into dt_slots.  Cells are (k, j, i) row-major on [lo, lo + n dx].
struct CorrectorArgs {
  double* luh;
  const void* lduh;     // (bulk)
  const void* lQhbnd;   // (bulk)
  const void* lFhbnd;   // (bulk)
  const double* dx;     // 3 doubles (device)
  const double* dt;     // device scalar: the corrector time step size
  const double* t;      // device scalar: the corrector time stamp
  double lo[3];         // domain origin (node coordinates of boundary values / adjustment)
  double hdx[3];        // the cell size on the host side of the call (node coordinates)
  double cfl_pnpm;      // CFL * PNPM[order]
  unsigned long long* dt_slots;
  int nx, ny, nz;
  int boundary;         // kBoundary*
  int adjust;           // apply the PDE's adjust_point_solution after the update
  // optional batch of independent periodic patches (cubes): cell -> patch p,
  // whose cells are patch_first[p] .. patch_first[p] + patch_n[p]^3 - 1; dx,
  // dt and t per patch (dx[3 p], dt[p], t[p]); slots_per_patch dt slots per
  // patch.  nx = the number of cells, ny = nz = 1.
  const int* cell_patch;
  const long* patch_first;
  const int* patch_n;
  int slots_per_patch;
};
hipError_t launch_corrector(const SolverDesc& sd, const CorrectorArgs& a, hipStream_t s);
// Whether the library's instantiation for this solver supports a boundary
// treatment and the per-step adjustment (the library's PDEs have neither hook).
bool corrector_supports(const SolverDesc& sd, int boundary, bool adjust);

// solutionAdjustment (solutionAdjustment.cpph:17-47) of every cell of a grid:
// the PDE's adjust_point_solution at every node, at time *t with step size dt_val.
struct AdjustArgs {
  double* luh;
  const double* t;  // device scalar
  double dt;
  double lo[3], dx[3];
  int nx, ny, nz;
};

struct VolumeArgs {
  double* lduh;
  const double* lFhi;
  const double* dx;
  long dx_stride;
};

constexpr int kMaxModes = 8;

struct ICArgs {
  double* luh;
  long ncells;
  int nx, ny, nz;
  int order, nvar, npar;
  double lo[3];
  double dx[3];
  // rho = rho0 + amp sin(2 pi k.x) + sum_m mode_amp sin(2 pi k_m.x + phi_m0),
  // u_i = vel_i + sum_m mode_amp sin(2 pi k_m.x + phi_mi), p constant.
  double rho0, amp, k[3], vel[3], p;
  int nmodes;
  double mode_amp;
  double mode_k[kMaxModes][3];
  double mode_phase[kMaxModes][4];
};

// Elastic workload state: zero stresses, Gaussian velocity pulse u = v = w =
// amp exp(-|x - c|^2 / (2 width^2)) at the centre c of [0,1]^3, and a seeded
// two-layer material (uniform draws r[6] from the seed): interface
// z0 = 0.5 + 0.1 sin(2 pi x + 2 pi r0) sin(2 pi y + 2 pi r1); rho = 2.05 + 0.9 r2|r3
// + 0.05 sin(2 pi (x+y+z)); cp = 4.1 + 1.8 r4|r5 + 0.1 sin(2 pi (x-y+z)),
// cs = cp / sqrt 3; parameters (lambda, mu, rho).
struct ElasticICArgs {
  double* luh;
  long ncells;
  int nx, ny, nz, order;
  double lo[3], dx[3];
  double r[6];
  double width, amp;
};
hipError_t launch_elastic_ic(const ElasticICArgs& a, hipStream_t s);

bool supported(const SolverDesc& sd);
// Kernels that only depend on (order, nvar): volume and surface integrals.
bool supported_sizes(const SolverDesc& sd);
hipError_t launch_predictor(const SolverDesc& sd, bool export_temps, const PredictorArgs& a,
                            long ncells, hipStream_t s);
// Predictor variants.  kPredictorAuto: the measured default, the vector-ALU
// kernel (predictor_kernel) with its 144 KB LDS budget; kPredictorValu64 /
// kPredictorValu144: that kernel with a 64 KB / 144 KB LDS budget for the
// Picard iterate (see PredictorLayout); kPredictorMfma: the FP32-bulk kernel
// with its contractions on the matrix cores (predictor_mfma.hip), where
// compiled (predictor_mfma_supported).
enum : int { kPredictorAuto = 0, kPredictorValu64 = 1, kPredictorValu144 = 2, kPredictorMfma = 3 };
constexpr int kPredictorVariants = 4;
// Bytes of PredictorArgs::ws one cell needs (0: no workspace), for a variant.
size_t predictor_ws_bytes_per_cell(const SolverDesc& sd, int variant = kPredictorAuto);
// Bit v set: variant v runs a kernel of its own for this solver (bit 0 always;
// kPredictorValu144 is the Auto kernel, so its bit is never set).
unsigned predictor_variants(const SolverDesc& sd);
// Whether a variant can be requested for this solver (it may share Auto's kernel).
bool predictor_variant_available(const SolverDesc& sd, int variant);
// The FP32-bulk predictor with its contractions on the matrix cores
// (predictor_mfma.hip): solvers it is compiled for, and its launch (no
// workspace, no exports).
bool predictor_mfma_supported(const SolverDesc& sd);
hipError_t launch_predictor_mfma(const SolverDesc& sd, const PredictorArgs& a, long ncells, hipStream_t s);
// Cells per workspace chunk the host allocates for (bounds the workspace size).
constexpr long kPredictorWsCells = 4096;
hipError_t launch_riemann(const SolverDesc& sd, const RiemannArgs& a, hipStream_t s);
// update_kernel modes: surface integral, solution update, write lduh back.
// With sd.linear the surface integral is surfaceIntegralLinear (fluctuations,
// F_{2d+1} FRCoeff + F_{2d} FLCoeff, divided by dx).
enum : int { kUpdSurface = 1, kUpdUpdate = 2, kUpdWriteLduh = 4 };
hipError_t launch_update(const SolverDesc& sd, int mode, const UpdateArgs& a, long ncells,
                         hipStream_t s);
hipError_t launch_dt_finalize(StepState* st, bool advance_time, hipStream_t s);
// fused time stepping: the stability check, next predictor dt and
// startNewTimeStep (init: dt = dtp = the stable dt)
hipError_t launch_dt_finalize_fused(StepState* st, double weight, bool init, hipStream_t s);
// fused time stepping of the limiting solver.  Phase 1 is the check after the
// update; without an irregular limiter-domain change (*irregular == 0) it is
// the check above, with one it only keeps the checked predictor dt
// (dtp_checked) for the repeat.  Phase 2 (a no-op without an irregular
// change) repeats the check with the recomputed dt candidates and sets the
// corrector and predictor dt to the new stable dt.
hipError_t launch_dt_finalize_fused_limiting(StepState* st, double weight, const int* irregular, int phase,
                                             hipStream_t s);
// Per-patch time steps of a batch: dt[p] = min of the patch's slots (then
// reset); with advance_time, t[p] += the old dt[p] first.
hipError_t launch_patch_dt_finalize(double* dt, double* t, unsigned long long* slots, int npatches,
                                    int slots_per_patch, bool advance_time, hipStream_t s);
hipError_t launch_volume_integral(const SolverDesc& sd, const VolumeArgs& a, long ncells, hipStream_t s);
// One cell: lFhi = the time average of the space-time fluxes lFi.
hipError_t launch_time_average_lfi(int order, int V, const double* lFi, double* lFhi, hipStream_t s);
hipError_t launch_euler_ic(const ICArgs& a, hipStream_t s);

// ---- linear ADER-DG kernels (linear_kernels.hip), FP64 ----
struct LinearPredictorArgs {
  const double* luh;  // [cells][B^3][V+P]
  double* lduh;       // [cells][B^3][V] volumeIntegralLinear of the prediction (may be null)
  double* lQhbnd;     // [cells][6][B^2][V+P]
  double* lFhbnd;     // [cells][6][B^2][V]
  const double* dx;   // dx[cell * dx_stride + d]
  long dx_stride;
  const double* dt;   // device scalar
  const double* PSi;  // point sources (t,z,y,x,var) (B+1) B^3 V of one cell, or null
  long cell0;
  // optional exports of the reference's temporaries (single-cell API only)
  double* lQi;    // (t,z,y,x,var) (B+1) B^3 (V+P)
  double* lFi;    // (t,z,y,x,dir,var) B^4 3V + a zeroed source block of B^4 V
  double* gradQ;  // (t,z,y,x,dir,var) B^4 3V
  double* lQhi;   // (z,y,x,var) B^3 (V+P)
  double* lFhi;   // (dir,z,y,x,var) 3 B^3 V + a zeroed source block of B^3 V
};
bool linear_supported(const SolverDesc& sd);
bool linear_sizes_supported(const SolverDesc& sd);  // (order, nvar) of the linear integrals
// pointSource's projection of one cell's (B+1) B^3 V forces; x0 per space-time
// node (3 doubles), geo = center[3], dx[3] (device); orders 1..7.
hipError_t launch_point_source(const SolverDesc& sd, double* PSi, const double* x0, const double* geo,
                               hipStream_t s);
hipError_t launch_linear_predictor(const SolverDesc& sd, bool export_temps, const LinearPredictorArgs& a,
                                   long ncells, hipStream_t s);
// riemannSolverLinear over a face list (fluctuations into both cells' blocks)
hipError_t launch_linear_riemann(const SolverDesc& sd, const RiemannArgs& a, hipStream_t s);
// volumeIntegralLinear of one or more cells: lduh = -w^3 sum_d lFhi_d
hipError_t launch_linear_volume(const SolverDesc& sd, double* lduh, const double* lFhi, long ncells,
                                hipStream_t s);
// surfaceIntegralLinear of one or more cells (lduh in/out, one dx for all)
hipError_t launch_linear_surface(const SolverDesc& sd, double* lduh, const double* lFhbnd, const double* dx,
                                 long ncells, hipStream_t s);
// The Cartesian VTK plotter's interpolation (output_kernels.hip): luh
// [cells][B^3][D] -> values at each cell's (N+1)^3 equidistant vertices,
// [cells][B^3][D] (x fastest); orders 1..6.
hipError_t launch_plot_vertices(int order, int D, const double* luh, double* out, long ncells, hipStream_t s);
// st->picard_total += sum(iters[0..ncells)), st->predictor_cells += ncells
hipError_t launch_count_iterations(const int* iters, long ncells, StepState* st, hipStream_t s);

}  // namespace exa
