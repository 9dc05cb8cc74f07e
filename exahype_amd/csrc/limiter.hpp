// limiter.hpp -- internal interface of the a-posteriori subcell limiter
// kernels (finite-volume patches, DMP, limiter status machine).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aderdg.hpp"

namespace exa {

constexpr int kLimGhost = 2;  // ghost layers of an FV patch (MUSCL-Hancock)
constexpr int kLimMaxOrder = 3;

__host__ __device__ constexpr int lim_L(int order) { return 2 * order + 1; }
__host__ __device__ constexpr int lim_S(int order) { return lim_L(order) + 2 * kLimGhost; }

// Projection tables of every order 0..9 (host-computed, uploaded once):
// uh2lim[o][k*L + i], uh2lob[o][j*B + i], lim2uh[o][i*B + j].
struct LimTables {
  double uh2lim[10][19 * 10];
  double uh2lob[10][10 * 10];
  double lim2uh[10][19 * 10];
};
// Builds the tables on the host (LimiterProjectionMatrices.cpp:39-223 restated).
bool build_limiter_tables(LimTables* t);
// Uploads them to the device constant bank once per device; thread-safe.
hipError_t ensure_limiter_tables();

// Per-cell state of a limiting grid.  FV patches are (S,S,S,V) with S =
// 2N+1+4; each cell owns two patch buffers (current / previous), selected
// by cur[c], so a swap is a flip of one byte.
struct LimArgs {
  double* luh;        // [cells][B^3][V]
  double* luh_prev;   // [cells][B^3][V]
  const double* lduh; // [cells][B^3][V] (volume integral of the prediction)
  const double* lFhbnd;  // [cells][6][B^2][V]
  double* lim;        // [cells][2][S^3][V]
  uint8_t* cur;       // [cells] current patch buffer
  uint8_t* has_patch; // [cells] patch "allocated" (reference heap semantics)
  int* st;            // [cells] limiter status
  int* pst;           // [cells] status before the current update
  int* st2;           // [cells] scratch for the status spreading sweeps
  int* itc;           // [cells] iterations to cure
  const int* nbr;     // [cells][6] face neighbour, -1 at a boundary face
  double* mn;         // [cells][6][V] solution minima per face block
  double* mx;         // [cells][6][V]
  double* limmm;      // [cells][2][V] min/max over the FV patch interior
  int* flags;         // [0] irregular change this step; [1] FV list size; [2] recompute list size
  int* fv_list;       // [cells]
  int* rc_list;       // [cells] cells with status > 0 after an irregular change (rolled back)
  double* loc_mm;     // [cells][2][V] local min/max of the DG solution after the update
  double* dt_cell;    // [cells] stable dt candidate of the DG solution after the update
  const double* dx;   // device [3]
  StepState* st_state;  // dt, dt slots
  double cfl_pnpm;
  double relax, scaling;
  int T, A, cure;
  int ncells;
  // optional: min over every FV face of 0.99/3 h_d / s_max (the admissible dt
  // musclhancock's solutionUpdate returns), as the bits of a positive double
  unsigned long long* fv_dt_max;
  // limiter-type godunov: the first-order update (godunov.cpph:35-143) in
  // place of MUSCL-Hancock; it reads only the first ghost layer
  int fv_godunov;
};

enum : int { kLimFlagIrregular = 0, kLimFlagFvCount = 1, kLimFlagRcCount = 2 };

// The launch_* functions below take the solver order (2..kLimMaxOrder, Euler,
// FP64).  "all" launches cover every cell; the FV launches walk fv_list.
bool limiter_supported(const SolverDesc& sd);
// pst = st; flags cleared; FV list of cells with st >= A.
hipError_t launch_lim_begin_step(const SolverDesc& sd, const LimArgs& a, hipStream_t s);
// per interior face: merge the min/max face blocks (mergeSolutionMinMaxOnFace)
hipError_t launch_lim_minmax_merge(const SolverDesc& sd, const LimArgs& a, const FaceDesc* faces, long nfaces,
                                   hipStream_t s);
// ghost layers of every listed FV cell (neighbour copies / transmissive);
// only_if_irregular: do nothing unless flags[irregular] is set
hipError_t launch_lim_ghosts(const SolverDesc& sd, const LimArgs& a, bool only_if_irregular, hipStream_t s);
// MUSCL-Hancock (or, with fv_godunov, Godunov) update of the listed FV cells
// + projection to DG (+ patch min/max)
hipError_t launch_lim_fv_update(const SolverDesc& sd, const LimArgs& a, bool only_if_irregular, hipStream_t s);
// DG update of cells with status < A, then dt, DMP, admissibility and the
// new limiter status of every cell
hipError_t launch_lim_update_status(const SolverDesc& sd, const LimArgs& a, hipStream_t s);
// irregular change: spreading sweeps, rollback, FV list rebuild
hipError_t launch_lim_spread_rollback(const SolverDesc& sd, const LimArgs& a, hipStream_t s);
// irregular change: recompute helpers, new dt and min/max of every cell
hipError_t launch_lim_recompute_finish(const SolverDesc& sd, const LimArgs& a, hipStream_t s);
// irregular change in fused time stepping: out[c] = 0 for the cells whose
// prediction is recomputed, T for the others (the predictor's status gate)
hipError_t launch_lim_repredict_mask(const LimArgs& a, int* out, hipStream_t s);
// initial state: DMP min/max, admissibility, spreading, FV initial condition
// into the patches of cells with status > 0 (lim_ic [cells][L^3][V] or null)
hipError_t launch_lim_init(const SolverDesc& sd, const LimArgs& a, const double* lim_ic, hipStream_t s);

// Single-cell style batched entry points (drop-in boundary of the reference's
// kernels::limiter calls).
hipError_t launch_project_on_fv(int order, int V, int g, const double* luh, double* lim, long ncells, hipStream_t s);
hipError_t launch_project_on_dg(int order, int V, int g, const double* lim, double* luh, long ncells, hipStream_t s);
// local min/max (GL nodes, Lobatto nodes, subcell means) -> mn, mx [cells][V];
// with bmin/bmax ([cells][6][V]) also the DMP: ok[cells] and faces 0/1 rewritten
hipError_t launch_dmp(int order, const double* luh, double* mn, double* mx, double* bmin, double* bmax, int* ok,
                      double relax, double scaling, long ncells, hipStream_t s);
// list[i] = i for i < n and flags[FV count] = n (drives launch_lim_fv_update
// over independent patches)
hipError_t launch_lim_iota(int* list, int n, int* flags, hipStream_t s);

}  // namespace exa
