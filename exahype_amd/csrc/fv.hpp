// fv.hpp -- internal interface of the finite-volume kernels: the single-patch
// Godunov update (reference layout) and the device-resident finite-volumes
// grid (dense subcell planes, first-order Godunov with Rusanov fluxes).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aderdg.hpp"

namespace exa {

// One patch (S,S,S,5), S = L + 2g, variable fastest (godunov.cpph:47).
// luh_new holds a copy of luh on entry and receives the reference's writes:
// the interior update and the boundary-face fluxes added to the first ghost
// layer.  dt_max (bits of a positive double, DBL_MAX on entry) receives the
// minimum over the patch's faces of cfl/3 h_d / s_max.
struct GodunovPatchArgs {
  const double* luh;
  double* luh_new;
  int L, g;
  double h[3];
  double dt;
  double cfl;
  unsigned long long* dt_max;
};
hipError_t launch_godunov_patch(const GodunovPatchArgs& a, hipStream_t s);

// Time-step state of a finite-volumes grid (device resident).  The new
// solution's stable dt, min over subcells and directions of
// (cfl/3 h_d) / max|lambda_d| (FVC:409-460), is formed as
// min over d of ch[d] / (max over subcells of max|lambda_d|): a correctly
// rounded division by a positive number is monotone, so both give the same
// double.  The update kernel only takes maxima.
struct FvState {
  double t;                       // time stamp
  double dt;                      // time step size of the next update
  unsigned long long lam_max[3];  // running maxima of max|lambda_d| (double bits; 0 = none)
  double ch[3];                   // cfl / 3 * h_d
  long long steps;
  int bad;                        // a non-finite or non-positive dt was produced
};

// The grid's solution is one plane set per variable, u[v][Z][Y][X] with
// X = nx L etc.: subcell (cell (i,j,k), local (x,y,z)) sits at global
// (k L + z, j L + y, i L + x).  Ghost layers do not exist; neighbour reads
// wrap (periodic) or clamp to the boundary subcell (transmissive with ghost
// width 1, where the reference's mirror copies that subcell).
struct FvGridArgs {
  const double* u;
  double* un;
  int X, Y, Z;
  int periodic;
  double h[3];   // subcell sizes dx_d / L
  FvState* state;
  int nsx, nry, nzc, zlen, ntiles;  // tiling of the update kernel
};
// One Godunov update u -> un with dt = state->dt, fused with the new
// solution's eigenvalue maxima (atomic maxima into state->lam_max).
hipError_t launch_fv_godunov_step(const FvGridArgs& a, hipStream_t s);
// The MUSCL-Hancock grid's patches [cells][2][S^3][5] (S = L + 2G, current
// buffer cur[cell]): the stable dt of the current interiors (begin: the
// first dt, t = 0; else the end of a step, t += dt); cell-major interiors
// [cells][L^3][5] to or from the current buffers; the next dt into a step
// state.
hipError_t launch_fv_patch_dt(const double* lim, const uint8_t* cur, long ncells, int L, int G, FvState* st,
                              bool begin, hipStream_t s);
hipError_t launch_fv_patch_io(double* cells, double* lim, const uint8_t* cur, long ncells, int L, int G, bool to_patch,
                              hipStream_t s);
hipError_t launch_fv_dt_to_step(const FvState* fs, StepState* ss, hipStream_t s);
// t += dt; dt = min_d ch[d] / lam_max[d]; lam_max = 0; steps += 1.
hipError_t launch_fv_finalize(FvState* st, hipStream_t s);
// Eigenvalue maxima of u, then t = 0 and dt as above.
hipError_t launch_fv_initial_dt(const FvGridArgs& a, hipStream_t s);
// Host layout [cells][L^3][5] (cell-major, subcells z,y,x) <-> planes.
hipError_t launch_fv_pack(const double* cells, double* planes, int nx, int ny, int nz, int L, hipStream_t s);
hipError_t launch_fv_unpack(const double* planes, double* cells, int nx, int ny, int nz, int L, hipStream_t s);
// Euler initial data (ICArgs' formula) at the subcell centres lo + (g + 1/2) h.
hipError_t launch_fv_euler_ic(const ICArgs& ic, double* planes, int X, int Y, int Z, const double* h,
                              hipStream_t s);

// Workgroup shape of the update kernel: kFvRows wavefronts, each a row of 64
// consecutive subcells along x; rows 1..kFvRows-2 and lanes 1..62 produce
// output, the outer rows and lanes supply neighbour values.
constexpr int kFvRows = 8;
constexpr int kFvOutX = 64 - 2;

}  // namespace exa
