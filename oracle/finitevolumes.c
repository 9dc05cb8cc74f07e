/*
 * finitevolumes.c -- CPU restatement of the reference's first-order Godunov
 * finite-volume update and of the stand-alone finite-volumes solver's time
 * step on one uniform level.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline; see oracle.h).
 *
 * Citations (paths relative to the reference checkout):
 *   GOD = ExaHyPE/kernels/finitevolumes/godunov/c/3d/godunov.cpph
 *   FVS = ExaHyPE/exahype/solvers/FiniteVolumesSolver.cpp
 *   FVC = ExaHyPE/kernels/finitevolumes/commons/c/commons.cpph
 *   RUN = ExaHyPE/exahype/runners/Runner.cpp
 * Patches are (S,S,S,D), S = L + 2g, variable fastest (GOD:47).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline double std_min(double a, double b) { return (b < a) ? b : a; }

/* godunov::c::solutionUpdate, GOD:35-143, for a PDE with flux only (no
 * source, no NCP).  luh_new must hold a copy of luh on entry (FVS:669-671);
 * the face loops also add the boundary-face fluxes to the first ghost layer
 * of luh_new, as the reference does.  dx is the patch (cell) size.  Returns
 * dt_max_allowed = min over faces of cfl/3 * h_d / s_max. */
double ora_godunov_update(const ora_solver* s, int L, int g, double* luh_new, const double* luh,
                          const double* dx, double dt, double cfl) {
  const int V = s->nvar, D = s->nvar + s->npar;
  const int pb = g, pe = g + L, S = L + 2 * g;
  const double cellSize[3] = {dx[0] / L, dx[1] / L, dx[2] / L};
  const double cdt[3] = {dt / cellSize[0], dt / cellSize[1], dt / cellSize[2]};
#define IDX(i, j, k, l) ((((long)(i) * S + (j)) * S + (k)) * D + (l))
  double fL[16], fR[16];
  double dt_max_allowed = DBL_MAX;
  /* x faces (GOD:61-83) */
  for (int i = pb; i < pe; i++)
    for (int j = pb; j < pe; j++)
      for (int k = pb - 1; k < pe; k++) {
        const double smax = ora_fv_rusanov(s, fL, fR, luh + IDX(i, j, k, 0), luh + IDX(i, j, k + 1, 0), 0);
        dt_max_allowed = std_min(dt_max_allowed, cfl / 3 * cellSize[0] / smax);
        for (int l = 0; l < V; ++l) {
          luh_new[IDX(i, j, k, l)] -= cdt[0] * fL[l];
          luh_new[IDX(i, j, k + 1, l)] += cdt[0] * fR[l];
        }
      }
  /* y faces (GOD:85-103) */
  for (int i = pb; i < pe; i++)
    for (int j = pb - 1; j < pe; j++)
      for (int k = pb; k < pe; k++) {
        const double smax = ora_fv_rusanov(s, fL, fR, luh + IDX(i, j, k, 0), luh + IDX(i, j + 1, k, 0), 1);
        dt_max_allowed = std_min(dt_max_allowed, cfl / 3 * cellSize[1] / smax);
        for (int l = 0; l < V; ++l) {
          luh_new[IDX(i, j, k, l)] -= cdt[1] * fL[l];
          luh_new[IDX(i, j + 1, k, l)] += cdt[1] * fR[l];
        }
      }
  /* z faces (GOD:105-123) */
  for (int i = pb - 1; i < pe; i++)
    for (int j = pb; j < pe; j++)
      for (int k = pb; k < pe; k++) {
        const double smax = ora_fv_rusanov(s, fL, fR, luh + IDX(i, j, k, 0), luh + IDX(i + 1, j, k, 0), 2);
        dt_max_allowed = std_min(dt_max_allowed, cfl / 3 * cellSize[2] / smax);
        for (int l = 0; l < V; ++l) {
          luh_new[IDX(i, j, k, l)] -= cdt[2] * fL[l];
          luh_new[IDX(i + 1, j, k, l)] += cdt[2] * fR[l];
        }
      }
#undef IDX
  return dt_max_allowed;
}

/* ------------------------------------------------------------------------ */
/* Stand-alone finite-volumes grid                                           */
/* ------------------------------------------------------------------------ */
/* One uniform level of FiniteVolumesSolver cells, each holding a patch of   */
/* L^3 subcells with g ghost layers, under global time stepping.  One time   */
/* step restates the non-fused Runner step (RUN:1116-1162) for this solver:  */
/*  1 Merging: per interior face ghostLayerFilling both ways (FVS:767-803,   */
/*    FVC:874-980); per boundary face boundaryLayerExtraction, transmissive  */
/*    boundaryValues and ghostLayerFillingAtBoundary (FVS:805-846).          */
/*  2 SolutionUpdate: updateSolution (FVS:662-715: copy to the previous      */
/*    solution, then the scheme's solutionUpdate when dt > 0) and            */
/*    startNewTimeStep (FVS:553-574: stableTimeStepSize of the new solution, */
/*    FVC:409-460, with the solver's CFL).                                   */
/*  3 The global dt is the minimum over cells (global time stepping,         */
/*    FVS:180-199); the time stamp advances by the dt just used.             */
/* Ghost layers never filled by a merge (edges, corners) stay zero, as the  */
/* reference's freshly allocated patches are (FVS:477-481).                  */
/* ------------------------------------------------------------------------ */
struct ora_fvgrid {
  ora_solver s;
  int scheme, L, g, S, nx, ny, nz, periodic;
  long ncells, n_lim;
  double dx[3];
  double t, dt;
  double *lim, *prev;
};

#define FV_PATCH(gr, c) ((gr)->lim + (c) * (gr)->n_lim)

ora_fvgrid* ora_fvgrid_create(const ora_solver* s, int scheme, int L, int nx, int ny, int nz, const double* lo,
                              const double* hi, int periodic) {
  if (scheme != ORA_FV_GODUNOV && scheme != ORA_FV_MUSCLHANCOCK) return NULL;
  ora_fvgrid* g = (ora_fvgrid*)calloc(1, sizeof(ora_fvgrid));
  g->s = *s;
  g->scheme = scheme;
  g->L = L;
  g->g = scheme == ORA_FV_GODUNOV ? 1 : 2; /* GenericFiniteVolumes{Godunov,MUSCLHancock}InC.java:12 */
  g->S = L + 2 * g->g;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->periodic = periodic;
  g->ncells = (long)nx * ny * nz;
  g->dx[0] = (hi[0] - lo[0]) / nx;
  g->dx[1] = (hi[1] - lo[1]) / ny;
  g->dx[2] = (hi[2] - lo[2]) / nz;
  g->n_lim = (long)g->S * g->S * g->S * (s->nvar + s->npar);
  g->lim = (double*)calloc(g->ncells * g->n_lim, sizeof(double));
  g->prev = (double*)calloc(g->n_lim, sizeof(double));
  return g;
}

void ora_fvgrid_destroy(ora_fvgrid* g) {
  if (!g) return;
  free(g->lim);
  free(g->prev);
  free(g);
}

long ora_fvgrid_num_cells(const ora_fvgrid* g) { return g->ncells; }
double ora_fvgrid_time(const ora_fvgrid* g) { return g->t; }
double ora_fvgrid_dt(const ora_fvgrid* g) { return g->dt; }

static long fv_cell(const ora_fvgrid* g, int i, int j, int k) {
  if (g->periodic) {
    i = (i + g->nx) % g->nx;
    j = (j + g->ny) % g->ny;
    k = (k + g->nz) % g->nz;
  } else if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz) {
    return -1;
  }
  return ((long)k * g->ny + j) * g->nx + i;
}

static double fv_min_dt(const ora_fvgrid* g) {
  double dt = DBL_MAX;
  for (long c = 0; c < g->ncells; c++)
    dt = std_min(dt, ora_fv_stable_dt(&g->s, g->L, g->g, FV_PATCH(g, c), g->dx, g->s.cfl));
  return dt;
}

/* Interior values [ncells][L^3][D] (z,y,x,var); ghosts zeroed.  Then the
 * first time step size of the initial solution. */
void ora_fvgrid_set_solution(ora_fvgrid* g, const double* interior) {
  const int L = g->L, G = g->g, S = g->S, D = g->s.nvar + g->s.npar;
  memset(g->lim, 0, sizeof(double) * g->ncells * g->n_lim);
  for (long c = 0; c < g->ncells; c++) {
    const double* src = interior + c * (long)L * L * L * D;
    for (int z = 0; z < L; z++)
      for (int y = 0; y < L; y++)
        memcpy(FV_PATCH(g, c) + (((long)(z + G) * S + (y + G)) * S + G) * D, src + ((long)z * L + y) * L * D,
               sizeof(double) * L * D);
  }
  g->t = 0.0;
  g->dt = fv_min_dt(g);
}

void ora_fvgrid_get_solution(const ora_fvgrid* g, double* interior) {
  const int L = g->L, G = g->g, S = g->S, D = g->s.nvar + g->s.npar;
  for (long c = 0; c < g->ncells; c++) {
    double* dst = interior + c * (long)L * L * L * D;
    for (int z = 0; z < L; z++)
      for (int y = 0; y < L; y++)
        memcpy(dst + ((long)z * L + y) * L * D, FV_PATCH(g, c) + (((long)(z + G) * S + (y + G)) * S + G) * D,
               sizeof(double) * L * D);
  }
}

void ora_fvgrid_get_patch(const ora_fvgrid* g, long c, double* patch) {
  memcpy(patch, FV_PATCH(g, c), sizeof(double) * g->n_lim);
}

static void fv_step(ora_fvgrid* g) {
  const int L = g->L, G = g->g, D = g->s.nvar + g->s.npar;
  /* 1. merging */
  for (int d = 0; d < 3; d++)
    for (long c = 0; c < g->ncells; c++) {
      const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
      const long r = fv_cell(g, i + (d == 0), j + (d == 1), k + (d == 2));
      if (r < 0) continue;
      ora_fv_ghost_fill(L, G, D, FV_PATCH(g, c), FV_PATCH(g, r), d, +1);
      ora_fv_ghost_fill(L, G, D, FV_PATCH(g, r), FV_PATCH(g, c), d, -1);
    }
  if (!g->periodic)
    for (long c = 0; c < g->ncells; c++) {
      const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
      const int at[6] = {i == 0, i == g->nx - 1, j == 0, j == g->ny - 1, k == 0, k == g->nz - 1};
      for (int f = 0; f < 6; f++)
        if (at[f]) ora_fv_boundary_transmissive(L, G, D, FV_PATCH(g, c), f);
    }
  /* 2. solution update and the new cell time step sizes */
  double dt_next = DBL_MAX;
  for (long c = 0; c < g->ncells; c++) {
    double* p = FV_PATCH(g, c);
    memcpy(g->prev, p, sizeof(double) * g->n_lim);
    if (g->dt > 0) {
      if (g->scheme == ORA_FV_GODUNOV)
        ora_godunov_update(&g->s, L, G, p, g->prev, g->dx, g->dt, g->s.cfl);
      else
        ora_muscl_update(&g->s, L, G, p, g->prev, g->dx, g->dt);
    }
    dt_next = std_min(dt_next, ora_fv_stable_dt(&g->s, L, G, p, g->dx, g->s.cfl));
  }
  /* 3. global time stepping */
  g->t += g->dt;
  g->dt = dt_next;
}

/* Advance nsteps; dt_out[nsteps] (may be NULL) receives the dt of each step. */
void ora_fvgrid_step(ora_fvgrid* g, int nsteps, double* dt_out) {
  for (int n = 0; n < nsteps; n++) {
    if (dt_out) dt_out[n] = g->dt;
    fv_step(g);
  }
}
