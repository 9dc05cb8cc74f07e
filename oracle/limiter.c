/*
 * limiter.c -- CPU restatement of the reference's a-posteriori subcell
 * limiter path (3-D): projection tables, DG <-> FV projections, the discrete
 * maximum principle and min/max searches, the MUSCL-Hancock finite-volume
 * update with its pointwise Rusanov solver, the FV ghost-layer and
 * boundary commons, the ADER-DG boundary conditions, and a limiting grid
 * driver that restates the limiter status machine, neighbour merging,
 * rollback and local recomputation of LimitingADERDGSolver.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); the HIP product never includes,
 * links or calls this file.  Paths cited relative to the reference checkout:
 *   LPM = ExaHyPE/kernels/LimiterProjectionMatrices.cpp
 *   LIM = ExaHyPE/kernels/limiter/generic/c/Limiter.cpp
 *   MH  = ExaHyPE/kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph
 *   RUS = ExaHyPE/kernels/finitevolumes/riemannsolvers/c/rusanov.cpph
 *   FVC = ExaHyPE/kernels/finitevolumes/commons/c/commons.cpph
 *   LS  = ExaHyPE/exahype/solvers/LimitingADERDGSolver.cpp
 *   AS  = ExaHyPE/exahype/solvers/ADERDGSolver.cpp
 *   FVS = ExaHyPE/exahype/solvers/FiniteVolumesSolver.cpp
 * std::min(a,b) is restated as (b < a) ? b : a and std::max(a,b) as
 * (a < b) ? b : a (their definitions), so NaN behaviour matches too.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXA_TABLE static const
#include "../include/exahype_amd/device/dg_tables.inc"

static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* Projection tables, LPM:39-223                                             */
/* ------------------------------------------------------------------------ */

/* BaseFunc1D, LPM:139-150: Lagrange basis on the Gauss-Legendre nodes. */
static void base_func_1d(double* phi, double xi, int B) {
  const double* x = exa_gl_nodes[B - 1];
  for (int i = 0; i < B; i++) phi[i] = 1;
  for (int m = 0; m < B; m++)
    for (int j = 0; j < B; j++) {
      if (j == m) continue;
      phi[m] = phi[m] * (xi - x[j]) / (x[m] - x[j]);
    }
}

/* matrixInverse, LPM:153-223, including its pivot search that stores the
 * signed (not absolute) value of a new pivot candidate (:181).  Returns 0
 * for a singular matrix. */
static int matrix_inverse(int n, const double* a, double* ia) {
  double* c = (double*)malloc(sizeof(double) * n * 2 * n);
#define C_(i, j) c[(i) * 2 * n + (j)]
  for (int i = 0; i < n; i++) {
    int j;
    for (j = 0; j < n; j++) C_(i, j) = a[j * n + i];
    for (; j < 2 * n; j++) C_(i, j) = 0.;
    C_(i, i + n) = 1.;
  }
  for (int i = 0; i < n; i++) {
    int ml = i;
    double mlV = fabs(C_(i, i));
    for (int j = i + 1; j < n; j++) {
      if (fabs(C_(j, i)) > mlV) {
        ml = j;
        mlV = C_(j, i);
      }
    }
    for (int k = 0; k < 2 * n; k++) {
      const double tmp = C_(ml, k);
      C_(ml, k) = C_(i, k);
      C_(i, k) = tmp;
    }
    if (C_(i, i) == 0) {
      free(c);
      return 0;
    }
    const double piv = 1. / C_(i, i);
    for (int k = 0; k < 2 * n; k++) C_(i, k) *= piv;
    for (int j = i + 1; j < n; j++) {
      const double tmp = C_(j, i);
      for (int k = 0; k < 2 * n; k++) C_(j, k) -= tmp * C_(i, k);
    }
  }
  for (int i = n - 1; i >= 0; i--)
    for (int j = i - 1; j >= 0; j--) {
      const double tmp = C_(j, i);
      for (int k = 0; k < 2 * n; k++) C_(j, k) -= tmp * C_(i, k);
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) ia[j * n + i] = C_(i, j + n);
#undef C_
  free(c);
  return 1;
}

/* initLimiterProjectionMatrices for one order, LPM:48-136.
 * uh2lim[k*L + i]: mean of basis k over subcell i (B x L);
 * uh2lob[j*B + i]: basis j at Lobatto node i (B x B);
 * lim2uh[i*B + j]: constrained least-squares inverse (L x B). */
int ora_limiter_tables(int order, double* uh2lim, double* uh2lob, double* lim2uh) {
  const int B = order + 1, L = 2 * order + 1;
  double phi[16];
  memset(uh2lim, 0, sizeof(double) * B * L);
  {
    const double dxi = 1. / L;
    for (int i = 0; i < L; i++) {
      const double xLeft = i * dxi;
      for (int j = 0; j < B; j++) {
        const double xi = xLeft + dxi * exa_gl_nodes[B - 1][j];
        base_func_1d(phi, xi, B);
        for (int k = 0; k < B; k++) uh2lim[k * L + i] += exa_gl_weights[B - 1][j] * phi[k];
      }
    }
  }
  for (int i = 0; i < B; i++) {
    base_func_1d(phi, exa_lob_nodes[B - 1][i], B);
    for (int j = 0; j < B; j++) uh2lob[j * B + i] = phi[j];
  }
  {
    const int n1 = B + 1;
    double lsqm[11 * 11], ilsqm[11 * 11], lsqrhs[19 * 11];
    const double dxi = 1.0 / L;
    for (int i = 0; i < B; i++) {
      for (int j = 0; j < B; j++) {
        lsqm[i * n1 + j] = 0.;
        for (int k = 0; k < L; k++) lsqm[i * n1 + j] += 2 * uh2lim[i * L + k] * uh2lim[j * L + k];
      }
      lsqm[i * n1 + B] = exa_gl_weights[B - 1][i];
    }
    for (int i = 0; i < B; i++) lsqm[B * n1 + i] = -exa_gl_weights[B - 1][i];
    lsqm[B * n1 + B] = 0.;
    if (!matrix_inverse(n1, lsqm, ilsqm)) return 0;
    for (int i = 0; i < L; i++) {
      for (int j = 0; j < B; j++) lsqrhs[i * n1 + j] = 2 * uh2lim[j * L + i];
      lsqrhs[i * n1 + B] = dxi;
    }
    for (int i = 0; i < L; i++)
      for (int j = 0; j < B; j++) {
        lim2uh[i * B + j] = 0.;
        for (int k = 0; k < B + 1; k++) lim2uh[i * B + j] += ilsqm[k * n1 + j] * lsqrhs[i * n1 + k];
      }
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Projections, LIM:23-112 (stride V for luh and lim, as the reference).      */
/* ------------------------------------------------------------------------ */
void ora_project_on_fv(int order, int V, int g, const double* uh2lim, const double* luh,
                       double* lim) {
  const int B = order + 1, L = 2 * order + 1, S = L + 2 * g;
  for (int z = g; z < L + g; z++)
    for (int y = g; y < L + g; y++)
      for (int x = g; x < L + g; x++)
        for (int v = 0; v < V; v++) {
          double* o = lim + (((long)z * S + y) * S + x) * V + v;
          *o = 0;
          for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
              for (int ix = 0; ix < B; ix++)
                *o += luh[((iz * B + iy) * B + ix) * V + v] * uh2lim[iz * L + (z - g)] *
                      uh2lim[iy * L + (y - g)] * uh2lim[ix * L + (x - g)];
        }
}

void ora_project_on_dg(int order, int V, int g, const double* lim2uh, const double* lim,
                       double* luh) {
  const int B = order + 1, L = 2 * order + 1, S = L + 2 * g;
  for (int z = 0; z < B; z++)
    for (int y = 0; y < B; y++)
      for (int x = 0; x < B; x++)
        for (int v = 0; v < V; v++) {
          double* o = luh + ((z * B + y) * B + x) * V + v;
          *o = 0;
          for (int iz = g; iz < L + g; iz++)
            for (int iy = g; iy < L + g; iy++)
              for (int ix = g; ix < L + g; ix++)
                *o += lim[(((long)iz * S + iy) * S + ix) * V + v] * lim2uh[(iz - g) * B + z] *
                      lim2uh[(iy - g) * B + y] * lim2uh[(ix - g) * B + x];
        }
}

/* ------------------------------------------------------------------------ */
/* Min/max searches and the DMP, LIM:114-368.  Observables are the first     */
/* nobs variables (the default mapDiscreteMaximumPrincipleObservables copy,   */
/* AbstractGenericADERDGSolverHeader.template:115-119).                       */
/* ------------------------------------------------------------------------ */
void ora_cell_local_min_max(int order, int V, int nobs, const double* uh2lim, const double* uh2lob,
                            const double* luh, double* mn, double* mx) {
  const int B = order + 1, L = 2 * order + 1;
  for (int v = 0; v < nobs; v++) {
    mn[v] = DBL_MAX;
    mx[v] = -DBL_MAX;
  }
  /* findCellLocalMinAndMax, LIM:210-245 */
  for (int n = 0; n < B * B * B; n++)
    for (int v = 0; v < nobs; v++) {
      mn[v] = std_min(mn[v], luh[n * V + v]);
      mx[v] = std_max(mx[v], luh[n * V + v]);
    }
  /* compareWithADERDGSolutionAtGaussLobattoNodes, LIM:254-305 */
  double val[64];
  for (int z = 0; z < B; z++)
    for (int y = 0; y < B; y++)
      for (int x = 0; x < B; x++) {
        for (int v = 0; v < V; v++) {
          val[v] = 0.0;
          for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
              for (int ix = 0; ix < B; ix++)
                val[v] += luh[((iz * B + iy) * B + ix) * V + v] * uh2lob[iz * B + z] * uh2lob[iy * B + y] *
                          uh2lob[ix * B + x];
        }
        for (int v = 0; v < nobs; v++) {
          mn[v] = std_min(mn[v], val[v]);
          mx[v] = std_max(mx[v], val[v]);
        }
      }
  /* compareWithADERDGSolutionAtFVSubcellCenters, LIM:311-368 */
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        for (int v = 0; v < V; v++) {
          val[v] = 0.0;
          for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
              for (int ix = 0; ix < B; ix++)
                val[v] += luh[((iz * B + iy) * B + ix) * V + v] * uh2lim[iz * L + z] * uh2lim[iy * L + y] *
                          uh2lim[ix * L + x];
        }
        for (int v = 0; v < nobs; v++) {
          mn[v] = std_min(mn[v], val[v]);
          mx[v] = std_max(mx[v], val[v]);
        }
      }
}

/* findCellLocalLimiterMinAndMax, LIM:167-205 (interior subcells). */
void ora_limiter_local_min_max(int order, int D, int nobs, int g, const double* lim, double* mn,
                               double* mx) {
  const int L = 2 * order + 1, S = L + 2 * g;
  for (int v = 0; v < nobs; v++) {
    mn[v] = DBL_MAX;
    mx[v] = -DBL_MAX;
  }
  for (int iz = g; iz < L + g; iz++)
    for (int iy = g; iy < L + g; iy++)
      for (int ix = g; ix < L + g; ix++) {
        const double* q = lim + (((long)iz * S + iy) * S + ix) * D;
        for (int v = 0; v < nobs; v++) {
          mn[v] = std_min(mn[v], q[v]);
          mx[v] = std_max(mx[v], q[v]);
        }
      }
}

/* discreteMaximumPrincipleAndMinAndMaxSearch, LIM:114-165.  bmin/bmax hold
 * 6 face blocks of nobs; faces 0 and 1 are overwritten with the local and
 * the boundary extrema.  Returns 1 when the DMP holds. */
int ora_dmp(int order, int V, int nobs, const double* uh2lim, const double* uh2lob, const double* luh,
            double relax, double scaling, double* bmin, double* bmax) {
  double lmin[64], lmax[64];
  ora_cell_local_min_max(order, V, nobs, uh2lim, uh2lob, luh, lmin, lmax);
  int ok = 1;
  for (int v = 0; v < nobs; v++) {
    double boundaryMin = bmin[v];
    for (int i = 1; i < 6; i++) boundaryMin = std_min(boundaryMin, bmin[i * nobs + v]);
    double boundaryMax = bmax[v];
    for (int i = 1; i < 6; i++) boundaryMax = std_max(boundaryMax, bmax[i * nobs + v]);
    double scaledDifference = (boundaryMax - boundaryMin) * scaling;
    scaledDifference = std_max(scaledDifference, relax);
    if ((lmin[v] < (boundaryMin - scaledDifference)) || (lmax[v] > (boundaryMax + scaledDifference))) ok = 0;
    bmin[v] = lmin[v];
    bmax[v] = lmax[v];
    bmin[v + nobs] = boundaryMin;
    bmax[v + nobs] = boundaryMax;
  }
  return ok;
}

/* Euler isPhysicallyAdmissible, Demonstrators/EulerADERDG/MyEulerSolver_ADERDG.cpp:147-160 */
int ora_euler_admissible(const double* mn) {
  if (mn[0] <= 0.0) return 0;
  if (mn[4] <= 0.0) return 0;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Finite volumes (Euler: flux only, V = D = 5)                              */
/* ------------------------------------------------------------------------ */

/* rusanov, RUS:98-156 (useFlux, no NCP).  Returns s_max. */
double ora_fv_rusanov(const ora_solver* s, double* fnL, double* fnR, const double* qL, const double* qR,
                      int d) {
  const int V = s->nvar;
  double sL[16] = {0.0}, sR[16] = {0.0};
  ora_eigenvalues(s, qL, d, sL);
  ora_eigenvalues(s, qR, d, sR);
  double s_max = -1.0;
  for (int i = 0; i < V; i++) s_max = std_max(fabs(sL[i]), s_max);
  for (int i = 0; i < V; i++) s_max = std_max(fabs(sR[i]), s_max);
  double FL3[3][16] = {{0.0}}, FR3[3][16] = {{0.0}};
  double* FL[3] = {FL3[0], FL3[1], FL3[2]};
  double* FR[3] = {FR3[0], FR3[1], FR3[2]};
  ora_flux(s, qL, FL);
  ora_flux(s, qR, FR);
  for (int i = 0; i < V; i++) {
    fnL[i] = 0.5 * s_max * (qL[i] - qR[i]);
    fnL[i] += 0.5 * (FL3[d][i] + FR3[d][i]);
    fnR[i] = fnL[i];
  }
  return s_max;
}

static inline double minmod(double a, double b) { /* MH:27-40 */
  if (a * b < 0.0) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

/* musclhancock::c::solutionUpdate, MH:49-369, for a PDE with flux only.
 * luh_new must hold a copy of luh on entry (FVS:669-671).  Patch (S,S,S,D),
 * S = L + 2g; dx is the DG cell size.  Returns dt_max_allowed. */
double ora_muscl_update(const ora_solver* s, int L, int g, double* luh_new, const double* luh,
                        const double* dx, double dt) {
  const int V = s->nvar, D = s->nvar + s->npar;
  const int pb = g, pe = g + L, S = L + 2 * g;
  const double cflFactor = 0.99; /* musclhancock.h:29 */
  const double cellSize[3] = {dx[0] / L, dx[1] / L, dx[2] / L};
  const double invCellSize[3] = {1. / cellSize[0], 1. / cellSize[1], 1. / cellSize[2]};
  const double invCellSizeTimesDt[3] = {dt * invCellSize[0], dt * invCellSize[1], dt * invCellSize[2]};
  const long nvp = (long)S * S * S * V;
#define IDX(i, j, k, l) ((((long)(i) * S + (j)) * S + (k)) * D + (l))
#define IDW(i, j, k, l) ((((long)(i) * S + (j)) * S + (k)) * V + (l))
#define IDS(i, j, k, d, l) (((((long)(i) * S + (j)) * S + (k)) * 3 + (d)) * V + (l))
  double* slope = (double*)calloc(nvp * 3, sizeof(double));
  double* w = (double*)calloc(nvp * 6, sizeof(double));
  double* wLx = w;
  double* wLy = w + nvp;
  double* wLz = w + 2 * nvp;
  double* wRx = w + 3 * nvp;
  double* wRy = w + 4 * nvp;
  double* wRz = w + 5 * nvp;
  double* slopet = (double*)calloc(nvp, sizeof(double));
  for (int i = pb - 1; i < pe + 1; i++)
    for (int j = pb - 1; j < pe + 1; j++)
      for (int k = pb - 1; k < pe + 1; k++)
        for (int l = 0; l < V; l++) {
          slope[IDS(i, j, k, 0, l)] =
              minmod(luh[IDX(i, j, k + 1, l)] - luh[IDX(i, j, k, l)], luh[IDX(i, j, k, l)] - luh[IDX(i, j, k - 1, l)]) *
              invCellSize[0];
          slope[IDS(i, j, k, 1, l)] =
              minmod(luh[IDX(i, j + 1, k, l)] - luh[IDX(i, j, k, l)], luh[IDX(i, j, k, l)] - luh[IDX(i, j - 1, k, l)]) *
              invCellSize[1];
          slope[IDS(i, j, k, 2, l)] =
              minmod(luh[IDX(i + 1, j, k, l)] - luh[IDX(i, j, k, l)], luh[IDX(i, j, k, l)] - luh[IDX(i - 1, j, k, l)]) *
              invCellSize[2];
        }
  for (int i = pb - 1; i < pe + 1; i++)
    for (int j = pb - 1; j < pe + 1; j++)
      for (int k = pb - 1; k < pe + 1; k++)
        for (int l = 0; l < V; l++) {
          const double u = luh[IDX(i, j, k, l)];
          wLx[IDW(i, j, k, l)] = u - 0.5 * cellSize[0] * slope[IDS(i, j, k, 0, l)];
          wRx[IDW(i, j, k, l)] = u + 0.5 * cellSize[0] * slope[IDS(i, j, k, 0, l)];
          wLy[IDW(i, j, k, l)] = u - 0.5 * cellSize[1] * slope[IDS(i, j, k, 1, l)];
          wRy[IDW(i, j, k, l)] = u + 0.5 * cellSize[1] * slope[IDS(i, j, k, 1, l)];
          wLz[IDW(i, j, k, l)] = u - 0.5 * cellSize[2] * slope[IDS(i, j, k, 2, l)];
          wRz[IDW(i, j, k, l)] = u + 0.5 * cellSize[2] * slope[IDS(i, j, k, 2, l)];
        }
  double flx[3 * 16] = {0.0};
  double* F[3] = {flx, flx + 16, flx + 32};
  for (int i = pb - 1; i < pe + 1; i++)
    for (int j = pb - 1; j < pe + 1; j++)
      for (int k = pb - 1; k < pe + 1; k++) {
        /* MH:181-196, with its (never true) patchEnd+1 terms kept */
        if ((j == pb - 1 && k == pb - 1) || (j == pb - 1 && k == pe + 1) || (j == pe + 1 && k == pb - 1) ||
            (j == pe + 1 && k == pe + 1) || (j == pb - 1 && i == pb - 1) || (j == pb - 1 && i == pe + 1) ||
            (j == pe + 1 && i == pb - 1) || (j == pe + 1 && i == pe + 1) || (i == pb - 1 && k == pb - 1) ||
            (i == pb - 1 && k == pe + 1) || (i == pe + 1 && k == pb - 1) || (i == pe + 1 && k == pe + 1))
          continue;
        /* slopet is indexed with the data stride (MH:200); D == V here */
        double* Qt = slopet + IDX(i, j, k, 0);
        ora_flux(s, wLx + IDW(i, j, k, 0), F);
        for (int l = 0; l < V; l++) Qt[l] += F[0][l] * invCellSize[0];
        ora_flux(s, wRx + IDW(i, j, k, 0), F);
        for (int l = 0; l < V; l++) Qt[l] -= F[0][l] * invCellSize[0];
        ora_flux(s, wLy + IDW(i, j, k, 0), F);
        for (int l = 0; l < V; l++) Qt[l] += F[1][l] * invCellSize[1];
        ora_flux(s, wRy + IDW(i, j, k, 0), F);
        for (int l = 0; l < V; l++) Qt[l] -= F[1][l] * invCellSize[1];
        ora_flux(s, wLz + IDW(i, j, k, 0), F);
        for (int l = 0; l < V; l++) Qt[l] += F[2][l] * invCellSize[2];
        ora_flux(s, wRz + IDW(i, j, k, 0), F);
        for (int l = 0; l < V; l++) Qt[l] -= F[2][l] * invCellSize[2];
        for (int l = 0; l < V; l++) {
          wLx[IDW(i, j, k, l)] += 0.5 * dt * Qt[l];
          wRx[IDW(i, j, k, l)] += 0.5 * dt * Qt[l];
          wLy[IDW(i, j, k, l)] += 0.5 * dt * Qt[l];
          wRy[IDW(i, j, k, l)] += 0.5 * dt * Qt[l];
          wLz[IDW(i, j, k, l)] += 0.5 * dt * Qt[l];
          wRz[IDW(i, j, k, l)] += 0.5 * dt * Qt[l];
        }
      }
  double fL[16], fR[16];
  double dt_max_allowed = DBL_MAX;
  for (int i = pb; i < pe; i++)
    for (int j = pb; j < pe; j++)
      for (int k = pb - 1; k < pe; k++) {
        const double smax = ora_fv_rusanov(s, fL, fR, wRx + IDW(i, j, k, 0), wLx + IDW(i, j, k + 1, 0), 0);
        dt_max_allowed = std_min(dt_max_allowed, cflFactor / 3 * cellSize[0] / smax);
        for (int l = 0; l < V; ++l) {
          luh_new[IDX(i, j, k, l)] -= invCellSizeTimesDt[0] * fL[l];
          luh_new[IDX(i, j, k + 1, l)] += invCellSizeTimesDt[0] * fR[l];
        }
      }
  for (int i = pb; i < pe; i++)
    for (int j = pb - 1; j < pe; j++)
      for (int k = pb; k < pe; k++) {
        const double smax = ora_fv_rusanov(s, fL, fR, wRy + IDW(i, j, k, 0), wLy + IDW(i, j + 1, k, 0), 1);
        dt_max_allowed = std_min(dt_max_allowed, cflFactor / 3 * cellSize[1] / smax);
        for (int l = 0; l < V; ++l) {
          luh_new[IDX(i, j, k, l)] -= invCellSizeTimesDt[1] * fL[l];
          luh_new[IDX(i, j + 1, k, l)] += invCellSizeTimesDt[1] * fR[l];
        }
      }
  for (int i = pb - 1; i < pe; i++)
    for (int j = pb; j < pe; j++)
      for (int k = pb; k < pe; k++) {
        const double smax = ora_fv_rusanov(s, fL, fR, wRz + IDW(i, j, k, 0), wLz + IDW(i + 1, j, k, 0), 2);
        dt_max_allowed = std_min(dt_max_allowed, cflFactor / 3 * cellSize[2] / smax);
        for (int l = 0; l < V; ++l) {
          luh_new[IDX(i, j, k, l)] -= invCellSizeTimesDt[2] * fL[l];
          luh_new[IDX(i + 1, j, k, l)] += invCellSizeTimesDt[2] * fR[l];
        }
      }
#undef IDS
  free(slope);
  free(w);
  free(slopet);
  return dt_max_allowed;
}

/* ghostLayerFilling, FVC:874-980.  (dir, side) is the neighbour's position
 * relative to the cell: side -1 or +1 along axis dir (0 = x, 1 = y, 2 = z). */
void ora_fv_ghost_fill(int L, int g, int D, double* luh, const double* nb, int dir, int side) {
  const int pb = g, pe = g + L, S = L + 2 * g;
  for (int j = pb; j < pe; ++j)
    for (int i = pb; i < pe; ++i)
      for (int k = 0; k < g; ++k) {
        long dst, src;
        const int gd = side < 0 ? k : pe + k;           /* ghost layer */
        const int sd = side < 0 ? pe - g + k : pb + k;  /* neighbour interior layer */
        if (dir == 0) {
          dst = IDX(j, i, gd, 0);
          src = IDX(j, i, sd, 0);
        } else if (dir == 1) {
          dst = IDX(j, gd, i, 0);
          src = IDX(j, sd, i, 0);
        } else {
          dst = IDX(gd, j, i, 0);
          src = IDX(sd, j, i, 0);
        }
        memcpy(luh + dst, nb + src, sizeof(double) * D);
      }
}

/* Transmissive FV boundary on one face of a patch: boundaryLayerExtraction
 * (FVC:499-609), boundaryConditions with boundaryValues = copy (FVC:723-871,
 * mirror K = g-k-1), ghostLayerFillingAtBoundary (FVC:611-721).  The
 * extraction at the lower y and z faces reads layer patchBegin for every k
 * (no +k, FVC:558,588), which this restatement keeps. */
void ora_fv_boundary_transmissive(int L, int g, int D, double* luh, int face) {
  const int pb = g, pe = g + L, S = L + 2 * g;
  const int dir = face / 2, side = (face % 2) ? +1 : -1;
  double* bnd = (double*)malloc(sizeof(double) * L * L * g * D);
#define IB(a, b, k) ((((long)(a) * L + (b)) * g + (k)) * D)
  /* extraction: bnd(a,b,k) = interior layer */
  for (int j = pb; j < pe; ++j)
    for (int i = pb; i < pe; ++i)
      for (int k = 0; k < g; k++) {
        int layer;
        if (side < 0)
          layer = (dir == 0) ? pb + k : pb; /* lower y/z faces omit +k */
        else
          layer = pe - g + k;
        long src;
        if (dir == 0)
          src = IDX(j, i, layer, 0);
        else if (dir == 1)
          src = IDX(j, layer, i, 0);
        else
          src = IDX(layer, j, i, 0);
        memcpy(bnd + IB(j - pb, i - pb, k), luh + src, sizeof(double) * D);
      }
  /* boundary values (copy) with the mirror index, then filling */
  for (int j = pb; j < pe; ++j)
    for (int i = pb; i < pe; ++i)
      for (int k = 0; k < g; ++k) {
        const int K = g - k - 1;
        const double* in = bnd + IB(j - pb, i - pb, K);
        const int gl = side < 0 ? k : pe + k;
        long dst;
        if (dir == 0)
          dst = IDX(j, i, gl, 0);
        else if (dir == 1)
          dst = IDX(j, gl, i, 0);
        else
          dst = IDX(gl, j, i, 0);
        memcpy(luh + dst, in, sizeof(double) * D);
      }
#undef IB
  free(bnd);
}
#undef IDX
#undef IDW

/* FV stableTimeStepSize, FVC:409-460 (cflFactor = the FV solver's CFL). */
double ora_fv_stable_dt(const ora_solver* s, int L, int g, const double* luh, const double* dx, double cfl) {
  const int V = s->nvar, D = s->nvar + s->npar, S = L + 2 * g;
  double lambda[16] = {0.0};
  double dt = DBL_MAX;
  for (int i = g; i < g + L; i++)
    for (int j = g; j < g + L; j++)
      for (int k = g; k < g + L; k++)
        for (int d = 0; d < 3; d++) {
          ora_eigenvalues(s, luh + (((long)i * S + j) * S + k) * D, d, lambda);
          double smax = fabs(lambda[0]);
          for (int v = 1; v < V; v++) smax = std_max(smax, fabs(lambda[v])); /* max_element of |lambda| */
          dt = std_min(dt, cfl / 3 * (dx[d] / L) / smax);
        }
  return dt;
}

/* ------------------------------------------------------------------------ */
/* Limiting grid driver                                                      */
/* ------------------------------------------------------------------------ */
/* Restates, for one uniform level without mesh refinement, the non-fused    */
/* time step of Runner::runOneTimeStepWithThreeSeparateAlgorithmicSteps      */
/* (Runner.cpp:1116-1162) with a LimitingADERDGSolver:                        */
/*  1 Merging: per interior face mergeNeighboursBasedOnLimiterStatus          */
/*    (LS:1560-1628), mergeSolutionMinMaxOnFace (LS:1654-1690) and            */
/*    mergeNeighboursLimiterStatus (AS:2544-2590); per boundary face          */
/*    mergeWithBoundaryDataBasedOnLimiterStatus (LS:1708-1740).               */
/*  2 SolutionUpdate: updateSolution (LS:843-907), startNewTimeStep,          */
/*    updateLimiterStatusAndMinAndMaxAfterSolutionUpdate (LS:909-994).         */
/*  3 On an irregular limiter-domain change (Runner.cpp:866-946):             */
/*    T = 2h+1 status spreading sweeps (LimiterStatusSpreading.cpp,           */
/*    LS:392-405), reinitialiseSolversLocally (LS:1343-1381), a merging       */
/*    sweep with isRecomputation (LocalRecomputation.cpp:385-415), then per   */
/*    cell recomputeSolutionLocally (LS:1396-1436), startNewTimeStep and      */
/*    determineMinAndMax (LocalRecomputation.cpp:236-285).                    */
/*  4 Prediction for cells with status < T (Prediction.cpp:171-181).          */
/* Limiter patches are dense (every cell owns one); "allocation" zeroes a    */
/* patch (FVS:477-481) and is tracked per cell, as the reference's heap does. */
/* ------------------------------------------------------------------------ */
enum { ORA_REGULAR = 0, ORA_IRREGULAR = 1 };

struct ora_lgrid {
  ora_solver s;
  int nx, ny, nz, periodic;
  long ncells;
  double dx[3], lo[3];
  double t, dt;
  int h, T, A, cure, nobs, L, g, S, fv_scheme;
  double relax, scaling;
  long n_luh, n_lim, n_qb, n_fb, n_du;
  double *luh, *luh_prev, *lduh, *lQhbnd, *lFhbnd, *lim, *lim_prev, *mn, *mx, *dtc;
  int *st, *pst, *fs, *itc, *has_patch, *iters;
  double uh2lim[19 * 10], uh2lob[100], lim2uh[19 * 10];
  int have_prediction;
  int last_irregular;
  /* fused time stepping: dt is the corrector time step size, dtp the
   * predictor's; violated = the last stability check failed */
  int fused, violated;
  double weight, dtp;
  long reruns;
};

#define LG_LUH(g, c) ((g)->luh + (c) * (g)->n_luh)
#define LG_LIM(g, c) ((g)->lim + (c) * (g)->n_lim)

static long lg_cell(const ora_lgrid* g, int i, int j, int k) {
  if (g->periodic) {
    i = (i + g->nx) % g->nx;
    j = (j + g->ny) % g->ny;
    k = (k + g->nz) % g->nz;
  } else if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz) {
    return -1;
  }
  return ((long)k * g->ny + j) * g->nx + i;
}

ora_lgrid* ora_lgrid_create(const ora_solver* s, int nx, int ny, int nz, const double* lo, const double* hi,
                            int periodic, int helper_layers, int steps_till_cured, double relax,
                            double scaling, int fv_scheme) {
  ora_lgrid* g = (ora_lgrid*)calloc(1, sizeof(ora_lgrid));
  g->s = *s;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->periodic = periodic;
  g->ncells = (long)nx * ny * nz;
  for (int d = 0; d < 3; d++) g->lo[d] = lo[d];
  g->dx[0] = (hi[0] - lo[0]) / nx;
  g->dx[1] = (hi[1] - lo[1]) / ny;
  g->dx[2] = (hi[2] - lo[2]) / nz;
  g->h = helper_layers;
  g->T = 2 * helper_layers + 1; /* AS:461-462 */
  g->A = helper_layers + 1;
  g->cure = steps_till_cured;
  g->relax = relax;
  g->scaling = scaling;
  g->nobs = s->nvar;
  g->L = 2 * s->order + 1;
  g->fv_scheme = fv_scheme;
  g->g = fv_scheme == ORA_FV_GODUNOV ? 1 : 2;
  g->S = g->L + 2 * g->g;
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  g->n_luh = B * B * B * D;
  g->n_du = B * B * B * V;
  g->n_qb = 6 * B * B * D;
  g->n_fb = 6 * B * B * V;
  g->n_lim = (long)g->S * g->S * g->S * D;
  g->luh = (double*)calloc(g->ncells * g->n_luh, sizeof(double));
  g->luh_prev = (double*)calloc(g->ncells * g->n_luh, sizeof(double));
  g->lduh = (double*)calloc(g->ncells * g->n_du, sizeof(double));
  g->lQhbnd = (double*)calloc(g->ncells * g->n_qb, sizeof(double));
  g->lFhbnd = (double*)calloc(g->ncells * g->n_fb, sizeof(double));
  g->lim = (double*)calloc(g->ncells * g->n_lim, sizeof(double));
  g->lim_prev = (double*)calloc(g->ncells * g->n_lim, sizeof(double));
  g->mn = (double*)calloc(g->ncells * 6 * g->nobs, sizeof(double));
  g->mx = (double*)calloc(g->ncells * 6 * g->nobs, sizeof(double));
  g->dtc = (double*)calloc(g->ncells, sizeof(double));
  g->st = (int*)calloc(g->ncells, sizeof(int));
  g->pst = (int*)calloc(g->ncells, sizeof(int));
  g->fs = (int*)calloc(g->ncells * 6, sizeof(int));
  g->itc = (int*)calloc(g->ncells, sizeof(int));
  g->has_patch = (int*)calloc(g->ncells, sizeof(int));
  g->iters = (int*)calloc(g->ncells, sizeof(int));
  ora_limiter_tables(s->order, g->uh2lim, g->uh2lob, g->lim2uh);
  return g;
}

void ora_lgrid_destroy(ora_lgrid* g) {
  if (!g) return;
  free(g->luh);
  free(g->luh_prev);
  free(g->lduh);
  free(g->lQhbnd);
  free(g->lFhbnd);
  free(g->lim);
  free(g->lim_prev);
  free(g->mn);
  free(g->mx);
  free(g->dtc);
  free(g->st);
  free(g->pst);
  free(g->fs);
  free(g->itc);
  free(g->has_patch);
  free(g->iters);
  free(g);
}

long ora_lgrid_num_cells(const ora_lgrid* g) { return g->ncells; }
double ora_lgrid_time(const ora_lgrid* g) { return g->t; }
double ora_lgrid_dt(const ora_lgrid* g) { return g->dt; }
int ora_lgrid_last_irregular(const ora_lgrid* g) { return g->last_irregular; }

void ora_lgrid_get_solution(const ora_lgrid* g, double* luh) {
  memcpy(luh, g->luh, sizeof(double) * g->ncells * g->n_luh);
}
void ora_lgrid_get_status(const ora_lgrid* g, int* status) { memcpy(status, g->st, sizeof(int) * g->ncells); }
void ora_lgrid_get_limiter_solution(const ora_lgrid* g, long c, double* lim) {
  memcpy(lim, LG_LIM(g, c), sizeof(double) * g->n_lim);
}
void ora_lgrid_get_min_max(const ora_lgrid* g, double* mn, double* mx) {
  memcpy(mn, g->mn, sizeof(double) * g->ncells * 6 * g->nobs);
  memcpy(mx, g->mx, sizeof(double) * g->ncells * 6 * g->nobs);
}

/* ensureRequiredLimiterPatchIsAllocated (LS:1254-1271): returns 1 when a
 * (zeroed) patch was allocated now. */
static int lg_ensure_patch(ora_lgrid* g, long c) {
  if (!g->has_patch[c] && g->st[c] > 0) {
    memset(LG_LIM(g, c), 0, sizeof(double) * g->n_lim);
    memset(g->lim_prev + c * g->n_lim, 0, sizeof(double) * g->n_lim);
    g->has_patch[c] = 1;
    return 1;
  }
  return 0;
}

/* ensureNoUnrequiredLimiterPatchIsAllocatedOnComputeCell (LS:1200-1212) */
static void lg_release_patch(ora_lgrid* g, long c) {
  if (g->has_patch[c] && g->st[c] == 0) g->has_patch[c] = 0;
}

static void lg_swap(double* a, double* b, long n) {
  for (long i = 0; i < n; i++) {
    const double t = a[i];
    a[i] = b[i];
    b[i] = t;
  }
}

static void lg_project_fv(ora_lgrid* g, long c) {
  ora_project_on_fv(g->s.order, g->s.nvar, g->g, g->uh2lim, LG_LUH(g, c), LG_LIM(g, c));
}
static void lg_project_dg(ora_lgrid* g, long c) {
  ora_project_on_dg(g->s.order, g->s.nvar, g->g, g->lim2uh, LG_LIM(g, c), LG_LUH(g, c));
}

/* determineSolverMinAndMax (LS:1089-1125) / determineLimiterMinAndMax
 * (LS:1127-1159): local extrema copied to all six face blocks. */
static void lg_solver_min_max(ora_lgrid* g, long c) {
  double* mn = g->mn + c * 6 * g->nobs;
  double* mx = g->mx + c * 6 * g->nobs;
  ora_cell_local_min_max(g->s.order, g->s.nvar, g->nobs, g->uh2lim, g->uh2lob, LG_LUH(g, c), mn, mx);
  for (int f = 1; f < 6; f++) {
    memcpy(mn + f * g->nobs, mn, sizeof(double) * g->nobs);
    memcpy(mx + f * g->nobs, mx, sizeof(double) * g->nobs);
  }
}
static void lg_limiter_min_max(ora_lgrid* g, long c) {
  double* mn = g->mn + c * 6 * g->nobs;
  double* mx = g->mx + c * 6 * g->nobs;
  ora_limiter_local_min_max(g->s.order, g->s.nvar + g->s.npar, g->nobs, g->g, LG_LIM(g, c), mn, mx);
  for (int f = 1; f < 6; f++) {
    memcpy(mn + f * g->nobs, mn, sizeof(double) * g->nobs);
    memcpy(mx + f * g->nobs, mx, sizeof(double) * g->nobs);
  }
}

/* ADERDGSolver::updateSolution (AS:2185-2248): previous solution, surface
 * integral, solution update (adjustSolution is a no-op for t > 0). */
static void lg_dg_update(ora_lgrid* g, long c) {
  memcpy(g->luh_prev + c * g->n_luh, LG_LUH(g, c), sizeof(double) * g->n_luh);
  ora_surface_integral_nonlinear(&g->s, g->lduh + c * g->n_du, g->lFhbnd + c * g->n_fb, g->dx);
  ora_solution_update(&g->s, LG_LUH(g, c), g->lduh + c * g->n_du, g->dt);
}

/* FiniteVolumesSolver::updateSolution (FVS:662-700) + projectOnDGSpace */
static void lg_fv_update(ora_lgrid* g, long c) {
  memcpy(g->lim_prev + c * g->n_lim, LG_LIM(g, c), sizeof(double) * g->n_lim);
  if (g->dt > 0) {
    if (g->fv_scheme == ORA_FV_GODUNOV) /* the FV solver's CFL only scales the unused dt_max_allowed */
      ora_godunov_update(&g->s, g->L, g->g, LG_LIM(g, c), g->lim_prev + c * g->n_lim, g->dx, g->dt, 0.9);
    else
      ora_muscl_update(&g->s, g->L, g->g, LG_LIM(g, c), g->lim_prev + c * g->n_lim, g->dx, g->dt);
  }
  lg_project_dg(g, c);
}

/* DG boundary face with transmissive boundaryValues (QOut = QIn, FOut = FIn;
 * boundaryConditions.cpph:24-85), then the Riemann solve of AS:2952-2960. */
static void lg_dg_boundary(ora_lgrid* g, long c, int face) {
  const long B = g->s.order + 1, D = g->s.nvar + g->s.npar, V = g->s.nvar;
  const long B2 = B * B;
  double* FIn = g->lFhbnd + c * g->n_fb + face * B2 * V;
  const double* QIn = g->lQhbnd + c * g->n_qb + face * B2 * D;
  double* FOut = (double*)malloc(sizeof(double) * B2 * V);
  double* QOut = (double*)malloc(sizeof(double) * B2 * D);
  memcpy(FOut, FIn, sizeof(double) * B2 * V);
  memcpy(QOut, QIn, sizeof(double) * B2 * D);
  if (face % 2 == 0)
    ora_riemann_solver_nonlinear(&g->s, FOut, FIn, QOut, QIn, g->dt, face / 2);
  else
    ora_riemann_solver_nonlinear(&g->s, FIn, FOut, QIn, QOut, g->dt, face / 2);
  free(FOut);
  free(QOut);
}

static void lg_dg_riemann(ora_lgrid* g, long cl, long cr, int d) {
  const long B = g->s.order + 1, D = g->s.nvar + g->s.npar, V = g->s.nvar;
  const long B2 = B * B;
  ora_riemann_solver_nonlinear(&g->s, g->lFhbnd + cl * g->n_fb + (2 * d + 1) * B2 * V,
                               g->lFhbnd + cr * g->n_fb + (2 * d) * B2 * V,
                               g->lQhbnd + cl * g->n_qb + (2 * d + 1) * B2 * D,
                               g->lQhbnd + cr * g->n_qb + (2 * d) * B2 * D, g->dt, d);
}

static void lg_fv_exchange(ora_lgrid* g, long cl, long cr, int d) {
  const int D = g->s.nvar + g->s.npar;
  ora_fv_ghost_fill(g->L, g->g, D, LG_LIM(g, cl), LG_LIM(g, cr), d, +1);
  ora_fv_ghost_fill(g->L, g->g, D, LG_LIM(g, cr), LG_LIM(g, cl), d, -1);
}

/* mergeWithLimiterStatus (AS:2547-2563) for both sides of a face */
static void lg_merge_status(ora_lgrid* g, long cl, long cr, int d) {
  const int T = g->T;
  const int sl = imin(g->st[cl], T), sr = imin(g->st[cr], T);
  g->fs[cl * 6 + 2 * d + 1] = imax(0, imax(sl, sr) - 1);
  g->fs[cr * 6 + 2 * d] = imax(0, imax(sr, sl) - 1);
}

static void lg_merge_min_max(ora_lgrid* g, long cl, long cr, int d) {
  const int n = g->nobs;
  double* mn1 = g->mn + (cl * 6 + 2 * d + 1) * n;
  double* mn2 = g->mn + (cr * 6 + 2 * d) * n;
  double* mx1 = g->mx + (cl * 6 + 2 * d + 1) * n;
  double* mx2 = g->mx + (cr * 6 + 2 * d) * n;
  for (int i = 0; i < n; i++) {
    const double lo = std_min(mn1[i], mn2[i]);
    const double hi = std_max(mx1[i], mx2[i]);
    mn1[i] = lo;
    mn2[i] = lo;
    mx1[i] = hi;
    mx2[i] = hi;
  }
}

static void lg_merging(ora_lgrid* g, int recomputation) {
  const int A = g->A, D = g->s.nvar + g->s.npar;
  /* within one direction the faces write disjoint face slots and ghost slabs */
  for (int d = 0; d < 3; d++)
#pragma omp parallel for schedule(static)
    for (long c = 0; c < g->ncells; c++) {
      const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
      const long r = lg_cell(g, i + (d == 0), j + (d == 1), k + (d == 2));
      if (r < 0) continue;
      const int dg_l = g->st[c] < A, dg_r = g->st[r] < A;
      if (dg_l && dg_r) {
        if (!recomputation) lg_dg_riemann(g, c, r, d);
      } else if (!dg_l && !dg_r) {
        lg_fv_exchange(g, c, r, d);
      } else {
        lg_fv_exchange(g, c, r, d);
        if (!recomputation) lg_dg_riemann(g, c, r, d);
      }
      if (!recomputation) {
        lg_merge_min_max(g, c, r, d);
        lg_merge_status(g, c, r, d);
      }
    }
  if (g->periodic) return;
  /* boundary faces */
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g->ncells; c++) {
    const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
    const int at[6] = {i == 0, i == g->nx - 1, j == 0, j == g->ny - 1, k == 0, k == g->nz - 1};
    for (int f = 0; f < 6; f++) {
      if (!at[f]) continue;
      if (g->st[c] < A) {
        if (!recomputation) lg_dg_boundary(g, c, f);
      } else {
        ora_fv_boundary_transmissive(g->L, g->g, D, LG_LIM(g, c), f);
      }
    }
  }
}

static double lg_dg_dt(ora_lgrid* g, long c) { return ora_stable_time_step_size(&g->s, LG_LUH(g, c), g->dx); }

/* Prediction (Prediction.cpp:177) of the cells with status < troubled
 * (select 0), or recomputePredictorLocally (LS:1438-1464) after an irregular
 * change in fused time stepping (select 1): the cells with status < troubled
 * whose solution was recomputed from an FV patch -- FV-active neighbours
 * (status >= A) and cells troubled in the previous step. */
static void lg_predict(ora_lgrid* g, double dt, int select) {
  const ora_solver* s = &g->s;
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B3 = B * B * B, B4 = B3 * B;
#pragma omp parallel
  {
    double* lQi = (double*)malloc(sizeof(double) * B4 * D);
    double* rhs = (double*)malloc(sizeof(double) * B4 * D);
    double* lFi = (double*)malloc(sizeof(double) * B4 * 4 * V);
    double* gradQ = (double*)malloc(sizeof(double) * B4 * 3 * V);
    double* lQhi = (double*)malloc(sizeof(double) * B3 * D);
    double* lFhi = (double*)malloc(sizeof(double) * B3 * 4 * V);
#pragma omp for schedule(dynamic, 16)
    for (long c = 0; c < g->ncells; c++) {
      if (select == 1 && !(g->st[c] < g->T && (g->st[c] >= g->A || g->pst[c] >= g->T) && dt > 0)) continue;
      if (g->st[c] >= g->T) {
        g->iters[c] = 0;
        continue;
      }
      g->iters[c] = ora_space_time_predictor_nonlinear(s, g->lQhbnd + c * g->n_qb, g->lFhbnd + c * g->n_fb, lQi, rhs,
                                                       lFi, gradQ, lQhi, lFhi, LG_LUH(g, c), g->dx, dt);
      ora_volume_integral_nonlinear(s, g->lduh + c * g->n_du, lFi, lFhi, g->dx);
    }
    free(lQi);
    free(rhs);
    free(lFi);
    free(gradQ);
    free(lQhi);
    free(lFhi);
  }
}

static void lg_spread(ora_lgrid* g) {
  for (int sweep = 0; sweep < g->T; sweep++) {
    for (int d = 0; d < 3; d++)
#pragma omp parallel for schedule(static)
      for (long c = 0; c < g->ncells; c++) {
        const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
        const long r = lg_cell(g, i + (d == 0), j + (d == 1), k + (d == 2));
        if (r >= 0) lg_merge_status(g, c, r, d);
      }
    /* updateLimiterStatusDuringLimiterStatusSpreading, LS:392-405 */
#pragma omp parallel for schedule(static)
    for (long c = 0; c < g->ncells; c++) {
      if (g->st[c] < g->T) {
        int m = 0;
        for (int f = 0; f < 6; f++) m = imax(m, g->fs[c * 6 + f]);
        g->st[c] = m;
      } else {
        g->st[c] = g->T;
        g->itc[c] = g->cure + 1;
      }
      for (int f = 0; f < 6; f++) g->fs[c * 6 + f] = 0;
      lg_ensure_patch(g, c);
    }
  }
}

/* Initial state: the DG initial condition luh at the nodes, then
 * updateLimiterStatusAndMinAndMaxAfterSetInitialConditions (LS:1031-1043);
 * initially troubled cells are spread and every cell with a limiter patch
 * gets the FV initial condition at its subcell centres (LS:746-762 calls the
 * FV solver's setInitialConditions, i.e. solutionAdjustment FVC:462-497):
 * lim_ic holds those values for every cell, [ncells][L^3][V+P] (z,y,x,var);
 * when lim_ic is NULL the patch gets the projection of luh instead.  Then
 * the first dt and the first prediction. */
void ora_lgrid_set_solution(ora_lgrid* g, const double* luh, const double* lim_ic) {
  memcpy(g->luh, luh, sizeof(double) * g->ncells * g->n_luh);
  memset(g->st, 0, sizeof(int) * g->ncells);
  memset(g->fs, 0, sizeof(int) * g->ncells * 6);
  memset(g->has_patch, 0, sizeof(int) * g->ncells);
  int any = 0;
#pragma omp parallel for schedule(static) reduction(| : any)
  for (long c = 0; c < g->ncells; c++) {
    lg_solver_min_max(g, c);
    g->itc[c] = 0;
    if (!ora_euler_admissible(g->mn + c * 6 * g->nobs)) {
      g->itc[c] = g->cure + 1;
      g->st[c] = g->T;
      any = 1;
    }
  }
  if (any) {
    lg_spread(g);
    const int L = g->L, G = g->g, S = g->S, D = g->s.nvar + g->s.npar;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < g->ncells; c++)
      if (g->st[c] > 0) {
        lg_ensure_patch(g, c);
        if (!lim_ic) {
          lg_project_fv(g, c);
          continue;
        }
        const double* src = lim_ic + c * (long)L * L * L * D;
        for (int z = 0; z < L; z++)
          for (int y = 0; y < L; y++)
            memcpy(LG_LIM(g, c) + (((long)(z + G) * S + (y + G)) * S + G) * D, src + ((long)z * L + y) * L * D,
                   sizeof(double) * L * D);
      }
  }
  memcpy(g->pst, g->st, sizeof(int) * g->ncells);
  g->t = 0.0;
  double dt = DBL_MAX;
#pragma omp parallel for schedule(static) reduction(min : dt)
  for (long c = 0; c < g->ncells; c++) dt = std_min(dt, lg_dg_dt(g, c));
  g->dt = dt;
  g->dtp = dt; /* fused: the first predictor and corrector dt, unweighted */
  g->violated = 0;
  g->reruns = 0;
  lg_predict(g, g->dt, 0);
  g->have_prediction = 1;
}

/* Solution update and limiter status of every cell (SolutionUpdate with
 * LS:843-937); returns 1 on an irregular limiter-domain change. */
static int lg_update_status(ora_lgrid* g) {
  const int T = g->T, A = g->A, nobs = g->nobs;
  int irregular = 0;
  /* per cell: reads its own face statuses, writes only its own data */
#pragma omp parallel for schedule(dynamic, 16) reduction(| : irregular)
  for (long c = 0; c < g->ncells; c++) {
    g->pst[c] = g->st[c];
    const int s = g->st[c];
    if (s == 0) {
      lg_dg_update(g, c);
    } else if (s < A) {
      lg_dg_update(g, c);
      lg_swap(LG_LIM(g, c), g->lim_prev + c * g->n_lim, g->n_lim);
      lg_project_fv(g, c);
    } else {
      lg_fv_update(g, c);
    }
    lg_release_patch(g, c);
    g->dtc[c] = lg_dg_dt(g, c);
    /* updateLimiterStatusAndMinAndMaxAfterSolutionUpdate */
    double* mn = g->mn + c * 6 * nobs;
    double* mx = g->mx + c * 6 * nobs;
    int valid = ora_dmp(g->s.order, g->s.nvar, nobs, g->uh2lim, g->uh2lob, LG_LUH(g, c), g->relax, g->scaling, mn, mx);
    if (valid) valid = ora_euler_admissible(mn);
    if (s >= A) lg_limiter_min_max(g, c);
    /* determineLimiterStatusAfterSolutionUpdate, LS:955-994 */
    if (!valid) {
      g->itc[c] = g->cure + 1;
      if (s <= A) irregular = 1;
      g->st[c] = T;
    } else if (g->pst[c] >= T) {
      g->st[c] = T;
      g->itc[c] -= 1;
      if (g->itc[c] == 0) {
        g->st[c] = T - 1;
        g->itc[c] = g->cure + 1;
      }
    } else {
      int m = 0;
      for (int f = 0; f < 6; f++) m = imax(m, g->fs[c * 6 + f]);
      g->st[c] = m;
    }
    /* allocateLimiterPatchAfterSolutionUpdate, LS:940-953 */
    if (lg_ensure_patch(g, c) && g->st[c] < A) lg_project_fv(g, c);
    for (int f = 0; f < 6; f++) g->fs[c * 6 + f] = 0;
  }
  return irregular;
}

/* Irregular change (Runner.cpp:866-946): status spreading, rollback
 * (reinitialiseSolversLocally), merging and local recomputation with the
 * new dt candidates of every cell. */
static void lg_irregular(ora_lgrid* g) {
  const int T = g->T, A = g->A;
  lg_spread(g);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g->ncells; c++) { /* reinitialiseSolversLocally */
    lg_ensure_patch(g, c);
    if (g->st[c] > 0) {
      if (g->pst[c] >= A) {
        lg_swap(LG_LIM(g, c), g->lim_prev + c * g->n_lim, g->n_lim);
        lg_project_dg(g, c);
      } else {
        lg_swap(LG_LUH(g, c), g->luh_prev + c * g->n_luh, g->n_luh);
        lg_project_fv(g, c);
      }
      if (g->st[c] >= T) g->itc[c] = g->cure + 1;
    }
    lg_release_patch(g, c);
  }
  lg_merging(g, 1);
#pragma omp parallel for schedule(dynamic, 16)
  for (long c = 0; c < g->ncells; c++) { /* recomputeSolutionLocally */
    const int s = g->st[c];
    if (s > 0) {
      if (s >= A) {
        lg_fv_update(g, c);
      } else if (g->pst[c] < A) {
        lg_swap(LG_LUH(g, c), g->luh_prev + c * g->n_luh, g->n_luh);
        lg_project_fv(g, c);
      } else {
        lg_swap(LG_LIM(g, c), g->lim_prev + c * g->n_lim, g->n_lim);
        lg_project_dg(g, c);
      }
    }
    g->dtc[c] = lg_dg_dt(g, c);
    if (s < A)
      lg_solver_min_max(g, c);
    else
      lg_limiter_min_max(g, c);
  }
}

static double lg_min_dt(const ora_lgrid* g) {
  double dt = DBL_MAX;
#pragma omp parallel for schedule(static) reduction(min : dt)
  for (long c = 0; c < g->ncells; c++) dt = std_min(dt, g->dtc[c]);
  return dt;
}

/* One time step; returns 1 when it had an irregular limiter-domain change. */
static int lg_step(ora_lgrid* g) {
  /* 1. merging */
  lg_merging(g, 0);
  /* 2. solution update */
  const int irregular = lg_update_status(g);
  /* 3. irregular change: spreading, rollback, local recomputation */
  if (irregular) lg_irregular(g);
  g->last_irregular = irregular;
  /* 4. new time step size and prediction */
  g->t += g->dt;
  g->dt = lg_min_dt(g);
  lg_predict(g, g->dt, 0);
  return irregular;
}

/* One fused time step of the limiting solver (LS:806-836 per cell, the
 * SolutionUpdate endIteration's stability check TimeStepSizeComputation.cpp:
 * 167-204, Runner.cpp:1053-1110), in a one-process run: merging, update and
 * status, prediction of the cells with status < troubled with the predictor
 * dt, the stability check of that dt against the new stable dt and
 * startNewTimeStep.  After an irregular change the spreading, rollback
 * (which takes the time stamps back one step) and local recomputation run
 * (updateMeshAndSubdomains, Runner.cpp:866-946); the recomputation re-predicts
 * the cells whose solution came from an FV patch with the predictor dt the
 * cells were rolled back to (the one this step predicted with), the check is
 * repeated with the recomputed dt candidates (LocalRecomputation.cpp:168-222),
 * and the adapter's TimeStepSizeComputation (updateTimeStepSizesFused,
 * ADERDGSolver.cpp:594-612) then sets the corrector and predictor dt to the
 * new stable dt.  A failed (last) check re-runs the prediction with the new
 * dt.  Returns 1 on an irregular change. */
static int lg_step_fused(ora_lgrid* g) {
  lg_merging(g, 0);
  const int irregular = lg_update_status(g);
  lg_predict(g, g->dtp, 0);
  const double stable = lg_min_dt(g);
  double next = stable;
  double used = g->dtp;
  if (used == 0.0) used = stable;
  int violated = used > stable;
  double dtp_checked = g->dtp;
  if (violated) {
    next = fmin(next, g->weight * stable);
    dtp_checked = g->weight * stable;
  } else {
    next = fmin(next, 0.5 * (used + g->weight * stable));
  }
  if (!irregular) {
    g->t += g->dt;
    g->dt = dtp_checked;
    g->dtp = next;
  } else {
    lg_irregular(g);
    lg_predict(g, g->dtp, 1);
    const double stable2 = lg_min_dt(g);
    double used2 = dtp_checked;
    if (used2 == 0.0) used2 = stable2;
    violated = used2 > stable2;
    g->t += g->dt;
    g->dt = stable2;
    g->dtp = stable2;
  }
  g->last_irregular = irregular;
  g->violated = violated;
  if (violated) {
    lg_predict(g, g->dt, 0);
    g->reruns++;
  }
  return irregular;
}

void ora_lgrid_set_fused(ora_lgrid* g, int on, double weight) {
  g->fused = on;
  g->weight = weight;
}

long ora_lgrid_predictor_reruns(const ora_lgrid* g) { return g->reruns; }

/* The cell's current prediction: its face blocks lQhbnd [6][B^2][D] and
 * lFhbnd [6][B^2][V] (test access). */
void ora_lgrid_get_prediction(const ora_lgrid* g, long c, double* lQhbnd, double* lFhbnd) {
  memcpy(lQhbnd, g->lQhbnd + c * g->n_qb, sizeof(double) * g->n_qb);
  memcpy(lFhbnd, g->lFhbnd + c * g->n_fb, sizeof(double) * g->n_fb);
}

/* Advance nsteps; status_out[nsteps*ncells] (may be NULL) receives the
 * limiter status after each step, dt_out[nsteps] the dt used by each step,
 * irregular_out[nsteps] whether the step recomputed. */
void ora_lgrid_step(ora_lgrid* g, int nsteps, int* status_out, double* dt_out, int* irregular_out) {
  for (int n = 0; n < nsteps; n++) {
    if (dt_out) dt_out[n] = g->dt;
    const int irr = g->fused ? lg_step_fused(g) : lg_step(g);
    if (irregular_out) irregular_out[n] = irr;
    if (status_out) memcpy(status_out + (long)n * g->ncells, g->st, sizeof(int) * g->ncells);
  }
}
