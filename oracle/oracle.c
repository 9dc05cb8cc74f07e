/*
 * oracle.c -- CPU restatement of the reference's generic 3-D ADER-DG path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the HIP
 * product never includes, links or calls this file.
 *
 * Each function below follows one reference function line for line (loop
 * order, accumulation order, expression trees); the citation is given at
 * the function.  Paths are relative to the reference checkout:
 *   K = ExaHyPE/kernels/aderdg/generic/c/3d
 * Build: see oracle/Makefile (-O3 -ffp-contract=off -fopenmp).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXA_TABLE static const
#include "../include/exahype_amd/device/dg_tables.inc"

#define W(o, i) exa_gl_weights[o][i]
#define FLC(o, i) exa_fl_coeff[o][i]
#define FRC(o, i) exa_fr_coeff[o][i]
#define KXI(o, i, j) exa_kxi[o][i][j]
#define IK1(o, i, j) exa_ik1[o][i][j]
#define DUDX(o, i, j) exa_dudx[o][i][j]

#define ORA_INLINE static inline __attribute__((always_inline))

double ora_table(int which, int o, int i, int j) {
  switch (which) {
    case ORA_T_WEIGHTS: return exa_gl_weights[o][i];
    case ORA_T_NODES: return exa_gl_nodes[o][i];
    case ORA_T_FLCOEFF: return exa_fl_coeff[o][i];
    case ORA_T_FRCOEFF: return exa_fr_coeff[o][i];
    case ORA_T_KXI: return exa_kxi[o][i][j];
    case ORA_T_IK1: return exa_ik1[o][i][j];
    case ORA_T_DUDX: return exa_dudx[o][i][j];
    case ORA_T_LOB_WEIGHTS: return exa_lob_weights[o][i];
    case ORA_T_LOB_NODES: return exa_lob_nodes[o][i];
  }
  return 0.0;
}

/* ------------------------------------------------------------------------ */
/* PDEs                                                                       */
/* ------------------------------------------------------------------------ */

/* Compressible Euler, gamma = 1.4, as the reference's kernel test solver
 * (ExaHyPE/exahype/tests/kernels/c/GenericEulerKernelTest3D.cpp:26-57).
 * The demonstrator MyEulerSolver_ADERDG.cpp:102-135 computes the same
 * quantities with a different association order (rounding-level only). */
ORA_INLINE void euler_flux(const double* Q, double** F) {
  double* f = F[0];
  double* g = F[1];
  double* h = F[2];
  const double GAMMA = 1.4;
  const double irho = 1.0 / Q[0];
  const double p = (GAMMA - 1) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
  f[0] = Q[1];
  f[1] = irho * Q[1] * Q[1] + p;
  f[2] = irho * Q[1] * Q[2];
  f[3] = irho * Q[1] * Q[3];
  f[4] = irho * Q[1] * (Q[4] + p);
  g[0] = Q[2];
  g[1] = irho * Q[2] * Q[1];
  g[2] = irho * Q[2] * Q[2] + p;
  g[3] = irho * Q[2] * Q[3];
  g[4] = irho * Q[2] * (Q[4] + p);
  h[0] = Q[3];
  h[1] = irho * Q[3] * Q[1];
  h[2] = irho * Q[3] * Q[2];
  h[3] = irho * Q[3] * Q[3] + p;
  h[4] = irho * Q[3] * (Q[4] + p);
}

/* GenericEulerKernelTest3D.cpp:68-82 */
ORA_INLINE void euler_eigenvalues(const double* Q, int d, double* lambda) {
  const double GAMMA = 1.4;
  double irho = 1.0 / Q[0];
  double p = (GAMMA - 1) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
  double u_n = Q[d + 1] * irho;
  double c = sqrt(GAMMA * p * irho);
  lambda[0] = u_n - c;
  lambda[1] = u_n;
  lambda[2] = u_n;
  lambda[3] = u_n;
  lambda[4] = u_n + c;
}

/* Linear elasticity, velocity-stress form, 9 variables (sxx, syy, szz, sxy,
 * syz, sxz, u, v, w) and 3 parameters (lambda, mu, rho) in the layout of the
 * reference's elasticity test solver (ExaHyPE/exahype/tests/kernels/c/
 * ElasticityKernelTest2D.cpp:101-182): eigenvalues -cp,-cs,-cs,0,0,0,cs,cs,cp
 * (:101-129) and the non-conservative product B(Q) grad Q, whose x- and
 * y-blocks are :150-170 and whose z-block is the one written out there as a
 * comment (:172-180).  The nonlinear kernels take the sum over directions. */
ORA_INLINE void elastic_eigenvalues(const double* Q, double* lambda) {
  const double lam = Q[9], mu = Q[10], rho0 = Q[11];
  const double cp = sqrt((lam + 2 * mu) / rho0);
  const double cs = sqrt(mu / rho0);
  lambda[0] = -cp;
  lambda[1] = -cs;
  lambda[2] = -cs;
  lambda[3] = 0.0;
  lambda[4] = 0.0;
  lambda[5] = 0.0;
  lambda[6] = +cs;
  lambda[7] = +cs;
  lambda[8] = +cp;
}

ORA_INLINE void elastic_ncp(const double* Q, const double* gradQ, double* BgradQ) {
  const int nv = 9;
  const double lam = Q[nv], mu = Q[nv + 1], irho = 1.0 / Q[nv + 2];
  const double* gx = gradQ;
  const double* gy = gradQ + nv;
  const double* gz = gradQ + 2 * nv;
  BgradQ[0] = (-(lam + 2 * mu) * gx[6] + -lam * gy[7]) + -lam * gz[8];
  BgradQ[1] = (-lam * gx[6] + -(lam + 2 * mu) * gy[7]) + -lam * gz[8];
  BgradQ[2] = (-lam * gx[6] + -lam * gy[7]) + -(lam + 2 * mu) * gz[8];
  BgradQ[3] = (-mu * gx[7] + -mu * gy[6]) + 0.0;
  BgradQ[4] = (0.0 + -mu * gy[8]) + -mu * gz[7];
  BgradQ[5] = (-mu * gx[8] + 0.0) + -mu * gz[6];
  BgradQ[6] = (-irho * gx[0] + -irho * gy[3]) + -irho * gz[5];
  BgradQ[7] = (-irho * gx[3] + -irho * gy[1]) + -irho * gz[4];
  BgradQ[8] = (-irho * gx[5] + -irho * gy[4]) + -irho * gz[2];
}

void ora_flux(const ora_solver* s, const double* Q, double** F) {
  if (s->pde == ORA_PDE_ELASTIC) {
    for (int d = 0; d < 3; d++)
      for (int m = 0; m < s->nvar; m++) F[d][m] = 0.0;  /* NCP form: no flux */
    return;
  }
  euler_flux(Q, F);
}

void ora_eigenvalues(const ora_solver* s, const double* Q, int d, double* lambda) {
  if (s->pde == ORA_PDE_ELASTIC) {
    elastic_eigenvalues(Q, lambda);
    return;
  }
  euler_eigenvalues(Q, d, lambda);
}

/* GenericEulerKernelTest3D.cpp:59-66 and the elasticity test solver
 * (ElasticityKernelTest2D.cpp:95-99): zero sources. */
void ora_algebraic_source(const ora_solver* s, const double* Q, double* S) {
  (void)Q;
  for (int m = 0; m < s->nvar; m++) S[m] = 0.0;
}

/* Euler: GenericEulerKernelTest3D.cpp:84-107 with _setNcpAndMatrixBToZero
 * set, the state the predictor test runs it in (BgradQ = 0).  Elastic: above. */
void ora_ncp(const ora_solver* s, const double* Q, const double* gradQ, double* BgradQ) {
  if (s->pde == ORA_PDE_ELASTIC) {
    elastic_ncp(Q, gradQ, BgradQ);
    return;
  }
  (void)Q;
  (void)gradQ;
  for (int m = 0; m < s->nvar; m++) BgradQ[m] = 0.0;
}

/* ------------------------------------------------------------------------ */
/* Space-time predictor (K/spaceTimePredictorNonlinear.cpph)                 */
/* ------------------------------------------------------------------------ */

/* aderPicardLoopNonlinear, K/spaceTimePredictorNonlinear.cpph:42-329 */
ORA_INLINE int picard_loop(const ora_solver* s, const int order, const int V, const int P,
                           const double* luh, const double dt, const double* dx, double* lQi,
                           double* rhs, double* lFi, double* gradQ) {
  const int D = V + P;
  const int B = order + 1;
  const int B2 = B * B;
  const int B4 = B2 * B2;
  const int useFlux = s->use_flux, useSource = s->use_source, useNCP = s->use_ncp;
  const double invDx[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
#define I_LUH(i, j, k, m) ((((i) * B + (j)) * B + (k)) * D + (m))
#define I_LQI(i, j, k, l, m) (((((i) * B + (j)) * B + (k)) * B + (l)) * D + (m))
#define I_LFI(i, j, k, l, d, m) ((((((i) * B + (j)) * B + (k)) * B + (l)) * 4 + (d)) * V + (m))
#define I_RHS(i, j, k, l, m) (((((i) * B + (j)) * B + (k)) * B + (l)) * V + (m))
#define I_GQ(i, j, k, l, d, m) ((((((i) * B + (j)) * B + (k)) * B + (l)) * 3 + (d)) * V + (m))

  /* 1. trivial initial guess (:68-82) */
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          memcpy(lQi + I_LQI(i, j, k, l, 0), luh + I_LUH(i, j, k, 0), sizeof(double) * D);

  const int MaxIterations = 2 * (order + 1); /* :85 */
  int iter;
  for (iter = 0; iter < MaxIterations; iter++) {
    for (int i = 0; i < B; i++) { /* time DOF */
      if (useFlux) {            /* :96-122 */
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++)
            for (int l = 0; l < B; l++) {
              double* F[3];
              F[0] = lFi + I_LFI(i, j, k, l, 0, 0);
              F[1] = lFi + I_LFI(i, j, k, l, 1, 0);
              F[2] = lFi + I_LFI(i, j, k, l, 2, 0);
              ora_flux(s, lQi + I_LQI(j, k, l, i, 0), F);
            }
      }
      /* rhs0 (:124-141) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++) {
            const double weight = W(order, j) * W(order, k) * W(order, l);
            for (int m = 0; m < V; m++)
              rhs[I_RHS(i, j, k, l, m)] = weight * FLC(order, i) * luh[I_LUH(j, k, l, m)];
          }
      if (useNCP) memset(gradQ, 0, sizeof(double) * B4 * 3 * V); /* :145-147 */
      /* x direction (:152-178) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = W(order, i) * W(order, j) * W(order, k);
          const double updateSize = weight * dt * invDx[0];
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++)
              for (int n = 0; n < B; n++) {
                if (useFlux)
                  rhs[I_RHS(i, j, k, l, m)] -= updateSize * lFi[I_LFI(i, j, k, n, 0, m)] * KXI(order, n, l);
                if (useNCP)
                  gradQ[I_GQ(j, k, l, i, 0, m)] += 1.0 * invDx[0] * lQi[I_LQI(j, k, n, i, m)] * DUDX(order, l, n);
              }
        }
      /* y direction (:181-207) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = W(order, i) * W(order, j) * W(order, k);
          const double updateSize = weight * dt * invDx[1];
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++)
              for (int n = 0; n < B; n++) {
                if (useFlux)
                  rhs[I_RHS(i, j, l, k, m)] -= updateSize * lFi[I_LFI(i, j, n, k, 1, m)] * KXI(order, n, l);
                if (useNCP)
                  gradQ[I_GQ(j, l, k, i, 1, m)] += 1.0 * invDx[1] * lQi[I_LQI(j, n, k, i, m)] * DUDX(order, l, n);
              }
        }
      /* z direction (:210-236) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = W(order, i) * W(order, j) * W(order, k);
          const double updateSize = weight * dt * invDx[2];
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++)
              for (int n = 0; n < B; n++) {
                if (useFlux)
                  rhs[I_RHS(i, l, j, k, m)] -= updateSize * lFi[I_LFI(i, n, j, k, 2, m)] * KXI(order, n, l);
                if (useNCP)
                  gradQ[I_GQ(l, j, k, i, 2, m)] += 1.0 * invDx[2] * lQi[I_LQI(n, j, k, i, m)] * DUDX(order, l, n);
              }
        }
      /* source and non-conservative product (:238-276) */
      if (useSource || useNCP) {
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++)
            for (int l = 0; l < B; l++) {
              const double weight = W(order, i) * W(order, j) * W(order, k) * W(order, l);
              const double updateSize = weight * dt;
              double* S = lFi + I_LFI(i, j, k, l, 3, 0);
              if (useSource) {
                ora_algebraic_source(s, lQi + I_LQI(j, k, l, i, 0), S);
              } else {
                for (int m = 0; m < V; m++) S[m] = 0.0;
              }
              if (useNCP) {
                double ncp[64];
                ora_ncp(s, lQi + I_LQI(j, k, l, i, 0), gradQ + I_GQ(j, k, l, i, 0, 0), ncp);
                for (int m = 0; m < V; m++) S[m] -= ncp[m];
              }
              for (int m = 0; m < V; m++) rhs[I_RHS(i, j, k, l, m)] += updateSize * S[m];
            }
      }
    } /* time DOF */

    /* 4. multiply with iK1 (:280-311) */
    double sq_res = 0.0;
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = W(order, i) * W(order, j) * W(order, k);
          const double iweight = 1.0 / weight;
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++) {
              double lQi_new = 0;
              for (int n = 0; n < B; n++) lQi_new += iweight * rhs[I_RHS(n, i, j, k, m)] * IK1(order, l, n);
              sq_res += (lQi_new - lQi[I_LQI(i, j, k, l, m)]) * (lQi_new - lQi[I_LQI(i, j, k, l, m)]);
              lQi[I_LQI(i, j, k, l, m)] = lQi_new;
            }
        }
    /* 5. exit condition (:314-318) */
    const double tol = 1e-7;
    if (sq_res < tol * tol) break;
  }
  return iter < MaxIterations ? iter + 1 : MaxIterations;
}

/* aderPredictorNonlinear, K/spaceTimePredictorNonlinear.cpph:337-409 */
ORA_INLINE void ader_predictor(const ora_solver* s, const int order, const int V, const int P,
                               const double* lQi, const double* lFi, double* lQhi, double* lFhi_x,
                               double* lFhi_y, double* lFhi_z, double* lShi) {
  const int D = V + P;
  const int B = order + 1;
  const int B3 = B * B * B;
#define I_H(i, j, k, m, NN) ((((i) * B + (j)) * B + (k)) * (NN) + (m))
  memset(lQhi, 0, sizeof(double) * B3 * D);
  memset(lFhi_x, 0, sizeof(double) * B3 * V);
  memset(lFhi_y, 0, sizeof(double) * B3 * V);
  memset(lFhi_z, 0, sizeof(double) * B3 * V);
  memset(lShi, 0, sizeof(double) * B3 * V);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        for (int l = 0; l < V; l++)
          for (int m = 0; m < B; m++) {
            if (s->use_flux) {
              lFhi_x[I_H(i, j, k, l, V)] += lFi[I_LFI(m, i, j, k, 0, l)] * W(order, m);
              lFhi_y[I_H(i, k, j, l, V)] += lFi[I_LFI(m, i, j, k, 1, l)] * W(order, m);
              lFhi_z[I_H(j, k, i, l, V)] += lFi[I_LFI(m, i, j, k, 2, l)] * W(order, m);
            }
            if (s->use_source || s->use_ncp)
              lShi[I_H(i, j, k, l, V)] += lFi[I_LFI(m, i, j, k, 3, l)] * W(order, m);
          }
        for (int l = 0; l < D; l++)
          for (int m = 0; m < B; m++) lQhi[I_H(i, j, k, l, D)] += lQi[I_LQI(i, j, k, m, l)] * W(order, m);
      }
}

/* aderExtrapolatorNonlinear, K/spaceTimePredictorNonlinear.cpph:411-528 */
ORA_INLINE void ader_extrapolator(const ora_solver* s, const int order, const int V, const int P,
                                  const double* lQhi, const double* lFhi_x, const double* lFhi_y,
                                  const double* lFhi_z, double* lQhbnd, double* lFhbnd) {
  const int D = V + P;
  const int B = order + 1;
  const int B2 = B * B;
#define I_BND(f, a, b, m, NN) ((((f) * B + (a)) * B + (b)) * (NN) + (m))
  memset(lQhbnd, 0, sizeof(double) * 6 * B2 * D);
  memset(lFhbnd, 0, sizeof(double) * 6 * B2 * V);
  for (int i = 0; i < B; i++) /* x faces (:435-463) */
    for (int j = 0; j < B; j++) {
      if (s->use_flux)
        for (int k = 0; k < V; k++)
          for (int l = 0; l < B; l++) {
            lFhbnd[I_BND(0, i, j, k, V)] += lFhi_x[I_H(i, j, l, k, V)] * FLC(order, l);
            lFhbnd[I_BND(1, i, j, k, V)] += lFhi_x[I_H(i, j, l, k, V)] * FRC(order, l);
          }
      for (int k = 0; k < D; k++)
        for (int l = 0; l < B; l++) {
          lQhbnd[I_BND(0, i, j, k, D)] += lQhi[I_H(i, j, l, k, D)] * FLC(order, l);
          lQhbnd[I_BND(1, i, j, k, D)] += lQhi[I_H(i, j, l, k, D)] * FRC(order, l);
        }
    }
  for (int i = 0; i < B; i++) /* y faces (:466-495) */
    for (int j = 0; j < B; j++) {
      if (s->use_flux)
        for (int k = 0; k < V; k++)
          for (int l = 0; l < B; l++) {
            lFhbnd[I_BND(2, i, j, k, V)] += lFhi_y[I_H(i, j, l, k, V)] * FLC(order, l);
            lFhbnd[I_BND(3, i, j, k, V)] += lFhi_y[I_H(i, j, l, k, V)] * FRC(order, l);
          }
      for (int k = 0; k < D; k++)
        for (int l = 0; l < B; l++) {
          lQhbnd[I_BND(2, i, j, k, D)] += lQhi[I_H(i, l, j, k, D)] * FLC(order, l);
          lQhbnd[I_BND(3, i, j, k, D)] += lQhi[I_H(i, l, j, k, D)] * FRC(order, l);
        }
    }
  for (int i = 0; i < B; i++) /* z faces (:498-527) */
    for (int j = 0; j < B; j++) {
      if (s->use_flux)
        for (int k = 0; k < V; k++)
          for (int l = 0; l < B; l++) {
            lFhbnd[I_BND(4, i, j, k, V)] += lFhi_z[I_H(i, j, l, k, V)] * FLC(order, l);
            lFhbnd[I_BND(5, i, j, k, V)] += lFhi_z[I_H(i, j, l, k, V)] * FRC(order, l);
          }
      for (int k = 0; k < D; k++)
        for (int l = 0; l < B; l++) {
          lQhbnd[I_BND(4, i, j, k, D)] += lQhi[I_H(l, i, j, k, D)] * FLC(order, l);
          lQhbnd[I_BND(5, i, j, k, D)] += lQhi[I_H(l, i, j, k, D)] * FRC(order, l);
        }
    }
}

/* aderTimeAveragingExtrapolatorNonlinear, K/spaceTimePredictorNonlinear.cpph:530-657 */
ORA_INLINE void ader_time_averaging_extrapolator(const int order, const int V, const int P,
                                                 const double* lQi, const double* lFi,
                                                 double* lQhbnd, double* lFhbnd) {
  const int D = V + P;
  const int B = order + 1;
  const int B2 = B * B;
  memset(lQhbnd, 0, sizeof(double) * 6 * B2 * D);
  memset(lFhbnd, 0, sizeof(double) * 6 * B2 * V);
  for (int i = 0; i < B; i++) {
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        for (int l = 0; l < V; l++)
          for (int n = 0; n < B; n++) {
            lFhbnd[I_BND(0, j, k, l, V)] += lFi[I_LFI(i, j, k, n, 0, l)] * FLC(order, n) * W(order, i);
            lFhbnd[I_BND(1, j, k, l, V)] += lFi[I_LFI(i, j, k, n, 0, l)] * FRC(order, n) * W(order, i);
          }
        for (int l = 0; l < D; l++)
          for (int n = 0; n < B; n++) {
            lQhbnd[I_BND(0, j, k, l, D)] += lQi[I_LQI(j, k, n, i, l)] * FLC(order, n) * W(order, i);
            lQhbnd[I_BND(1, j, k, l, D)] += lQi[I_LQI(j, k, n, i, l)] * FRC(order, n) * W(order, i);
          }
      }
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        for (int l = 0; l < V; l++)
          for (int n = 0; n < B; n++) {
            lFhbnd[I_BND(2, j, k, l, V)] += lFi[I_LFI(i, j, n, k, 1, l)] * FLC(order, n) * W(order, i);
            lFhbnd[I_BND(3, j, k, l, V)] += lFi[I_LFI(i, j, n, k, 1, l)] * FRC(order, n) * W(order, i);
          }
        for (int l = 0; l < D; l++)
          for (int n = 0; n < B; n++) {
            lQhbnd[I_BND(2, j, k, l, D)] += lQi[I_LQI(j, n, k, i, l)] * FLC(order, n) * W(order, i);
            lQhbnd[I_BND(3, j, k, l, D)] += lQi[I_LQI(j, n, k, i, l)] * FRC(order, n) * W(order, i);
          }
      }
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        for (int l = 0; l < V; l++)
          for (int n = 0; n < B; n++) {
            lFhbnd[I_BND(4, j, k, l, V)] += lFi[I_LFI(i, n, j, k, 2, l)] * FLC(order, n) * W(order, i);
            lFhbnd[I_BND(5, j, k, l, V)] += lFi[I_LFI(i, n, j, k, 2, l)] * FRC(order, n) * W(order, i);
          }
        for (int l = 0; l < D; l++)
          for (int n = 0; n < B; n++) {
            lQhbnd[I_BND(4, j, k, l, D)] += lQi[I_LQI(n, j, k, i, l)] * FLC(order, n) * W(order, i);
            lQhbnd[I_BND(5, j, k, l, D)] += lQi[I_LQI(n, j, k, i, l)] * FRC(order, n) * W(order, i);
          }
      }
  }
}

/* spaceTimePredictorNonlinear, K/spaceTimePredictorNonlinear.cpph:661-717 */
ORA_INLINE int stp_body(const ora_solver* s, const int order, const int V, const int P,
                        double* lQhbnd, double* lFhbnd, double* lQi, double* rhs, double* lFi,
                        double* gradQ, double* lQhi, double* lFhi, const double* luh,
                        const double* dx, double dt) {
  const int B = order + 1;
  const int B3 = B * B * B;
  int iters = picard_loop(s, order, V, P, luh, dt, dx, lQi, rhs, lFi, gradQ);
  if (s->no_time_averaging) {
    ader_time_averaging_extrapolator(order, V, P, lQi, lFi, lQhbnd, lFhbnd);
  } else {
    ader_predictor(s, order, V, P, lQi, lFi, lQhi, lFhi + 0 * B3 * V, lFhi + 1 * B3 * V,
                   lFhi + 2 * B3 * V, lFhi + 3 * B3 * V);
    ader_extrapolator(s, order, V, P, lQhi, lFhi + 0 * B3 * V, lFhi + 1 * B3 * V,
                      lFhi + 2 * B3 * V, lQhbnd, lFhbnd);
  }
  return iters;
}

/* Specialise the common (order, V, P) combinations so the compiler sees
 * constant loop bounds, as the reference's template instantiation does. */
#define ORA_DISPATCH(CALL, ...)                                              \
  do {                                                                       \
    const int o_ = s->order, v_ = s->nvar, p_ = s->npar;                     \
    if (v_ == 5 && p_ == 0) {                                                \
      switch (o_) {                                                          \
        case 2: CALL(2, 5, 0, __VA_ARGS__); break;                           \
        case 3: CALL(3, 5, 0, __VA_ARGS__); break;                           \
        case 4: CALL(4, 5, 0, __VA_ARGS__); break;                           \
        case 5: CALL(5, 5, 0, __VA_ARGS__); break;                           \
        case 6: CALL(6, 5, 0, __VA_ARGS__); break;                           \
        case 7: CALL(7, 5, 0, __VA_ARGS__); break;                           \
        default: CALL(o_, v_, p_, __VA_ARGS__); break;                       \
      }                                                                      \
    } else {                                                                 \
      CALL(o_, v_, p_, __VA_ARGS__);                                         \
    }                                                                        \
  } while (0)

#define STP_CALL(O, V, P, ...) ret = stp_body(s, O, V, P, __VA_ARGS__)

int ora_space_time_predictor_nonlinear(const ora_solver* s, double* lQhbnd, double* lFhbnd,
                                       double* lQi, double* rhs, double* lFi, double* gradQ,
                                       double* lQhi, double* lFhi, const double* luh,
                                       const double* dx, double dt) {
  int ret = 0;
  ORA_DISPATCH(STP_CALL, lQhbnd, lFhbnd, lQi, rhs, lFi, gradQ, lQhi, lFhi, luh, dx, dt);
  return ret;
}

/* ------------------------------------------------------------------------ */
/* Volume integral, K/volumeIntegralNonlinear.cpph:31-226                      */
/* ------------------------------------------------------------------------ */
ORA_INLINE void volume_body(const ora_solver* s, const int order, const int V, double* lduh,
                            const double* lFi, const double* lFhi, const double* dx) {
  const int B = order + 1;
  const int B3 = B * B * B;
  const int useFlux = s->use_flux, useSourceOrNCP = s->use_source || s->use_ncp;
  const double invDx[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
#define I4(i, j, k, m) ((((i) * B + (j)) * B + (k)) * V + (m))
  memset(lduh, 0, sizeof(double) * B3 * V);
  if (s->no_time_averaging) {
    if (useFlux) { /* :44-110 */
      for (int i = 0; i < B; i++) {
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++) {
            const double weight = W(order, i) * W(order, j) * W(order, k);
            const double updateSize = weight * invDx[0];
            for (int l = 0; l < B; l++)
              for (int m = 0; m < V; m++)
                for (int n = 0; n < B; n++)
                  lduh[I4(j, k, l, m)] += KXI(order, l, n) * lFi[I_LFI(i, j, k, n, 0, m)] * updateSize;
          }
        for (int j = 0; j < B; j++)
          for (int l = 0; l < B; l++) {
            const double weight = W(order, i) * W(order, j) * W(order, l);
            const double updateSize = weight * invDx[1];
            for (int k = 0; k < B; k++)
              for (int m = 0; m < V; m++)
                for (int n = 0; n < B; n++)
                  lduh[I4(j, k, l, m)] += KXI(order, k, n) * lFi[I_LFI(i, j, n, l, 1, m)] * updateSize;
          }
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++) {
            const double weight = W(order, i) * W(order, k) * W(order, l);
            const double updateSize = weight * invDx[2];
            for (int j = 0; j < B; j++)
              for (int m = 0; m < V; m++)
                for (int n = 0; n < B; n++)
                  lduh[I4(j, k, l, m)] += KXI(order, j, n) * lFi[I_LFI(i, n, k, l, 2, m)] * updateSize;
          }
      }
    }
    if (useSourceOrNCP) { /* :113-134 */
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++)
            for (int l = 0; l < B; l++) {
              const double weight = W(order, i) * W(order, j) * W(order, k) * W(order, l);
              for (int m = 0; m < V; m++) lduh[I4(j, k, l, m)] += weight * lFi[I_LFI(i, j, k, l, 3, m)];
            }
    }
  } else {
    if (useFlux) { /* :138-204 */
      const int x_offset = 0 * B3 * V;
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++) {
          const double weight = W(order, i) * W(order, j);
          const double updateSize = weight * invDx[0];
          for (int k = 0; k < B; k++)
            for (int l = 0; l < V; l++)
              for (int m = 0; m < B; m++)
                lduh[I4(i, j, k, l)] += KXI(order, k, m) * lFhi[x_offset + I4(i, j, m, l)] * updateSize;
        }
      const int y_offset = 1 * B3 * V;
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++) {
          const double weight = W(order, i) * W(order, j);
          const double updateSize = weight * invDx[1];
          for (int k = 0; k < B; k++)
            for (int l = 0; l < V; l++)
              for (int m = 0; m < B; m++)
                lduh[I4(i, k, j, l)] += KXI(order, k, m) * lFhi[y_offset + I4(i, j, m, l)] * updateSize;
        }
      const int z_offset = 2 * B3 * V;
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++) {
          const double weight = W(order, i) * W(order, j);
          const double updateSize = weight * invDx[2];
          for (int k = 0; k < B; k++)
            for (int l = 0; l < V; l++)
              for (int m = 0; m < B; m++)
                lduh[I4(k, i, j, l)] += KXI(order, k, m) * lFhi[z_offset + I4(i, j, m, l)] * updateSize;
        }
    }
    if (useSourceOrNCP) { /* :206-224 */
      const int s_offset = 3 * B3 * V;
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++) {
            const double weight = W(order, i) * W(order, j) * W(order, k);
            for (int l = 0; l < V; l++) lduh[I4(i, j, k, l)] += weight * lFhi[s_offset + I4(i, j, k, l)];
          }
    }
  }
}

#define VOL_CALL(O, V, P, ...) volume_body(s, O, V, __VA_ARGS__)
void ora_volume_integral_nonlinear(const ora_solver* s, double* lduh, const double* lFi,
                                   const double* lFhi, const double* dx) {
  ORA_DISPATCH(VOL_CALL, lduh, lFi, lFhi, dx);
}

/* ------------------------------------------------------------------------ */
/* Surface integral, K/surfaceIntegralNonlinear.cpph:21-90                     */
/* ------------------------------------------------------------------------ */
ORA_INLINE void surface_body(const int order, const int V, double* lduh, const double* lFhbnd,
                             const double* dx) {
  const int B = order + 1;
  const double invDx[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      const double weight = W(order, i) * W(order, j);
      const double updateSize = weight * invDx[0];
      for (int k = 0; k < B; k++)
        for (int l = 0; l < V; l++)
          lduh[I4(i, j, k, l)] -= (lFhbnd[I_BND(1, i, j, l, V)] * FRC(order, k) -
                                   lFhbnd[I_BND(0, i, j, l, V)] * FLC(order, k)) * updateSize;
    }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = W(order, i) * W(order, k);
        const double updateSize = weight * invDx[1];
        for (int l = 0; l < V; l++)
          lduh[I4(i, j, k, l)] -= (lFhbnd[I_BND(3, i, k, l, V)] * FRC(order, j) -
                                   lFhbnd[I_BND(2, i, k, l, V)] * FLC(order, j)) * updateSize;
      }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = W(order, j) * W(order, k);
        const double updateSize = weight * invDx[2];
        for (int l = 0; l < V; l++)
          lduh[I4(i, j, k, l)] -= (lFhbnd[I_BND(5, j, k, l, V)] * FRC(order, i) -
                                   lFhbnd[I_BND(4, j, k, l, V)] * FLC(order, i)) * updateSize;
      }
}

#define SURF_CALL(O, V, P, ...) surface_body(O, V, __VA_ARGS__)
void ora_surface_integral_nonlinear(const ora_solver* s, double* lduh, const double* lFhbnd,
                                    const double* dx) {
  ORA_DISPATCH(SURF_CALL, lduh, lFhbnd, dx);
}

/* ------------------------------------------------------------------------ */
/* Rusanov Riemann solver, K/riemannSolverNonlinear.cpph:37-117                */
/* ------------------------------------------------------------------------ */
ORA_INLINE void riemann_body(const ora_solver* s, const int order, const int V, const int P,
                             double* FL, double* FR, const double* QL, const double* QR,
                             const int direction) {
  const int D = V + P;
  const int B = order + 1;
  double QavL[64] = {0.0}, QavR[64] = {0.0};
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      const double weight = W(order, i) * W(order, j);
      for (int k = 0; k < D; k++) {
        QavL[k] += weight * QL[(i * B + j) * D + k];
        QavR[k] += weight * QR[(i * B + j) * D + k];
      }
    }
  double LL[64] = {0.0}, LR[64] = {0.0};
  ora_eigenvalues(s, QavL, direction, LL);
  ora_eigenvalues(s, QavR, direction, LR);
  for (int k = 0; k < V; k++) {
    LL[k] = fabs(LL[k]);
    LR[k] = fabs(LR[k]);
  }
  double smax_L = LL[0], smax_R = LR[0]; /* std::max_element: first maximum */
  for (int k = 1; k < V; k++) {
    if (smax_L < LL[k]) smax_L = LL[k];
    if (smax_R < LR[k]) smax_R = LR[k];
  }
  const double smax = (smax_L < smax_R) ? smax_R : smax_L; /* std::max */
  double Qavg[64];
  double gradQ[3][64];
  double ncp[64] = {0.0};
  memset(gradQ, 0, sizeof(gradQ));
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      if (s->use_ncp) {
        for (int l = 0; l < V; l++) {
          gradQ[direction][l] = QR[(i * B + j) * D + l] - QL[(i * B + j) * D + l];
          Qavg[l] = 0.5 * (QR[(i * B + j) * D + l] + QL[(i * B + j) * D + l]);
        }
        /* The reference leaves Qavg[V..D) uninitialised (:81,:93-96); the
         * parameters are averaged here instead. */
        for (int l = V; l < D; l++) Qavg[l] = 0.5 * (QR[(i * B + j) * D + l] + QL[(i * B + j) * D + l]);
        double gq[3 * 64];
        for (int d = 0; d < 3; d++)
          for (int l = 0; l < V; l++) gq[d * V + l] = gradQ[d][l];
        ora_ncp(s, Qavg, gq, ncp);
      }
      for (int k = 0; k < V; k++) {
        const int f = (i * B + j) * V + k;
        const int q = (i * B + j) * D + k;
        FL[f] = 0.5 * (FR[f] + FL[f]) - 0.5 * smax * (QR[q] - QL[q]);
        if (s->use_ncp) {
          FR[f] = FL[f] - 0.5 * ncp[k];
          FL[f] = FL[f] + 0.5 * ncp[k];
        } else {
          FR[f] = FL[f];
        }
      }
    }
}

#define RIEM_CALL(O, V, P, ...) riemann_body(s, O, V, P, __VA_ARGS__)
void ora_riemann_solver_nonlinear(const ora_solver* s, double* FL, double* FR, const double* QL,
                                  const double* QR, double dt, int direction) {
  (void)dt;
  ORA_DISPATCH(RIEM_CALL, FL, FR, QL, QR, direction);
}

/* ------------------------------------------------------------------------ */
/* Solution update, K/solutionUpdate.cpph:23-47                              */
/* ------------------------------------------------------------------------ */
ORA_INLINE void update_body(const int order, const int V, const int P, double* luh,
                            const double* lduh, const double dt) {
  const int D = V + P;
  const int B = order + 1;
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = W(order, i) * W(order, j) * W(order, k);
        const double updateSize = dt / weight;
        for (int l = 0; l < V; l++)
          luh[(((i * B) + j) * B + k) * D + l] += lduh[(((i * B) + j) * B + k) * V + l] * updateSize;
      }
}

#define UPD_CALL(O, V, P, ...) update_body(O, V, P, __VA_ARGS__)
void ora_solution_update(const ora_solver* s, double* luh, const double* lduh, double dt) {
  ORA_DISPATCH(UPD_CALL, luh, lduh, dt);
}

/* ------------------------------------------------------------------------ */
/* Stable time step, K/stableTimeStepSize.cpph:21-63                          */
/* ------------------------------------------------------------------------ */
ORA_INLINE double dt_body(const ora_solver* s, const int order, const int V, const int P,
                          const double* luh, const double* dx) {
  const int D = V + P;
  const int B = order + 1;
  const double cflFactor = s->cfl;
  static const double PNPM[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};
  const double invDx[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
  double dt = DBL_MAX;
  double lambda[64] = {0.0};
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        double denominator = 0.0;
        for (int d = 0; d < 3; d++) {
          ora_eigenvalues(s, luh + (((i * B) + j) * B + k) * D, d, lambda);
          double maxEigenvalue = 0.0;
          for (int ivar = 0; ivar < V; ivar++) {
            const double a = fabs(lambda[ivar]);
            maxEigenvalue = (a < maxEigenvalue) ? maxEigenvalue : a; /* std::max(a, m) */
          }
          denominator += maxEigenvalue * invDx[d];
        }
        const double cand = cflFactor * PNPM[order] / denominator;
        dt = (cand < dt) ? cand : dt; /* std::min(dt, cand) */
      }
  return dt;
}

double ora_stable_time_step_size(const ora_solver* s, const double* luh, const double* dx) {
  double ret = 0.0;
#define DT_CALL(O, V, P, ...) ret = dt_body(s, O, V, P, __VA_ARGS__)
  ORA_DISPATCH(DT_CALL, luh, dx);
  return ret;
}

/* ------------------------------------------------------------------------ */
/* boundaryConditions (K/c/3d/boundaryConditions.cpph:24-85): the user's     */
/* boundaryValues at every node of the face, x on the face and the two free   */
/* coordinates from the nodes (i along d2, j along d1).                       */
/* ------------------------------------------------------------------------ */
void ora_boundary_conditions(const ora_solver* s, ora_boundary_values_fn bv, double* FOut, double* QOut,
                             const double* FIn, const double* QIn, const double* centre, const double* size,
                             double t, double dt, int faceIndex, int normalNonZero) {
  const int V = s->nvar, D = s->nvar + s->npar, B = s->order + 1;
  const int f = faceIndex - 2 * normalNonZero;
  int d1 = 0, d2 = 0;
  switch (normalNonZero) {
    case 0: d1 = 1; d2 = 2; break;
    case 1: d1 = 0; d2 = 2; break;
    default: d1 = 0; d2 = 1; break;
  }
  double x[3];
  x[normalNonZero] = centre[normalNonZero] + (-0.5 + f) * size[normalNonZero];
  for (int i = 0; i < B; i++) {
    x[d2] = centre[d2] + size[d2] * (ora_table(ORA_T_NODES, s->order, i, 0) - 0.5);
    for (int j = 0; j < B; j++) {
      x[d1] = centre[d1] + size[d1] * (ora_table(ORA_T_NODES, s->order, j, 0) - 0.5);
      bv(s, x, t, dt, faceIndex, normalNonZero, FIn + (i * B + j) * V, QIn + (i * B + j) * D,
         FOut + (i * B + j) * V, QOut + (i * B + j) * D);
    }
  }
}

/* solutionAdjustment (K/c/3d/solutionAdjustment.cpph:17-47): the user's     */
/* adjustPointSolution at every node, x = centre + dx (xi - 1/2).            */
void ora_solution_adjustment(const ora_solver* s, ora_adjust_fn adj, double* luh, const double* centre,
                             const double* dx, double t, double dt) {
  const int D = s->nvar + s->npar, B = s->order + 1;
  double x[3];
  for (int i = 0; i < B; i++) {
    const double qi = ora_table(ORA_T_NODES, s->order, i, 0);
    for (int j = 0; j < B; j++) {
      const double qj = ora_table(ORA_T_NODES, s->order, j, 0);
      for (int k = 0; k < B; k++) {
        const double qk = ora_table(ORA_T_NODES, s->order, k, 0);
        x[2] = centre[2] + dx[2] * (qi - 0.5);
        x[1] = centre[1] + dx[1] * (qj - 0.5);
        x[0] = centre[0] + dx[0] * (qk - 0.5);
        adj(s, x, t, dt, luh + ((i * B + j) * B + k) * D);
      }
    }
  }
}

/* The test solver's boundaryValues: the reference Euler demonstrator's      */
/* reflecting wall (Demonstrators/EulerADERDG/MyEulerSolver_ADERDG.cpp:       */
/* 63-88): the inside state with the normal momentum negated, its normal flux. */
void ora_bv_reflecting_wall(const ora_solver* s, const double* x, double t, double dt, int faceIndex,
                            int normalNonZero, const double* FIn, const double* QIn, double* FOut, double* QOut) {
  (void)x; (void)t; (void)dt; (void)faceIndex; (void)FIn;
  const int V = s->nvar;
  for (int m = 0; m < V; m++) QOut[m] = QIn[m];
  QOut[1 + normalNonZero] = -QOut[1 + normalNonZero];
  double f0[16], f1[16], f2[16];
  double* F[3] = {f0, f1, f2};
  ora_flux(s, QOut, F);
  for (int m = 0; m < V; m++) FOut[m] = F[normalNonZero][m];
}

/* The test solver's adjustPointSolution: at t = 0 the entropy wave          */
/* rho = 1 + 0.2 sin(2 pi (x + y + z)), u = v = w = 1, p = 1; afterwards      */
/* Q += dt/2 (Q_0(x) - Q).                                                    */
void ora_adjust_entropy_wave(const ora_solver* s, const double* x, double t, double dt, double* Q) {
  const double twopi = 6.283185307179586476925286766559;
  const double rho = 1.0 + 0.2 * sin(twopi * (1.0 * x[0] + 1.0 * x[1] + 1.0 * x[2]));
  const double Q0[5] = {rho, rho * 1.0, rho * 1.0, rho * 1.0,
                        1.0 / (1.4 - 1.0) + 0.5 * rho * (1.0 * 1.0 + 1.0 * 1.0 + 1.0 * 1.0)};
  const int V = s->nvar < 5 ? s->nvar : 5;
  if (t == 0.0) {
    for (int m = 0; m < V; m++) Q[m] = Q0[m];
  } else {
    for (int m = 0; m < V; m++) Q[m] += 0.5 * dt * (Q0[m] - Q[m]);
  }
}

/* ------------------------------------------------------------------------ */
/* Grid driver: non-fused global time stepping on a uniform periodic grid.   */
/* Restates the phase order of Runner::runOneTimeStepWithThreeSeparate-      */
/* AlgorithmicSteps (ExaHyPE/exahype/runners/Runner.cpp:1116-1162):           */
/*   Merging  -> ADERDGSolver::solveRiemannProblemAtInterface (cpp:2789):    */
/*               riemann(FL=lFhbnd_L[2d+1], FR=lFhbnd_R[2d], QL, QR, dt, d)   */
/*   Update   -> ADERDGSolver::updateSolution (cpp:2185): surface integral,  */
/*               solution update with the corrector dt                       */
/*   dt       -> startNewTimeStep (cpp:2038) + global min (SolutionUpdate.cpp:269) */
/*   Predict  -> performPredictionAndVolumeIntegral (cpp:1881)               */
/* ------------------------------------------------------------------------ */
struct ora_grid {
  ora_solver s;
  int nx, ny, nz;
  int transmissive;
  ora_boundary_values_fn bv;  /* boundary faces: the user's boundaryValues (NULL: none) */
  ora_adjust_fn adj;          /* the user's adjustPointSolution after every update (NULL: none) */
  double lo[3];
  long ncells;
  double dx[3];
  double t, dt;
  /* fused time stepping (fuse-algorithmic-steps): dt is the corrector time
   * step size, dtp the predictor's; weight is fuse-algorithmic-steps-factor */
  int fused;
  double weight, dtp;
  long reruns;
  int have_prediction;
  double *luh, *lduh, *lQhbnd, *lFhbnd;
  int* iters;
};

static long cell_index(const ora_grid* g, int i, int j, int k) {
  i = (i + g->nx) % g->nx;
  j = (j + g->ny) % g->ny;
  k = (k + g->nz) % g->nz;
  return ((long)k * g->ny + j) * g->nx + i;
}

ora_grid* ora_grid_create(const ora_solver* s, int nx, int ny, int nz, const double* lo,
                          const double* hi) {
  return ora_grid_create_bc(s, nx, ny, nz, lo, hi, 0);
}

ora_grid* ora_grid_create_bc(const ora_solver* s, int nx, int ny, int nz, const double* lo,
                             const double* hi, int transmissive) {
  ora_grid* g = (ora_grid*)calloc(1, sizeof(ora_grid));
  g->transmissive = transmissive;
  g->lo[0] = lo[0];
  g->lo[1] = lo[1];
  g->lo[2] = lo[2];
  g->s = *s;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->ncells = (long)nx * ny * nz;
  g->dx[0] = (hi[0] - lo[0]) / nx;
  g->dx[1] = (hi[1] - lo[1]) / ny;
  g->dx[2] = (hi[2] - lo[2]) / nz;
  const int B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B2 = B * B, B3 = B2 * B;
  g->luh = (double*)calloc(g->ncells * B3 * D, sizeof(double));
  g->lduh = (double*)calloc(g->ncells * B3 * V, sizeof(double));
  g->lQhbnd = (double*)calloc(g->ncells * 6 * B2 * D, sizeof(double));
  g->lFhbnd = (double*)calloc(g->ncells * 6 * B2 * V, sizeof(double));
  g->iters = (int*)calloc(g->ncells, sizeof(int));
  return g;
}

void ora_grid_destroy(ora_grid* g) {
  if (!g) return;
  free(g->luh);
  free(g->lduh);
  free(g->lQhbnd);
  free(g->lFhbnd);
  free(g->iters);
  free(g);
}

long ora_grid_num_cells(const ora_grid* g) { return g->ncells; }
double ora_grid_time(const ora_grid* g) { return g->t; }

void ora_grid_picard_iterations(const ora_grid* g, int* out) {
  memcpy(out, g->iters, sizeof(int) * g->ncells);
}

void ora_grid_set_solution(ora_grid* g, const double* luh) {
  const int B = g->s.order + 1, D = g->s.nvar + g->s.npar;
  memcpy(g->luh, luh, sizeof(double) * g->ncells * B * B * B * D);
  g->have_prediction = 0;
}

void ora_grid_get_solution(const ora_grid* g, double* luh) {
  const int B = g->s.order + 1, D = g->s.nvar + g->s.npar;
  memcpy(luh, g->luh, sizeof(double) * g->ncells * B * B * B * D);
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

double ora_grid_stable_dt(ora_grid* g, int nthreads) {
  const int B = g->s.order + 1, D = g->s.nvar + g->s.npar;
  const long B3 = (long)B * B * B;
  double dtmin = DBL_MAX;
  set_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(min : dtmin)
  for (long c = 0; c < g->ncells; c++) {
    const double d = ora_stable_time_step_size(&g->s, g->luh + c * B3 * D, g->dx);
    if (d < dtmin) dtmin = d;
  }
  return dtmin;
}

typedef struct {
  double *lQi, *rhs, *lFi, *gradQ, *lQhi, *lFhi;
} scratch_t;

static scratch_t scratch_alloc(const ora_solver* s) {
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B3 = B * B * B, B4 = B3 * B;
  scratch_t t;
  t.lQi = (double*)malloc(sizeof(double) * (B4 + B3) * D); /* (B+1) B^3 D for the linear predictor */
  t.rhs = (double*)malloc(sizeof(double) * B4 * D);
  t.lFi = (double*)malloc(sizeof(double) * B4 * 4 * V);
  t.gradQ = (double*)malloc(sizeof(double) * B4 * 3 * V);
  t.lQhi = (double*)malloc(sizeof(double) * B3 * D);
  t.lFhi = (double*)malloc(sizeof(double) * B3 * 4 * V);
  return t;
}

static void scratch_free(scratch_t* t) {
  free(t->lQi);
  free(t->rhs);
  free(t->lFi);
  free(t->gradQ);
  free(t->lQhi);
  free(t->lFhi);
}

/* Prediction adapter: predictor + volume integral on cells [c0, c1). */
static void predict_cells(ora_grid* g, long c0, long c1, double dt) {
  const ora_solver* s = &g->s;
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B2 = B * B, B3 = B2 * B;
#pragma omp parallel
  {
    scratch_t t = scratch_alloc(s);
#pragma omp for schedule(dynamic)
    for (long c = c0; c < c1; c++) {
      if (s->linear) { /* the glue's isLinear branch (AbstractGenericADERDGSolverImplementation.template:53-86) */
        ora_space_time_predictor_linear(s, g->lQhbnd + c * 6 * B2 * D, g->lFhbnd + c * 6 * B2 * V, t.lQi, t.lFi,
                                        t.gradQ, t.lQhi, t.lFhi, g->luh + c * B3 * D, g->dx, dt, NULL);
        ora_volume_integral_linear(s, g->lduh + c * B3 * V, t.lFhi);
        g->iters[c] = 0;
        continue;
      }
      g->iters[c] = ora_space_time_predictor_nonlinear(
          s, g->lQhbnd + c * 6 * B2 * D, g->lFhbnd + c * 6 * B2 * V, t.lQi, t.rhs, t.lFi, t.gradQ,
          t.lQhi, t.lFhi, g->luh + c * B3 * D, g->dx, dt);
      ora_volume_integral_nonlinear(s, g->lduh + c * B3 * V, t.lFi, t.lFhi, g->dx);
    }
    scratch_free(&t);
  }
}

static void riemann_face(ora_grid* g, long cl, long cr, int d) {
  const ora_solver* s = &g->s;
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B2 = B * B;
  double* FL = g->lFhbnd + (cl * 6 + 2 * d + 1) * B2 * V;
  double* FR = g->lFhbnd + (cr * 6 + 2 * d) * B2 * V;
  const double* QL = g->lQhbnd + (cl * 6 + 2 * d + 1) * B2 * D;
  const double* QR = g->lQhbnd + (cr * 6 + 2 * d) * B2 * D;
  if (s->linear)
    ora_riemann_solver_linear(s, FL, FR, QL, QR, g->dt, d);
  else
    ora_riemann_solver_nonlinear(s, FL, FR, QL, QR, g->dt, d);
}

static void cell_centre(const ora_grid* g, long c, double* centre) {
  const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
  centre[0] = g->lo[0] + (i + 0.5) * g->dx[0];
  centre[1] = g->lo[1] + (j + 0.5) * g->dx[1];
  centre[2] = g->lo[2] + (k + 0.5) * g->dx[2];
}

static void boundary_face(ora_grid* g, long c, int face) {
  const ora_solver* s = &g->s;
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B2 = B * B;
  double* FIn = g->lFhbnd + (c * 6 + face) * B2 * V;
  const double* QIn = g->lQhbnd + (c * 6 + face) * B2 * D;
  double FOut[10 * 10 * 16], QOut[10 * 10 * 16];
  memcpy(FOut, FIn, sizeof(double) * B2 * V);
  memcpy(QOut, QIn, sizeof(double) * B2 * D);
  if (g->bv) {
    double centre[3];
    cell_centre(g, c, centre);
    ora_boundary_conditions(s, g->bv, FOut, QOut, FIn, QIn, centre, g->dx, g->t, g->dt, face, face / 2);
  }
  if (s->linear) {
    if (face % 2 == 0)
      ora_riemann_solver_linear(s, FOut, FIn, QOut, QIn, g->dt, face / 2);
    else
      ora_riemann_solver_linear(s, FIn, FOut, QIn, QOut, g->dt, face / 2);
  } else if (face % 2 == 0) {
    ora_riemann_solver_nonlinear(s, FOut, FIn, QOut, QIn, g->dt, face / 2);
  } else {
    ora_riemann_solver_nonlinear(s, FIn, FOut, QIn, QOut, g->dt, face / 2);
  }
}

static void update_cells(ora_grid* g, long c0, long c1) {
  const ora_solver* s = &g->s;
  const long B = s->order + 1, D = s->nvar + s->npar, V = s->nvar;
  const long B2 = B * B, B3 = B2 * B;
#pragma omp parallel for schedule(static)
  for (long c = c0; c < c1; c++) {
    if (s->linear)
      ora_surface_integral_linear(s, g->lduh + c * B3 * V, g->lFhbnd + c * 6 * B2 * V, g->dx);
    else
      ora_surface_integral_nonlinear(s, g->lduh + c * B3 * V, g->lFhbnd + c * 6 * B2 * V, g->dx);
    ora_solution_update(s, g->luh + c * B3 * D, g->lduh + c * B3 * V, g->dt);
    if (g->adj) { /* adjustSolution(t + dt, dt) after the update (ADERDGSolver.cpp:2231-2236) */
      double centre[3];
      cell_centre(g, c, centre);
      ora_solution_adjustment(s, g->adj, g->luh + c * B3 * D, centre, g->dx, g->t + g->dt, g->dt);
    }
  }
}

void ora_grid_set_hooks(ora_grid* g, ora_boundary_values_fn bv, ora_adjust_fn adj) {
  g->bv = bv;
  if (bv) g->transmissive = 1; /* the boundary faces are not wrapped */
  g->adj = adj;
}

void ora_grid_set_user(ora_grid* g, int user_boundary, int adjust) {
  ora_grid_set_hooks(g, user_boundary ? ora_bv_reflecting_wall : NULL, adjust ? ora_adjust_entropy_wave : NULL);
}

void ora_grid_set_fused(ora_grid* g, int on, double weight) {
  g->fused = on;
  g->weight = weight;
  g->have_prediction = 0;
}

long ora_grid_predictor_reruns(const ora_grid* g) { return g->reruns; }

void ora_grid_adjust(ora_grid* g) {
  const long B = g->s.order + 1, D = g->s.nvar + g->s.npar;
  const long B3 = B * B * B;
  for (long c = 0; c < g->ncells; c++) {
    double centre[3];
    cell_centre(g, c, centre);
    ora_solution_adjustment(&g->s, g->adj ? g->adj : ora_adjust_entropy_wave, g->luh + c * B3 * D, centre, g->dx, g->t,
                            0.0);
  }
  g->have_prediction = 0;
}

/* Merging: each interior face once (periodic wrap included), then the
 * boundary faces (ADERDGSolver::applyBoundaryConditions, ADERDGSolver.cpp:
 * 2893-2966): boundaryValues (the inside state and flux when there is no
 * user hook: transmissive), then the Riemann solve with the outside state
 * left of even faces. */
static void merge_faces(ora_grid* g) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g->ncells; c++) {
    const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
    if (!g->transmissive || i + 1 < g->nx) riemann_face(g, c, cell_index(g, i + 1, j, k), 0);
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g->ncells; c++) {
    const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
    if (!g->transmissive || j + 1 < g->ny) riemann_face(g, c, cell_index(g, i, j + 1, k), 1);
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g->ncells; c++) {
    const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
    if (!g->transmissive || k + 1 < g->nz) riemann_face(g, c, cell_index(g, i, j, k + 1), 2);
  }
  if (g->transmissive) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < g->ncells; c++) {
      const int i = (int)(c % g->nx), j = (int)((c / g->nx) % g->ny), k = (int)(c / ((long)g->nx * g->ny));
      const int at[6] = {i == 0, i == g->nx - 1, j == 0, j == g->ny - 1, k == 0, k == g->nz - 1};
      for (int f = 0; f < 6; f++)
        if (at[f]) boundary_face(g, c, f);
    }
  }
}

/* Fused time stepping (fuse-algorithmic-steps, global time stepping, one
 * rank).  Start (Runner::initialiseMesh, TimeStepSizeComputation with
 * updateTimeStepSizesFused, TimeStepSizeComputation.cpp:226-237 and
 * ADERDGSolver.cpp:594-616): corrector and predictor dt = the stable dt of
 * the initial solution (the weighting at :230-234 is compiled only into MPI
 * builds), then the first Prediction sweep (Runner.cpp:691-699).  Each sweep
 * (Runner::runOneTimeStepWithFusedAlgorithmicSteps, Runner.cpp:1053-1110):
 * Merging with the corrector dt, then per cell fusedTimeStep
 * (ADERDGSolver.cpp:1853-1879): updateSolution with the corrector dt, the
 * prediction with the predictor dt, the admissible dt of the new solution;
 * then SolutionUpdate::endIteration (mappings/SolutionUpdate.cpp:307-318):
 * reinitialiseTimeStepDataIfLastPredictorTimeStepSizeWasInstable
 * (TimeStepSizeComputation.cpp:167-204) and startNewTimeStep
 * (ADERDGSolver.cpp:560-592); a violated stability condition re-runs the
 * prediction with the new predictor dt (Runner.cpp:1104-1110). */
static void grid_step_fused(ora_grid* g, int nsteps, int nthreads, double* dt_out, int* iters_out) {
  if (!g->have_prediction) {
    const double s = ora_grid_stable_dt(g, nthreads);
    g->dt = s;
    g->dtp = s;
    predict_cells(g, 0, g->ncells, g->dtp);
    g->have_prediction = 1;
  }
  for (int n = 0; n < nsteps; n++) {
    if (iters_out) memcpy(iters_out + (long)n * g->ncells, g->iters, sizeof(int) * g->ncells);
    if (dt_out) dt_out[n] = g->dt;
    merge_faces(g);
    update_cells(g, 0, g->ncells);
    predict_cells(g, 0, g->ncells, g->dtp);
    const double stable = ora_grid_stable_dt(g, nthreads); /* min over cells of startNewTimeStep(cell) */
    double next = stable;                                   /* _minNextPredictorTimeStepSize */
    double used = g->dtp;                                   /* _minPredictorTimeStepSize */
    if (used == 0.0) used = stable;
    const int violated = used > stable;
    if (violated) {
      next = fmin(next, g->weight * stable);
      g->dtp = g->weight * stable;
    } else {
      next = fmin(next, 0.5 * (used + g->weight * stable));
    }
    g->t += g->dt; /* corrector time stamp <- predictor time stamp */
    g->dt = g->dtp;
    g->dtp = next;
    if (violated) {
      predict_cells(g, 0, g->ncells, g->dtp);
      g->reruns++;
    }
  }
}

void ora_grid_step(ora_grid* g, int nsteps, int nthreads, double* dt_out, int* iters_out) {
  set_threads(nthreads);
  if (g->fused) {
    grid_step_fused(g, nsteps, nthreads, dt_out, iters_out);
    return;
  }
  if (!g->have_prediction) {
    g->dt = ora_grid_stable_dt(g, nthreads);
    predict_cells(g, 0, g->ncells, g->dt);
    g->have_prediction = 1;
  }
  for (int n = 0; n < nsteps; n++) {
    if (iters_out) memcpy(iters_out + (long)n * g->ncells, g->iters, sizeof(int) * g->ncells);
    if (dt_out) dt_out[n] = g->dt;
    merge_faces(g);
    /* SolutionUpdate, then the new global time step size. */
    update_cells(g, 0, g->ncells);
    g->t += g->dt;
    g->dt = ora_grid_stable_dt(g, nthreads);
    /* Prediction for the next step. */
    predict_cells(g, 0, g->ncells, g->dt);
  }
}

/* The standard scheme with prescribed time step sizes: step n predicts and
 * corrects with dts[n] (a test device: the fused scheme must equal it when
 * given the fused scheme's corrector dt sequence).  Leaves no prediction. */
void ora_grid_step_dts(ora_grid* g, int nsteps, const double* dts, int nthreads) {
  set_threads(nthreads);
  for (int n = 0; n < nsteps; n++) {
    g->dt = dts[n];
    predict_cells(g, 0, g->ncells, g->dt);
    merge_faces(g);
    update_cells(g, 0, g->ncells);
    g->t += g->dt;
  }
  g->have_prediction = 0;
}

void ora_grid_sample_step(ora_grid* g, long m, int nthreads) {
  if (m > g->ncells) m = g->ncells;
  set_threads(nthreads);
  if (!g->have_prediction) {
    g->dt = ora_grid_stable_dt(g, nthreads);
    g->have_prediction = 1;
  }
  predict_cells(g, 0, m, g->dt);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < m; c++) {
    riemann_face(g, c, (c + 1) % m, 0);
    riemann_face(g, c, (c + 2) % m, 1);
    riemann_face(g, c, (c + 3) % m, 2);
  }
  update_cells(g, 0, m);
  const ora_solver* s = &g->s;
  const long B = s->order + 1, D = s->nvar + s->npar;
  const long B3 = B * B * B;
  double dtmin = DBL_MAX;
#pragma omp parallel for schedule(static) reduction(min : dtmin)
  for (long c = 0; c < m; c++) {
    const double d = ora_stable_time_step_size(s, g->luh + c * B3 * D, g->dx);
    if (d < dtmin) dtmin = d;
  }
  (void)dtmin;
}
