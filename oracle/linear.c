/*
 * linear.c -- CPU restatement of the reference's generic 3-D ADER-DG kernels
 * for linear PDEs (Cauchy-Kowalevski predictor, linear Riemann solver, linear
 * volume and surface integrals) and of the PDE callbacks they use.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 * Citations (paths relative to the reference checkout, K =
 * ExaHyPE/kernels/aderdg/generic/c/3d):
 *   STPL = K/spaceTimePredictorLinear.cpph
 *   RSL  = K/riemannSolverLinear.cpph
 *   VIL  = K/volumeIntegralLinear.cpph
 *   SIL  = K/surfaceIntegralLinear.cpph
 *   ETS  = ExaHyPE/exahype/tests/kernels/c/ElasticityKernelTest2D.cpp
 * Layouts: luh (z,y,x,var) B^3 D; lQi (t,z,y,x,var) (B+1)B^3 D (time index
 * slowest, B+1 time derivatives); lFi (t,z,y,x,dir,var) B^4 3V (plus a
 * zeroed source block of B^4 V); gradQ (t,z,y,x,dir,var) B^4 3V; lQhi
 * (z,y,x,var) B^3 D; lFhi (dir,z,y,x,var) 3 B^3 V plus a zeroed source
 * block B^3 V -- all three directions in the same (z,y,x) order; PSi
 * (t,z,y,x,var) (B+1)B^3 V; faces (face,a,b,var).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline double std_max(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* PDE callbacks of the linear kernels                                       */
/* ------------------------------------------------------------------------ */

/* coefficientMatrix(Q, normalNonZero, Bn), Bn[l][k] = B_d(k, l) (the
 * transpose of the matrix of B(Q) grad Q, as RSL:119-121 reads it).
 * Elastic: ETS:183-228 for x and y, the z case written there as a comment
 * (:212-220).  Euler: the Euler test solver's is zero
 * (GenericEulerKernelTest3D.cpp:108-132). */
void ora_coefficient_matrix(const ora_solver* s, const double* Q, int d, double* Bn) {
  const int V = s->nvar;
  for (int i = 0; i < V * V; i++) Bn[i] = 0.0;
  if (s->pde != ORA_PDE_ELASTIC) return;
  const double lam = Q[9], mu = Q[10], irho = 1.0 / Q[11];
#define BN(l, k) Bn[(l) * V + (k)]
  if (d == 0) {
    BN(6, 0) = -(lam + 2 * mu);
    BN(6, 1) = -lam;
    BN(6, 2) = -lam;
    BN(7, 3) = -mu;
    BN(8, 5) = -mu;
    BN(0, 6) = -irho;
    BN(3, 7) = -irho;
    BN(5, 8) = -irho;
  } else if (d == 1) {
    BN(7, 0) = -lam;
    BN(7, 1) = -(lam + 2 * mu);
    BN(7, 2) = -lam;
    BN(6, 3) = -mu;
    BN(8, 4) = -mu;
    BN(3, 6) = -irho;
    BN(1, 7) = -irho;
    BN(4, 8) = -irho;
  } else {
    BN(8, 0) = -lam;
    BN(8, 1) = -lam;
    BN(8, 2) = -(lam + 2 * mu);
    BN(7, 4) = -mu;
    BN(6, 5) = -mu;
    BN(5, 6) = -irho;
    BN(4, 7) = -irho;
    BN(2, 8) = -irho;
  }
#undef BN
}

/* nonConservativeProduct of a linear solver: BgradQ is a dim x nVar tensor,
 * BgradQ[d] = B_d grad_d Q (STPL:32, ETS:130-181; the z block is the comment
 * at :172-180).  Euler: zero. */
void ora_ncp_directional(const ora_solver* s, const double* Q, const double* gradQ, double* BgradQ) {
  const int V = s->nvar;
  for (int i = 0; i < 3 * V; i++) BgradQ[i] = 0.0;
  if (s->pde != ORA_PDE_ELASTIC) return;
  const double lam = Q[V], mu = Q[V + 1], irho = 1.0 / Q[V + 2];
  const double* gx = gradQ;
  const double* gy = gradQ + V;
  const double* gz = gradQ + 2 * V;
  double* bx = BgradQ;
  double* by = BgradQ + V;
  double* bz = BgradQ + 2 * V;
  bx[0] = -(lam + 2 * mu) * gx[6];
  bx[1] = -lam * gx[6];
  bx[2] = -lam * gx[6];
  bx[3] = -mu * gx[7];
  bx[4] = 0.0;
  bx[5] = -mu * gx[8];
  bx[6] = -irho * gx[0];
  bx[7] = -irho * gx[3];
  bx[8] = -irho * gx[5];
  by[0] = -lam * gy[7];
  by[1] = -(lam + 2 * mu) * gy[7];
  by[2] = -lam * gy[7];
  by[3] = -mu * gy[6];
  by[4] = -mu * gy[8];
  by[5] = 0.0;
  by[6] = -irho * gy[3];
  by[7] = -irho * gy[1];
  by[8] = -irho * gy[4];
  bz[0] = -lam * gz[8];
  bz[1] = -lam * gz[8];
  bz[2] = -(lam + 2 * mu) * gz[8];
  bz[3] = 0.0;
  bz[4] = -mu * gz[7];
  bz[5] = -mu * gz[6];
  bz[6] = -irho * gz[5];
  bz[7] = -irho * gz[4];
  bz[8] = -irho * gz[2];
}

/* ------------------------------------------------------------------------ */
/* Kernels                                                                   */
/* ------------------------------------------------------------------------ */

/* volumeIntegralLinear, VIL:21-53: lduh = -w_i w_j w_k sum_d lFhi_d. */
void ora_volume_integral_linear(const ora_solver* s, double* lduh, const double* lFhi) {
  const int V = s->nvar, B = s->order + 1, o = s->order;
  memset(lduh, 0, sizeof(double) * B * B * B * V);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight =
            ora_table(ORA_T_WEIGHTS, o, i, 0) * ora_table(ORA_T_WEIGHTS, o, j, 0) * ora_table(ORA_T_WEIGHTS, o, k, 0);
        for (int l = 0; l < V; l++)
          for (int m = 0; m < 3; m++)
            lduh[((i * B + j) * B + k) * V + l] -= weight * lFhi[(((m * B + i) * B + j) * B + k) * V + l];
      }
}

/* surfaceIntegralLinear, SIL:21-90: lduh -= (F_right FRCoeff + F_left FLCoeff)
 * w_a w_b / dx per direction (a division by dx, where the nonlinear kernel
 * multiplies by 1/dx). */
void ora_surface_integral_linear(const ora_solver* s, double* lduh, const double* lFhbnd, const double* dx) {
  const int V = s->nvar, B = s->order + 1, o = s->order;
#define LDUH(i, j, k, l) lduh[(((i) * B + (j)) * B + (k)) * V + (l)]
#define FB(f, a, b, l) lFhbnd[(((f) * B + (a)) * B + (b)) * V + (l)]
#define WT(i) ora_table(ORA_T_WEIGHTS, o, i, 0)
#define FRC(i) ora_table(ORA_T_FRCOEFF, o, i, 0)
#define FLC(i) ora_table(ORA_T_FLCOEFF, o, i, 0)
  for (int i = 0; i < B; i++) /* x faces */
    for (int j = 0; j < B; j++) {
      const double weight = WT(i) * WT(j);
      const double updateSize = weight / dx[0];
      for (int k = 0; k < B; k++)
        for (int l = 0; l < V; l++)
          LDUH(i, j, k, l) -= (FB(1, i, j, l) * FRC(k) + FB(0, i, j, l) * FLC(k)) * updateSize;
    }
  for (int i = 0; i < B; i++) /* y faces */
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = WT(i) * WT(k);
        const double updateSize = weight / dx[1];
        for (int l = 0; l < V; l++)
          LDUH(i, j, k, l) -= (FB(3, i, k, l) * FRC(j) + FB(2, i, k, l) * FLC(j)) * updateSize;
      }
  for (int i = 0; i < B; i++) /* z faces */
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = WT(j) * WT(k);
        const double updateSize = weight / dx[2];
        for (int l = 0; l < V; l++)
          LDUH(i, j, k, l) -= (FB(5, j, k, l) * FRC(i) + FB(4, j, k, l) * FLC(i)) * updateSize;
      }
#undef LDUH
#undef FB
#undef WT
#undef FRC
#undef FLC
}

/* riemannSolverLinear, RSL:23-132: Rusanov with scalar dissipation around
 * the coefficient matrix of the averaged face state.  FL = 0.5 (Bn - smax)^T
 * (QR - QL), FR = 0.5 (Bn + smax)^T (QR - QL) (fluctuations, not fluxes).
 * dt is unused (as in the reference). */
void ora_riemann_solver_linear(const ora_solver* s, double* FL, double* FR, const double* QL, const double* QR,
                               double dt, int d) {
  (void)dt;
  const int V = s->nvar, D = s->nvar + s->npar, B = s->order + 1, o = s->order;
  double QavL[32] = {0.0}, QavR[32] = {0.0}, Qav[32];
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      const double weight = ora_table(ORA_T_WEIGHTS, o, i, 0) * ora_table(ORA_T_WEIGHTS, o, j, 0);
      for (int k = 0; k < D; k++) {
        QavL[k] += weight * QL[(i * B + j) * D + k];
        QavR[k] += weight * QR[(i * B + j) * D + k];
      }
    }
  double LL[32] = {0.0}, LR[32] = {0.0};
  ora_eigenvalues(s, QavL, d, LL);
  ora_eigenvalues(s, QavR, d, LR);
  /* max_element of |LL|, of |LR|, then std::max (RSL:73-78) */
  double smax_L = fabs(LL[0]), smax_R = fabs(LR[0]);
  for (int i = 1; i < V; i++) {
    if (smax_L < fabs(LL[i])) smax_L = fabs(LL[i]);
    if (smax_R < fabs(LR[i])) smax_R = fabs(LR[i]);
  }
  const double smax = std_max(smax_L, smax_R);
  for (int k = 0; k < D; k++) Qav[k] = 0.5 * (QavL[k] + QavR[k]);
  double Bn[32 * 32], Bm[32 * 32], Bp[32 * 32];
  ora_coefficient_matrix(s, Qav, d, Bn);
  memcpy(Bm, Bn, sizeof(double) * V * V);
  memcpy(Bp, Bn, sizeof(double) * V * V);
  for (int i = 0; i < V; i++) {
    Bm[i * V + i] -= smax;
    Bp[i * V + i] += smax;
  }
  double* Q = (double*)malloc(sizeof(double) * B * B * D);
  for (int i = 0; i < B * B * D; i++) Q[i] = 0.5 * (QR[i] - QL[i]);
  memset(FL, 0, sizeof(double) * B * B * V);
  memset(FR, 0, sizeof(double) * B * B * V);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < V; k++)
        for (int l = 0; l < V; l++) {
          FL[(i * B + j) * V + k] += Bm[l * V + k] * Q[(i * B + j) * D + l];
          FR[(i * B + j) * V + k] += Bp[l * V + k] * Q[(i * B + j) * D + l];
        }
  free(Q);
}

/* spaceTimePredictorLinear, STPL:34-443, Cauchy-Kowalevski: lQi[0] = luh
 * (+ dt sum_n w_n PSi[n+1]; lQi[1] += PSi[0] with point sources); for each
 * time derivative i < B: gradQ = dudx contractions of lQi[i] / dx per
 * direction, lFi[i] = ncp(lQi[i], gradQ) per direction, lQi[i+1] =
 * -sum_d lFi[i][d], parameters copied; time averages with the Taylor
 * factors dt^i/(i+1)!; boundary extrapolation with FLCoeff / FRCoeff.
 * lQi/lFi/gradQ/lQhi/lFhi are the reference's temporaries (outputs here).
 * PSi may be NULL (usePointSource = false). */
void ora_space_time_predictor_linear(const ora_solver* s, double* lQbnd, double* lFbnd, double* lQi, double* lFi,
                                     double* gradQ, double* lQhi, double* lFhi, const double* luh, const double* dx,
                                     double dt, const double* PSi) {
  const int V = s->nvar, D = s->nvar + s->npar, B = s->order + 1, o = s->order;
  const int B2 = B * B, B3 = B2 * B, B4 = B3 * B;
#define IQ(t, i, j, k, m) (((((long)(t) * B + (i)) * B + (j)) * B + (k)) * D + (m))
#define IPS(t, i, j, k, m) (((((long)(t) * B + (i)) * B + (j)) * B + (k)) * V + (m))
#define IF(t, i, j, k, d, m) ((((((long)(t) * B + (i)) * B + (j)) * B + (k)) * 3 + (d)) * V + (m))
#define IH(i, j, k, m, NN) ((((i) * B + (j)) * B + (k)) * (NN) + (m))
#define IFH(d, i, j, k, m) (((((d) * B + (i)) * B + (j)) * B + (k)) * V + (m))
#define IB(f, a, b, m, NN) ((((f) * B + (a)) * B + (b)) * (NN) + (m))
#define WT(i) ora_table(ORA_T_WEIGHTS, o, i, 0)
#define DUDX(l, n) ora_table(ORA_T_DUDX, o, l, n)
#define FLC(i) ora_table(ORA_T_FLCOEFF, o, i, 0)
#define FRC(i) ora_table(ORA_T_FRCOEFF, o, i, 0)
  memset(lQi, 0, sizeof(double) * (B + 1) * B3 * D);
  for (int i = 0; i < B; i++) /* STPL:62-77 */
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) memcpy(lQi + IQ(0, i, j, k, 0), luh + IH(i, j, k, 0, D), sizeof(double) * D);
  if (PSi) { /* STPL:79-105 */
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++) {
            double sum = 0.0;
            for (int n = 1; n < B + 1; n++) sum = sum + WT(n - 1) * PSi[IPS(n, j, k, l, m)];
            lQi[IQ(0, j, k, l, m)] += dt * sum;
          }
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int n = 0; n < V; n++) lQi[IQ(1, j, k, l, n)] += PSi[IPS(0, j, k, l, n)];
  }
  memset(lFi + (long)B4 * 3 * V, 0, sizeof(double) * B4 * V); /* STPL:107 source block */
  memset(gradQ, 0, sizeof(double) * B4 * 3 * V);
  double BgradQ[3 * 32];
  for (int i = 0; i < B; i++) { /* time derivatives, STPL:118-233 */
    for (int j = 0; j < B; j++) /* x */
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++)
            for (int n = 0; n < B; n++)
              gradQ[IF(i, j, k, l, 0, m)] += 1.0 / dx[0] * lQi[IQ(i, j, k, n, m)] * DUDX(l, n);
    for (int j = 0; j < B; j++) /* y */
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++)
            for (int n = 0; n < B; n++)
              gradQ[IF(i, j, l, k, 1, m)] += 1.0 / dx[1] * lQi[IQ(i, j, n, k, m)] * DUDX(l, n);
    for (int j = 0; j < B; j++) /* z */
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++)
            for (int n = 0; n < B; n++)
              gradQ[IF(i, l, j, k, 2, m)] += 1.0 / dx[2] * lQi[IQ(i, n, j, k, m)] * DUDX(l, n);
    for (int j = 0; j < B; j++) /* fluxes */
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++) {
          ora_ncp_directional(s, lQi + IQ(i, j, k, l, 0), gradQ + IF(i, j, k, l, 0, 0), BgradQ);
          memcpy(lFi + IF(i, j, k, l, 0, 0), BgradQ, sizeof(double) * 3 * V);
        }
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < 3; m++)
            for (int n = 0; n < V; n++) lQi[IQ(i + 1, j, k, l, n)] -= lFi[IF(i, j, k, l, m, n)];
    if (s->npar > 0) /* copy parameters, STPL:221-232 */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++)
            for (int n = V; n < D; n++) lQi[IQ(i + 1, j, k, l, n)] = lQi[IQ(i, j, k, l, n)];
  }
  /* time averages, STPL:235-290 */
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < D; l++) lQhi[IH(i, j, k, l, D)] = lQi[IQ(0, i, j, k, l)];
  double dtavFac = 0.5 * dt;
  for (int i = 1; i < B; i++) {
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++) lQhi[IH(j, k, l, m, D)] += dtavFac * lQi[IQ(i, j, k, l, m)];
    dtavFac *= dt / (i + 2);
  }
  memset(lFhi + (long)B3 * 3 * V, 0, sizeof(double) * B3 * V); /* STPL:256 source block */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++) lFhi[IFH(i, j, k, l, m)] = lFi[IF(0, j, k, l, i, m)];
  dtavFac = 0.5 * dt;
  for (int i = 1; i < B; i++) {
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < B; m++)
            for (int n = 0; n < V; n++) lFhi[IFH(j, k, l, m, n)] += dtavFac * lFi[IF(i, k, l, m, j, n)];
    dtavFac *= dt / (i + 2);
  }
  if (PSi) /* STPL:292-309 */
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < B; m++)
            for (int n = 0; n < V; n++) {
              double sum = 0.0;
              for (int n0 = 1; n0 < B + 1; n0++) sum = sum + WT(n0 - 1) * PSi[IPS(n0, k, l, m, n)];
              lFhi[IFH(j, k, l, m, n)] += sum / 3;
            }
  /* boundary extrapolation, STPL:312-436 */
  memset(lQbnd, 0, sizeof(double) * 6 * B2 * D);
  memset(lFbnd, 0, sizeof(double) * 6 * B2 * V);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      for (int k = 0; k < D; k++)
        for (int l = 0; l < B; l++) {
          lQbnd[IB(0, i, j, k, D)] += lQhi[IH(i, j, l, k, D)] * FLC(l);
          lQbnd[IB(1, i, j, k, D)] += lQhi[IH(i, j, l, k, D)] * FRC(l);
        }
    }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < V; k++)
        for (int l = 0; l < B; l++) {
          lFbnd[IB(0, i, j, k, V)] += lFhi[IFH(0, i, j, l, k)] * FLC(l);
          lFbnd[IB(1, i, j, k, V)] += lFhi[IFH(0, i, j, l, k)] * FRC(l);
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < D; k++)
        for (int l = 0; l < B; l++) {
          lQbnd[IB(2, i, j, k, D)] += lQhi[IH(i, l, j, k, D)] * FLC(l);
          lQbnd[IB(3, i, j, k, D)] += lQhi[IH(i, l, j, k, D)] * FRC(l);
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < V; k++)
        for (int l = 0; l < B; l++) {
          lFbnd[IB(2, i, j, k, V)] += lFhi[IFH(1, i, l, j, k)] * FLC(l);
          lFbnd[IB(3, i, j, k, V)] += lFhi[IFH(1, i, l, j, k)] * FRC(l);
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < D; k++)
        for (int l = 0; l < B; l++) {
          lQbnd[IB(4, i, j, k, D)] += lQhi[IH(l, i, j, k, D)] * FLC(l);
          lQbnd[IB(5, i, j, k, D)] += lQhi[IH(l, i, j, k, D)] * FRC(l);
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < V; k++)
        for (int l = 0; l < B; l++) {
          lFbnd[IB(4, i, j, k, V)] += lFhi[IFH(2, l, i, j, k)] * FLC(l);
          lFbnd[IB(5, i, j, k, V)] += lFhi[IFH(2, l, i, j, k)] * FRC(l);
        }
#undef IQ
#undef IPS
#undef IF
#undef IH
#undef IFH
#undef IB
#undef WT
#undef DUDX
#undef FLC
#undef FRC
}

/* Lagrange basis function n on the Gauss-Legendre nodes of [0,1], at x (the
 * reference's basisFunctions[order][n], DGBasisFunctions.cpp, printed as
 * 15-digit polynomials; the product form agrees with them to ~1e-14). */
static double gl_basis(int o, int n, double x) {
  double v = 1.0;
  for (int k = 0; k <= o; k++)
    if (k != n) v *= (x - ora_table(ORA_T_NODES, o, k, 0)) / (ora_table(ORA_T_NODES, o, n, 0) - ora_table(ORA_T_NODES, o, k, 0));
  return v;
}

/* The projection step of pointSource, K/pointSource.cpph:40-126: the user's
 * pointSource has filled PSi[(n,i,j,k)] (V values, time node n = 0 at t and
 * n >= 1 at t + dt x_{n-1}; space node (z,y,x) = (i,j,k)) and returned the
 * source position x0 for that node (x0[node*3 + {x,y,z}]); each value becomes
 * phi_i(xRef_z) phi_j(xRef_y) phi_k(xRef_x) F / (w_i dz w_j dy w_k dx) with
 * xRef = (x0 - (center - dx/2)) / dx, or zero when x0 lies outside the cell. */
void ora_point_source_projection(const ora_solver* s, double* PSi, const double* x0, const double* center,
                                 const double* dx) {
  const int V = s->nvar, B = s->order + 1, o = s->order;
  for (int n = 0; n < B + 1; n++)
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const long node = (((long)n * B + i) * B + j) * B + k;
          double xr[3];
          for (int d = 0; d < 3; d++) xr[d] = (x0[node * 3 + d] - (center[d] - dx[d] / 2.)) / dx[d];
          const int inside = xr[0] >= 0. && xr[1] >= 0. && xr[2] >= 0. && xr[0] <= 1.0 && xr[1] <= 1.0 && xr[2] <= 1.0;
          for (int m = 0; m < V; m++) {
            double* p = PSi + node * V + m;
            if (inside) {
              double value = gl_basis(o, i, xr[2]) * gl_basis(o, j, xr[1]) * gl_basis(o, k, xr[0]) * (*p);
              value /= (ora_table(ORA_T_WEIGHTS, o, i, 0) * dx[2] * ora_table(ORA_T_WEIGHTS, o, j, 0) * dx[1] *
                        ora_table(ORA_T_WEIGHTS, o, k, 0) * dx[0]);
              *p = value;
            } else {
              *p = 0.0;
            }
          }
        }
}
