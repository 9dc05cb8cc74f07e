/*
 * amr.c -- CPU restatement of the reference's 3-D ADER-DG AMR routines: face
 * and volume unknowns prolongation (coarse -> fine) and restriction (fine ->
 * coarse, accumulated) over one or more levels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 * Citations (paths relative to the reference checkout):
 *   AMR = ExaHyPE/kernels/aderdg/generic/c/3d/amrRoutines.cpph
 *   P   = fineGridProjector1d[order][j][n][m] (ExaHyPE/kernels/DGMatrices.cpp,
 *         printed literals): coarse basis function n at fine node m of
 *         subinterval j, here exa_fine_proj.
 * Layouts: faces (a, b, var) B^2 NV; volumes (z, y, x, var) B^3 NV.  A
 * subinterval index in [0, 3^levelDelta) per direction: faces (x-like, y-like)
 * = the face's two in-plane directions in their storage order, volumes (x, y, z).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define EXA_TABLE static const
#include "../include/exahype_amd/device/dg_tables.inc"
#undef EXA_TABLE

double ora_fine_proj(int order, int j, int n, int m) { return exa_fine_proj[order][j][n][m]; }

/* 3^e */
static int pow3(int e) {
  int r = 1;
  while (e-- > 0) r *= 3;
  return r;
}

/* singleLevelFaceUnknownsProlongation, AMR:23-46 (fine += ...; the caller
 * zeroes fine). */
static void face_prolong_1(int NV, int B, int o, double* fine, const double* coarse, const int j[2]) {
  for (int m2 = 0; m2 < B; ++m2)
    for (int m1 = 0; m1 < B; ++m1)
      for (int v = 0; v < NV; ++v)
        for (int n2 = 0; n2 < B; ++n2)
          for (int n1 = 0; n1 < B; ++n1)
            fine[(m2 * B + m1) * NV + v] +=
                coarse[(n2 * B + n1) * NV + v] * exa_fine_proj[o][j[0]][n1][m1] * exa_fine_proj[o][j[1]][n2][m2];
}

/* singleLevelFaceUnknownsRestriction, AMR:145-172. */
static void face_restrict_1(int NV, int B, int o, double* coarse, const double* fine, const int j[2]) {
  const double* w = exa_gl_weights[o];
  for (int m2 = 0; m2 < B; ++m2)
    for (int m1 = 0; m1 < B; ++m1)
      for (int v = 0; v < NV; ++v)
        for (int n2 = 0; n2 < B; ++n2)
          for (int n1 = 0; n1 < B; ++n1)
            coarse[(m2 * B + m1) * NV + v] += w[n1] * w[n2] * exa_fine_proj[o][j[0]][m1][n1] *
                                              exa_fine_proj[o][j[1]][m2][n2] * fine[(n2 * B + n1) * NV + v] /
                                              w[m1] / w[m2] / 9.0;
}

/* singleLevelVolumeUnknownsProlongation, AMR:285-311. */
static void volume_prolong_1(int NV, int B, int o, double* fine, const double* coarse, const int j[3]) {
  for (int m3 = 0; m3 < B; ++m3)
    for (int m2 = 0; m2 < B; ++m2)
      for (int m1 = 0; m1 < B; ++m1)
        for (int v = 0; v < NV; ++v)
          for (int n3 = 0; n3 < B; ++n3)
            for (int n2 = 0; n2 < B; ++n2)
              for (int n1 = 0; n1 < B; ++n1)
                fine[((m3 * B + m2) * B + m1) * NV + v] += coarse[((n3 * B + n2) * B + n1) * NV + v] *
                                                           exa_fine_proj[o][j[0]][n1][m1] *
                                                           exa_fine_proj[o][j[1]][n2][m2] *
                                                           exa_fine_proj[o][j[2]][n3][m3];
}

/* singleLevelVolumeUnknownsRestriction, AMR:386-418. */
static void volume_restrict_1(int NV, int B, int o, double* coarse, const double* fine, const int j[3]) {
  const double* w = exa_gl_weights[o];
  for (int m3 = 0; m3 < B; ++m3)
    for (int m2 = 0; m2 < B; ++m2)
      for (int m1 = 0; m1 < B; ++m1)
        for (int v = 0; v < NV; ++v)
          for (int n3 = 0; n3 < B; ++n3)
            for (int n2 = 0; n2 < B; ++n2)
              for (int n1 = 0; n1 < B; ++n1)
                coarse[((m3 * B + m2) * B + m1) * NV + v] +=
                    w[n1] * w[n2] * w[n3] * exa_fine_proj[o][j[0]][m1][n1] * exa_fine_proj[o][j[1]][m2][n2] *
                    exa_fine_proj[o][j[2]][m3][n3] * fine[((n3 * B + n2) * B + n1) * NV + v] / w[m1] / w[m2] /
                    w[m3] / 27.0;
}

/* The digits of a subinterval index over levelDelta levels.  Prolongation
 * takes them from the most significant (the level just below the coarse
 * cell, AMR:91-99); restriction from the least significant (the level just
 * above the fine cell, AMR:221-228). */
static void digits(int idx, int level_delta, int l, int from_top, int* out) {
  const int sig = from_top ? pow3(level_delta - 1 - l) : pow3(l);
  *out = (idx / sig) % 3;
}

/* A multi-level operator: level l maps buf_l -> buf_{l+1} with the single
 * level operator (targets zeroed first, AMR:101-103), the first level reading
 * the input; prolongation writes the last level into the output, restriction
 * adds it to the output (AMR:257-259). */
typedef void (*single_fn)(int NV, int B, int o, double* dst, const double* src, const int* j);

static void multi_level(single_fn f, int ndim, int NV, int B, int o, int n, double* out, const double* in,
                        int level_delta, const int* idx, int from_top, int accumulate) {
  if (level_delta < 1) return;
  double* a = (double*)calloc((size_t)n, sizeof(double));
  double* b = (double*)calloc((size_t)n, sizeof(double));
  const double* src = in;
  for (int l = 0; l < level_delta; ++l) {
    int j[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d) digits(idx[d], level_delta, l, from_top, &j[d]);
    double* dst = (l % 2 == 0) ? a : b;
    memset(dst, 0, sizeof(double) * n);
    f(NV, B, o, dst, src, j);
    src = dst;
  }
  for (int i = 0; i < n; ++i) out[i] = accumulate ? out[i] + src[i] : src[i];
  free(a);
  free(b);
}

static void fp1(int NV, int B, int o, double* d, const double* s, const int* j) { face_prolong_1(NV, B, o, d, s, j); }
static void fr1(int NV, int B, int o, double* d, const double* s, const int* j) { face_restrict_1(NV, B, o, d, s, j); }
static void vp1(int NV, int B, int o, double* d, const double* s, const int* j) { volume_prolong_1(NV, B, o, d, s, j); }
static void vr1(int NV, int B, int o, double* d, const double* s, const int* j) { volume_restrict_1(NV, B, o, d, s, j); }

/* faceUnknownsProlongation, AMR:48-143: lQhbnd with V+P values per node,
 * lFhbnd with V; the fine arrays are overwritten. */
void ora_face_unknowns_prolongation(const ora_solver* s, double* lQhbndFine, double* lFhbndFine,
                                    const double* lQhbndCoarse, const double* lFhbndCoarse, int coarse_level,
                                    int fine_level, const int* subface) {
  const int B = s->order + 1, D = s->nvar + s->npar, V = s->nvar, o = s->order;
  const int dl = fine_level - coarse_level;
  multi_level(fp1, 2, D, B, o, B * B * D, lQhbndFine, lQhbndCoarse, dl, subface, 1, 0);
  multi_level(fp1, 2, V, B, o, B * B * V, lFhbndFine, lFhbndCoarse, dl, subface, 1, 0);
}

/* faceUnknownsRestriction, AMR:183-260: the restricted values are added to
 * the coarse arrays. */
void ora_face_unknowns_restriction(const ora_solver* s, double* lQhbndCoarse, double* lFhbndCoarse,
                                   const double* lQhbndFine, const double* lFhbndFine, int coarse_level,
                                   int fine_level, const int* subface) {
  const int B = s->order + 1, D = s->nvar + s->npar, V = s->nvar, o = s->order;
  const int dl = fine_level - coarse_level;
  multi_level(fr1, 2, D, B, o, B * B * D, lQhbndCoarse, lQhbndFine, dl, subface, 0, 1);
  multi_level(fr1, 2, V, B, o, B * B * V, lFhbndCoarse, lFhbndFine, dl, subface, 0, 1);
}

/* volumeUnknownsProlongation, AMR:313-384 (luhFine overwritten). */
void ora_volume_unknowns_prolongation(const ora_solver* s, double* luhFine, const double* luhCoarse,
                                      int coarse_level, int fine_level, const int* subcell) {
  const int B = s->order + 1, D = s->nvar + s->npar;
  multi_level(vp1, 3, D, B, s->order, B * B * B * D, luhFine, luhCoarse, fine_level - coarse_level, subcell, 1, 0);
}

/* volumeUnknownsRestriction, AMR:420-491 (added to luhCoarse). */
void ora_volume_unknowns_restriction(const ora_solver* s, double* luhCoarse, const double* luhFine,
                                     int coarse_level, int fine_level, const int* subcell) {
  const int B = s->order + 1, D = s->nvar + s->npar;
  multi_level(vr1, 3, D, B, s->order, B * B * B * D, luhCoarse, luhFine, fine_level - coarse_level, subcell, 0, 1);
}
