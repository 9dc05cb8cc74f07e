/*
 * output.c -- CPU restatement of the interpolation step of the reference's
 * Cartesian VTK plotter for ADER-DG cells.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 * ADERDG2CartesianVTK::plotVertexData
 * (ExaHyPE/exahype/plotters/VTK/ADERDG2CartesianVTK.cpp:288-336): the cell's
 * values at the (N+1)^3 equidistant vertices offset + i h / N, interpolated
 * from the Gauss-Legendre nodal values with equidistantGridProjector1d
 * (ExaHyPE/kernels/DGMatrices.cpp, printed literals; here exa_equi_proj,
 * [order][GL node n][equidistant point m]).  The identity post-processing
 * (every stored value written) is assumed.
 */
#include "oracle.h"

#define EXA_TABLE static const
#include "../include/exahype_amd/device/dg_tables.inc"
#undef EXA_TABLE

double ora_equi_proj(int order, int n, int m) { return exa_equi_proj[order][n][m]; }

/* luh [ncells][B^3][D] -> out [ncells][B^3][D], vertices (i2, i1, i0) with i0
 * (x) fastest, the reference's dfor order and product order. */
void ora_plot_vertices(const ora_solver* s, long ncells, const double* luh, double* out) {
  const int B = s->order + 1, D = s->nvar + s->npar, o = s->order;
  const long B3 = (long)B * B * B;
  for (long c = 0; c < ncells; c++) {
    const double* u = luh + c * B3 * D;
    for (int i2 = 0; i2 < B; i2++)
      for (int i1 = 0; i1 < B; i1++)
        for (int i0 = 0; i0 < B; i0++)
          for (int v = 0; v < D; v++) {
            double interp = 0.0;
            for (int ii2 = 0; ii2 < B; ii2++)
              for (int ii1 = 0; ii1 < B; ii1++)
                for (int ii0 = 0; ii0 < B; ii0++)
                  interp += exa_equi_proj[o][ii1][i1] * exa_equi_proj[o][ii0][i0] * exa_equi_proj[o][ii2][i2] *
                            u[((ii2 * B + ii1) * B + ii0) * D + v];
            out[c * B3 * D + ((i2 * B + i1) * B + i0) * D + v] = interp;
          }
  }
}
