// output_kernels.hip -- the interpolation step of the Cartesian VTK plotter on
// gfx950, FP64: every cell's values at its (N+1)^3 equidistant vertices
// (ADERDG2CartesianVTK::plotVertexData, exahype/plotters/VTK/
// ADERDG2CartesianVTK.cpp:288-336), computed where the solution lives.
//
// One workgroup per cell, one thread per vertex.  The reference sums the
// B^3 nodal values per vertex with the product of three 1-D projectors
// (equidistantGridProjector1d); here the same tensor product runs as three
// B-term passes through LDS (the same linear map, summed in another order).
#include "aderdg.hpp"
#include "exa_device.hpp"

namespace exa {

template <int B>
__global__ __launch_bounds__(node_threads(B)) void plot_vertices_kernel(const double* luh, double* out, int D,
                                                                       long cell0) {
  constexpr int N = B - 1, B3 = B * B * B, NT = node_threads(B);
  extern __shared__ double smem[];
  double* sA = smem;
  double* sB = sA + B3 * D;
  double* sP = sB + B3 * D;  // [n][m]
  const int tid = threadIdx.x;
  const bool active = tid < B3;
  const int node = active ? tid : 0;
  const int c[3] = {node % B, (node / B) % B, node / (B * B)};
  const size_t cell = (size_t)blockIdx.x + cell0;
  const double* u = luh + cell * (size_t)(B3 * D);
  for (int e = tid; e < B3 * D; e += NT) sA[e] = u[e];
  for (int e = tid; e < B * B; e += NT) sP[e] = exa_equi_proj[N][e / B][e % B];
  __syncthreads();
  double* in = sA;
  double* o = sB;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int stride = d == 0 ? 1 : (d == 1 ? B : B * B);
    const int base = node - c[d] * stride;
    if (active) {
      for (int v = 0; v < D; ++v) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < B; ++n) s += in[(base + n * stride) * D + v] * sP[n * B + c[d]];
        o[node * D + v] = s;
      }
    }
    __syncthreads();
    double* t = in;
    in = o;
    o = t;
  }
  double* dst = out + cell * (size_t)(B3 * D);
  for (int e = tid; e < B3 * D; e += NT) dst[e] = in[e];
}

hipError_t launch_plot_vertices(int order, int D, const double* luh, double* out, long ncells, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (D < 1) return hipErrorInvalidValue;
  auto go = [&]<int B>() -> hipError_t {
    const size_t lds = sizeof(double) * (2 * (size_t)B * B * B * D + B * B);
    if (lds > 64 * 1024) return hipErrorNotSupported;
    hipLaunchKernelGGL((plot_vertices_kernel<B>), dim3((unsigned)ncells), dim3(node_threads(B)), lds, s, luh, out, D,
                       0L);
    return hipGetLastError();
  };
  switch (order) {
    case 1: return go.template operator()<2>();
    case 2: return go.template operator()<3>();
    case 3: return go.template operator()<4>();
    case 4: return go.template operator()<5>();
    case 5: return go.template operator()<6>();
    case 6: return go.template operator()<7>();
    default: return hipErrorNotSupported;
  }
}

}  // namespace exa
