// amr_kernels.hip -- ADER-DG AMR transfer operators on gfx950, FP64: face and
// volume unknowns prolongation (coarse -> fine) and restriction (fine ->
// coarse, accumulated), over one or more levels.
//
// Reference semantics (K = ExaHyPE/kernels/aderdg/generic/c/3d):
//   K/amrRoutines.cpph:23-143   faceUnknownsProlongation (+ single level)
//   K/amrRoutines.cpph:145-260  faceUnknownsRestriction
//   K/amrRoutines.cpph:285-384  volumeUnknownsProlongation
//   K/amrRoutines.cpph:386-491  volumeUnknownsRestriction
//
// One workgroup per transfer (restriction: per destination, its transfers in
// list order), one thread per node of the B^NDIM block (padded to whole
// wavefronts).  A level is the tensor product of three (two for
// faces) 1-D operators; it runs as NDIM passes through LDS, each a B-term
// contraction along one direction, instead of the reference's B^NDIM-term sum
// per node (the same linear map; the sums round in another order).  The 1-D
// operator of direction d at a level is fineGridProjector1d[j_d] (prolongation:
// out_m = sum_n in_n P[j][n][m]) or w_n P[j][m][n] / (3 w_m) (restriction), j_d
// the subinterval digit of that level.  Restriction adds the restricted blocks
// of a destination's transfers, in list order, to the destination block.
#include "amr.hpp"
#include "exa_device.hpp"

namespace exa {

template <int B, int NDIM, bool RESTRICT>
__global__ __launch_bounds__(round_up(NDIM == 3 ? B * B * B : B * B, kWave)) void amr_kernel(AmrArgs a) {
  constexpr int N = B - 1;
  constexpr int NN = NDIM == 3 ? B * B * B : B * B;
  constexpr int NT = round_up(NN, kWave);
  extern __shared__ double smem[];
  const int nv = a.nv;
  double* sA = smem;
  double* sB = sA + NN * nv;
  double* sAcc = sB + NN * nv;                   // restriction: the destination block
  double* sM = sAcc + (RESTRICT ? NN * nv : 0);  // [NDIM][B][B] 1-D operators of a level
  const int tid = threadIdx.x;
  const bool active = tid < NN;
  const int node = active ? tid : 0;
  const int c[3] = {node % B, (node / B) % B, node / (B * B)};
  // Workgroups are dealt to the 8 XCDs round-robin; each XCD takes one
  // contiguous run of transfers, so the children of a coarse block (adjacent
  // in the list) read it through the same L2.  A bijection on [0, gridDim).
  const long g = gridDim.x, b = blockIdx.x, q = g / 8, r = g % 8, xcd = b % 8, idx = b / 8;
  const long k = xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
  const long i0 = RESTRICT ? a.ptr[k] : k;
  const long i1 = RESTRICT ? a.ptr[k + 1] : k + 1;
  if (i0 < 0 || i1 < i0 || i1 > a.nitems) {  // a malformed range: nothing is read
    if (tid == 0) atomicOr(a.error, 1);
    return;
  }
  const long ndig = a.ndigits;  // 3^level_delta
  long dst_blk = -1;
  for (long it = i0; it < i1; ++it) {
    const int* item = a.items + 5 * it;
    const long db = item[0], sb = item[1];
    bool ok = db >= 0 && db < a.dst_blocks && sb >= 0 && sb < a.src_blocks &&
              (!RESTRICT || dst_blk < 0 || db == dst_blk);
#pragma unroll
    for (int d = 0; d < NDIM; ++d) ok = ok && item[2 + d] >= 0 && item[2 + d] < ndig;
    if (!ok) {  // uniform over the workgroup: the item is skipped and reported
      if (tid == 0) atomicOr(a.error, 1);
      continue;
    }
    if (!RESTRICT || dst_blk < 0) {
      // claim the destination block: a second workgroup naming it writes nothing
      __shared__ unsigned prev;
      if (tid == 0) prev = atomicOr(a.claimed + (db >> 5), 1u << (db & 31));
      __syncthreads();
      if (prev & (1u << (db & 31))) {
        if (tid == 0) atomicOr(a.error, 1);
        return;  // uniform: every thread read the same prev
      }
    }
    if (RESTRICT && dst_blk < 0) {
      dst_blk = db;
      const double* dst = a.dst + db * a.dst_block;
      for (int e = tid; e < NN * nv; e += NT) sAcc[e] = dst[e];
    }
    const double* src = a.src + sb * a.src_block;
    for (int e = tid; e < NN * nv; e += NT) sA[e] = src[e];
    double* in = sA;
    double* out = sB;
    long sig = 1;  // prolongation decodes from the top digit, restriction from the bottom
    if (!RESTRICT)
      for (int l = 1; l < a.level_delta; ++l) sig *= 3;
    for (int l = 0; l < a.level_delta; ++l) {
      __syncthreads();  // the previous level's operator no longer read; the input block complete
      for (int e = tid; e < NDIM * B * B; e += NT) {
        const int d = e / (B * B), n = (e / B) % B, m = e % B;
        const int j = (int)((item[2 + d] / sig) % 3);
        sM[e] = RESTRICT ? exa_gl_weights[N][n] * exa_fine_proj[N][j][m][n] / exa_gl_weights[N][m] / 3.0
                         : exa_fine_proj[N][j][n][m];
      }
      __syncthreads();
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        const int stride = d == 0 ? 1 : (d == 1 ? B : B * B);
        const int base = node - c[d] * stride;
        const double* M = sM + d * B * B;
        if (active) {
          for (int v = 0; v < nv; ++v) {
            double s = 0.0;
#pragma unroll
            for (int n = 0; n < B; ++n) s += in[(base + n * stride) * nv + v] * M[n * B + c[d]];
            out[node * nv + v] = s;
          }
        }
        __syncthreads();
        double* t = in;
        in = out;
        out = t;
      }
      sig = RESTRICT ? sig * 3 : sig / 3;
    }
    if (!RESTRICT) {
      double* dst = a.dst + db * a.dst_block;
      for (int e = tid; e < NN * nv; e += NT) dst[e] = in[e];
    } else {  // added in list order, as successive single calls add
      for (int e = tid; e < NN * nv; e += NT) sAcc[e] += in[e];
    }
    __syncthreads();  // sA, sB free for the next transfer
  }
  if (RESTRICT && dst_blk >= 0) {
    double* dst = a.dst + dst_blk * a.dst_block;
    for (int e = tid; e < NN * nv; e += NT) dst[e] = sAcc[e];
  }
}

template <int NDIM, bool RESTRICT>
static hipError_t launch_amr_b(const AmrArgs& a, long ngroups, hipStream_t s) {
  auto go = [&]<int B>() -> hipError_t {
    constexpr int NN = NDIM == 3 ? B * B * B : B * B;
    const size_t lds = sizeof(double) * ((RESTRICT ? 3 : 2) * (size_t)NN * a.nv + NDIM * B * B);
    if (lds > 64 * 1024) return hipErrorNotSupported;
    hipLaunchKernelGGL((amr_kernel<B, NDIM, RESTRICT>), dim3((unsigned)ngroups), dim3(round_up(NN, kWave)), lds, s,
                       a);
    return hipGetLastError();
  };
  switch (a.order) {
    case 1: return go.template operator()<2>();
    case 2: return go.template operator()<3>();
    case 3: return go.template operator()<4>();
    case 4: return go.template operator()<5>();
    case 5: return go.template operator()<6>();
    default: return hipErrorNotSupported;
  }
}

hipError_t launch_amr(const AmrArgs& a, int ndim, bool restrict_, long ngroups, hipStream_t s) {
  if (ngroups <= 0) return hipSuccess;
  if (a.nv < 1 || a.level_delta < 1 || a.level_delta > kAmrMaxLevelDelta) return hipErrorInvalidValue;
  if (ndim == 3) return restrict_ ? launch_amr_b<3, true>(a, ngroups, s) : launch_amr_b<3, false>(a, ngroups, s);
  if (ndim == 2) return restrict_ ? launch_amr_b<2, true>(a, ngroups, s) : launch_amr_b<2, false>(a, ngroups, s);
  return hipErrorInvalidValue;
}

}  // namespace exa
