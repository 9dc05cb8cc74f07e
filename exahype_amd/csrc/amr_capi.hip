// amr_capi.hip -- C-ABI of the AMR transfer operators and the device-buffer
// helpers their batched forms take.
#include <cstring>
#include <mutex>

#include "../../include/exahype_amd.h"
#include "amr.hpp"
#include "capi_util.hpp"

// Scratch of exa_volume_unknowns_restriction_batch_noalloc, one copy per GPU
// for the life of the process: word 0 is the call's error flag, words
// kNoallocBitmapWord.. the claim bitmap (one bit per destination block).  In
// host code `scratch` only names the symbol; its device address on each GPU
// comes from hipGetSymbolAddress (noalloc_scratch below).
constexpr long kNoallocBitmapWord = 64;                        // 256 bytes: the flag's own line
constexpr long kNoallocBitmapWords = 1L << 18;                 // 1 MiB of bitmap
constexpr long kNoallocMaxBlocks = 32L * kNoallocBitmapWords;  // 8,388,608 destination blocks
static __device__ __align__(256) unsigned scratch[kNoallocBitmapWord + kNoallocBitmapWords];

namespace {

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mutex;
unsigned* g_noalloc_scratch[kNoallocMaxDevices] = {};  // device address of `scratch` per GPU

// Device address of `scratch` on the current GPU, resolved once per GPU.
int noalloc_scratch(unsigned** out) {
  int dev = 0;
  EXA_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoallocMaxDevices) return EXA_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return EXA_E_HIP;
    if (!p) return EXA_E_HIP;
    g_noalloc_scratch[dev] = static_cast<unsigned*>(p);
  }
  *out = g_noalloc_scratch[dev];
  return EXA_OK;
}

// Zeroes the error flag and the first `words` bitmap words (a kernel, not a
// memset, so that the call records no memset node if a caller captures it).
__global__ __launch_bounds__(256) void noalloc_clear_kernel(unsigned* s, long words) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i == 0) s[0] = 0u;
  if (i < words) s[kNoallocBitmapWord + i] = 0u;
}

long ipow3(int e) {
  long r = 1;
  while (e-- > 0) r *= 3;
  return r;
}

// Sizes of a transfer: blocks of B^ndim nodes.
struct AmrSizes {
  int B, D, V;
  long nodes(int ndim) const { return ndim == 3 ? (long)B * B * B : (long)B * B; }
};

bool amr_solver(const exa_solver* s, AmrSizes* z) {
  SolverDesc sd;
  if (!to_desc(s, &sd)) return false;
  z->B = sd.order + 1;
  z->D = sd.nvar + sd.npar;
  z->V = sd.nvar;
  return true;
}

// Runs one launch on the null stream and waits for it; optionally times it
// with HIP events on that stream.  Items the kernel rejected -> EXA_E_INVALID.
int run_amr(exa::AmrArgs a, int ndim, bool restrict_, long ngroups, float* elapsed_ms) {
  DevBuf d_err, d_claim;
  EXA_TRY(d_err.alloc(sizeof(int)));
  EXA_TRY(hipMemset(d_err.p, 0, sizeof(int)));
  a.error = static_cast<int*>(d_err.p);
  const size_t claim_bytes = sizeof(unsigned) * (size_t)((a.dst_blocks + 31) / 32);
  EXA_TRY(d_claim.alloc(claim_bytes));
  EXA_TRY(hipMemset(d_claim.p, 0, claim_bytes ? claim_bytes : 8));
  a.claimed = static_cast<unsigned*>(d_claim.p);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    EXA_TRY(hipEventCreate(&e0));
    EXA_TRY(hipEventCreate(&e1));
    EXA_TRY(hipEventRecord(e0, nullptr));
  }
  const hipError_t le = exa::launch_amr(a, ndim, restrict_, ngroups, nullptr);
  if (elapsed_ms && le == hipSuccess) (void)hipEventRecord(e1, nullptr);
  const hipError_t se = hipDeviceSynchronize();
  if (elapsed_ms) {
    if (le == hipSuccess && se == hipSuccess) (void)hipEventElapsedTime(elapsed_ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  EXA_TRY(le);
  EXA_TRY(se);
  int err = 0;
  EXA_TRY(hipMemcpy(&err, d_err.p, sizeof(int), hipMemcpyDeviceToHost));
  return err ? EXA_E_INVALID : EXA_OK;
}

exa::AmrArgs amr_args(const AmrSizes& z, int ndim, int nv, double* dst, long dst_blocks, const double* src,
                      long src_blocks, const int* items, long nitems, const int* ptr, int level_delta) {
  exa::AmrArgs a{};
  a.dst = dst;
  a.src = src;
  a.dst_block = a.src_block = z.nodes(ndim) * nv;
  a.dst_blocks = dst_blocks;
  a.src_blocks = src_blocks;
  a.items = items;
  a.nitems = nitems;
  a.ptr = ptr;
  a.nv = nv;
  a.order = z.B - 1;
  a.level_delta = level_delta;
  a.ndigits = ipow3(level_delta);
  return a;
}

// One transfer with host arrays (the reference's per-cell call): the blocks
// go to the device, one item, back.
int single_transfer(const AmrSizes& z, int ndim, bool restrict_, int nv, double* dst, const double* src,
                    int level_delta, const int* sub) {
  const long n = z.nodes(ndim) * nv;
  DevBuf d_dst, d_src, d_item, d_ptr;
  EXA_TRY(d_dst.alloc(sizeof(double) * n));
  EXA_TRY(d_src.alloc(sizeof(double) * n));
  EXA_TRY(d_item.alloc(sizeof(int) * 5));
  EXA_TRY(d_ptr.alloc(sizeof(int) * 2));
  const int item[5] = {0, 0, sub[0], sub[1], ndim == 3 ? sub[2] : 0};
  const int ptr[2] = {0, 1};
  EXA_TRY(hipMemcpy(d_item.p, item, sizeof(item), hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_ptr.p, ptr, sizeof(ptr), hipMemcpyHostToDevice));
  EXA_TRY(hipMemcpy(d_src.p, src, sizeof(double) * n, hipMemcpyHostToDevice));
  if (restrict_) EXA_TRY(hipMemcpy(d_dst.p, dst, sizeof(double) * n, hipMemcpyHostToDevice));
  const exa::AmrArgs a = amr_args(z, ndim, nv, d_dst.d(), 1, d_src.d(), 1, static_cast<const int*>(d_item.p), 1,
                                  static_cast<const int*>(d_ptr.p), level_delta);
  const int st = run_amr(a, ndim, restrict_, 1, nullptr);
  if (st != EXA_OK) return st;
  EXA_TRY(hipMemcpy(dst, d_dst.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return EXA_OK;
}

// Level difference and subinterval indices of a single call.  Prolongation
// over zero levels changes nothing (the reference's loop does not run).
int check_single(int coarse_level, int fine_level, const int* sub, int ndim, bool restrict_, int* level_delta) {
  *level_delta = fine_level - coarse_level;
  if (!sub || *level_delta < (restrict_ ? 1 : 0)) return EXA_E_INVALID;
  if (*level_delta > exa::kAmrMaxLevelDelta) return EXA_E_UNSUPPORTED;
  for (int d = 0; d < ndim; ++d)
    if (sub[d] < 0 || sub[d] >= ipow3(*level_delta)) return EXA_E_INVALID;
  return EXA_OK;
}

}  // namespace

namespace exa {
int noalloc_scratch_setup() {
  unsigned* p = nullptr;
  return noalloc_scratch(&p);
}
}  // namespace exa

extern "C" int exa_volume_unknowns_restriction_batch_noalloc(const exa_solver* s, double* luhCoarse, long ncoarse,
                                                             const double* luhFine, long nfine, const int* ptr,
                                                             long ndst, const int* items, long nitems,
                                                             int levelDelta) {
  AmrSizes z;
  if (!amr_solver(s, &z) || !luhFine || !luhCoarse || (ndst > 0 && (!items || !ptr)) || ndst < 0 || nitems < 0 ||
      nfine < 0 || ncoarse < 0 || levelDelta < 1)
    return EXA_E_INVALID;
  if (levelDelta > exa::kAmrMaxLevelDelta) return EXA_E_UNSUPPORTED;
  if (ncoarse > kNoallocMaxBlocks) return EXA_E_INVALID;  // beyond the claim bitmap
  unsigned* sp = nullptr;
  if (const int st = noalloc_scratch(&sp)) return st;
  const long words = (ncoarse + 31) / 32;
  hipLaunchKernelGGL(noalloc_clear_kernel, dim3((unsigned)((words > 1 ? words : 1) + 255) / 256), dim3(256), 0,
                     nullptr, sp, words);
  EXA_TRY(hipGetLastError());
  exa::AmrArgs a = amr_args(z, 3, z.D, luhCoarse, ncoarse, luhFine, nfine, items, nitems, ptr, levelDelta);
  a.error = reinterpret_cast<int*>(sp);
  a.claimed = sp + kNoallocBitmapWord;
  EXA_TRY(exa::launch_amr(a, 3, true, ndst, nullptr));
  EXA_TRY(hipDeviceSynchronize());
  unsigned err = 0;
  EXA_TRY(hipMemcpy(&err, sp, sizeof(unsigned), hipMemcpyDeviceToHost));
  return err ? EXA_E_INVALID : EXA_OK;
}

extern "C" void* exa_device_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
  return p;
}

extern "C" int exa_device_free(void* p) {
  EXA_TRY(hipFree(p));
  return EXA_OK;
}

extern "C" int exa_copy_to_device(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return EXA_E_INVALID;
  EXA_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return EXA_OK;
}

extern "C" int exa_copy_to_host(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return EXA_E_INVALID;
  EXA_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_face_unknowns_prolongation(const exa_solver* s, double* lQhbndFine, double* lFhbndFine,
                                              const double* lQhbndCoarse, const double* lFhbndCoarse,
                                              int coarseGridLevel, int fineGridLevel, const int* subfaceIndex) {
  AmrSizes z;
  int dl = 0;
  if (!amr_solver(s, &z) || !lQhbndFine || !lFhbndFine || !lQhbndCoarse || !lFhbndCoarse) return EXA_E_INVALID;
  if (const int st = check_single(coarseGridLevel, fineGridLevel, subfaceIndex, 2, false, &dl)) return st;
  if (dl == 0) return EXA_OK;
  if (const int st = single_transfer(z, 2, false, z.D, lQhbndFine, lQhbndCoarse, dl, subfaceIndex)) return st;
  return single_transfer(z, 2, false, z.V, lFhbndFine, lFhbndCoarse, dl, subfaceIndex);
}

extern "C" int exa_face_unknowns_restriction(const exa_solver* s, double* lQhbndCoarse, double* lFhbndCoarse,
                                             const double* lQhbndFine, const double* lFhbndFine,
                                             int coarseGridLevel, int fineGridLevel, const int* subfaceIndex) {
  AmrSizes z;
  int dl = 0;
  if (!amr_solver(s, &z) || !lQhbndFine || !lFhbndFine || !lQhbndCoarse || !lFhbndCoarse) return EXA_E_INVALID;
  if (const int st = check_single(coarseGridLevel, fineGridLevel, subfaceIndex, 2, true, &dl)) return st;
  if (const int st = single_transfer(z, 2, true, z.D, lQhbndCoarse, lQhbndFine, dl, subfaceIndex)) return st;
  return single_transfer(z, 2, true, z.V, lFhbndCoarse, lFhbndFine, dl, subfaceIndex);
}

extern "C" int exa_volume_unknowns_prolongation(const exa_solver* s, double* luhFine, const double* luhCoarse,
                                                int coarseGridLevel, int fineGridLevel, const int* subcellIndex) {
  AmrSizes z;
  int dl = 0;
  if (!amr_solver(s, &z) || !luhFine || !luhCoarse) return EXA_E_INVALID;
  if (const int st = check_single(coarseGridLevel, fineGridLevel, subcellIndex, 3, false, &dl)) return st;
  if (dl == 0) return EXA_OK;
  return single_transfer(z, 3, false, z.D, luhFine, luhCoarse, dl, subcellIndex);
}

extern "C" int exa_volume_unknowns_restriction(const exa_solver* s, double* luhCoarse, const double* luhFine,
                                               int coarseGridLevel, int fineGridLevel, const int* subcellIndex) {
  AmrSizes z;
  int dl = 0;
  if (!amr_solver(s, &z) || !luhFine || !luhCoarse) return EXA_E_INVALID;
  if (const int st = check_single(coarseGridLevel, fineGridLevel, subcellIndex, 3, true, &dl)) return st;
  return single_transfer(z, 3, true, z.D, luhCoarse, luhFine, dl, subcellIndex);
}

extern "C" int exa_volume_unknowns_prolongation_batch(const exa_solver* s, double* luhFine, long nfine,
                                                      const double* luhCoarse, long ncoarse, const int* items,
                                                      long nitems, int levelDelta, float* elapsed_ms) {
  AmrSizes z;
  if (!amr_solver(s, &z) || !luhFine || !luhCoarse || (nitems > 0 && !items) || nitems < 0 || nfine < 0 ||
      ncoarse < 0 || levelDelta < 1)
    return EXA_E_INVALID;
  if (levelDelta > exa::kAmrMaxLevelDelta) return EXA_E_UNSUPPORTED;
  const exa::AmrArgs a = amr_args(z, 3, z.D, luhFine, nfine, luhCoarse, ncoarse, items, nitems, nullptr, levelDelta);
  return run_amr(a, 3, false, nitems, elapsed_ms);
}

extern "C" int exa_volume_unknowns_restriction_batch(const exa_solver* s, double* luhCoarse, long ncoarse,
                                                     const double* luhFine, long nfine, const int* ptr, long ndst,
                                                     const int* items, long nitems, int levelDelta,
                                                     float* elapsed_ms) {
  AmrSizes z;
  if (!amr_solver(s, &z) || !luhFine || !luhCoarse || (ndst > 0 && (!items || !ptr)) || ndst < 0 || nitems < 0 ||
      nfine < 0 || ncoarse < 0 || levelDelta < 1)
    return EXA_E_INVALID;
  if (levelDelta > exa::kAmrMaxLevelDelta) return EXA_E_UNSUPPORTED;
  const exa::AmrArgs a = amr_args(z, 3, z.D, luhCoarse, ncoarse, luhFine, nfine, items, nitems, ptr, levelDelta);
  return run_amr(a, 3, true, ndst, elapsed_ms);
}

extern "C" int exa_face_unknowns_prolongation_batch(const exa_solver* s, double* lQhbndFine, double* lFhbndFine,
                                                    long nfine_faces, const double* lQhbndCoarse,
                                                    const double* lFhbndCoarse, long ncoarse_faces,
                                                    const int* items, long nitems, int levelDelta,
                                                    float* elapsed_ms) {
  AmrSizes z;
  if (!amr_solver(s, &z) || !lQhbndFine || !lFhbndFine || !lQhbndCoarse || !lFhbndCoarse ||
      (nitems > 0 && !items) || nitems < 0 || nfine_faces < 0 || ncoarse_faces < 0 || levelDelta < 1)
    return EXA_E_INVALID;
  if (levelDelta > exa::kAmrMaxLevelDelta) return EXA_E_UNSUPPORTED;
  float t1 = 0.0f, t2 = 0.0f;
  exa::AmrArgs a = amr_args(z, 2, z.D, lQhbndFine, nfine_faces, lQhbndCoarse, ncoarse_faces, items, nitems,
                            nullptr, levelDelta);
  if (const int st = run_amr(a, 2, false, nitems, elapsed_ms ? &t1 : nullptr)) return st;
  a = amr_args(z, 2, z.V, lFhbndFine, nfine_faces, lFhbndCoarse, ncoarse_faces, items, nitems, nullptr,
               levelDelta);
  if (const int st = run_amr(a, 2, false, nitems, elapsed_ms ? &t2 : nullptr)) return st;
  if (elapsed_ms) *elapsed_ms = t1 + t2;
  return EXA_OK;
}

extern "C" int exa_face_unknowns_restriction_batch(const exa_solver* s, double* lQhbndCoarse, double* lFhbndCoarse,
                                                   long ncoarse_faces, const double* lQhbndFine,
                                                   const double* lFhbndFine, long nfine_faces, const int* ptr,
                                                   long ndst, const int* items, long nitems, int levelDelta,
                                                   float* elapsed_ms) {
  AmrSizes z;
  if (!amr_solver(s, &z) || !lQhbndFine || !lFhbndFine || !lQhbndCoarse || !lFhbndCoarse ||
      (ndst > 0 && (!items || !ptr)) || ndst < 0 || nitems < 0 || nfine_faces < 0 || ncoarse_faces < 0 ||
      levelDelta < 1)
    return EXA_E_INVALID;
  if (levelDelta > exa::kAmrMaxLevelDelta) return EXA_E_UNSUPPORTED;
  float t1 = 0.0f, t2 = 0.0f;
  exa::AmrArgs a = amr_args(z, 2, z.D, lQhbndCoarse, ncoarse_faces, lQhbndFine, nfine_faces, items, nitems,
                            ptr, levelDelta);
  if (const int st = run_amr(a, 2, true, ndst, elapsed_ms ? &t1 : nullptr)) return st;
  a = amr_args(z, 2, z.V, lFhbndCoarse, ncoarse_faces, lFhbndFine, nfine_faces, items, nitems, ptr,
               levelDelta);
  if (const int st = run_amr(a, 2, true, ndst, elapsed_ms ? &t2 : nullptr)) return st;
  if (elapsed_ms) *elapsed_ms = t1 + t2;
  return EXA_OK;
}
