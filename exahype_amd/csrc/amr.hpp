// amr.hpp -- internal interface of the AMR transfer kernels (face and volume
// unknowns prolongation and restriction).
#pragma once

#include <hip/hip_runtime.h>

namespace exa {

constexpr int kAmrMaxLevelDelta = 6;  // 3^6 = 729 subintervals per direction

// A batch of transfers between blocks of B^NDIM nodes x nv values (volumes:
// (z,y,x,var); faces: (b,a,var)).  items: 5 ints per transfer (destination
// block, source block, subinterval index per direction in [0, 3^level_delta);
// faces use the first two).  Prolongation: one workgroup per item, the
// destination overwritten.  Restriction: one workgroup per destination k over
// items [ptr[k], ptr[k+1]), which must share one destination block; their
// restricted blocks are added to it in list order.  Items with an index out of
// range (or a mismatched destination) are skipped and set *error.  Every
// destination block may appear once per launch (one item of a prolongation,
// one range of a restriction): each workgroup claims its destination's bit in
// `claimed` (dst_blocks bits, zero on entry) before it writes, and a workgroup
// that finds the bit taken writes nothing and sets *error.
struct AmrArgs {
  double* dst;
  const double* src;
  long dst_block, src_block;    // doubles per block
  long dst_blocks, src_blocks;  // blocks in each array (bounds of the item indices)
  const int* items;
  long nitems;     // items in the list (bounds of the ptr ranges)
  const int* ptr;
  int* error;
  unsigned* claimed;  // one bit per destination block
  int nv;
  int order;
  int level_delta;
  long ndigits;  // 3^level_delta
};

// ndim 3 (volumes) or 2 (faces); ngroups = items (prolongation) or
// destinations (restriction).
hipError_t launch_amr(const AmrArgs& a, int ndim, bool restrict_, long ngroups, hipStream_t s);

// Resolves the current GPU's address of the allocation-free restriction's
// scratch (amr_capi.hip); called by exa_init.
int noalloc_scratch_setup();

}  // namespace exa
