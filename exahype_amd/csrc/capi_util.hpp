// capi_util.hpp -- helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/exahype_amd.h"
#include "aderdg.hpp"

using exa::FaceDesc;
using exa::SolverDesc;

namespace {

// PNPM[order], stableTimeStepSize.cpph:33-34
const double kPNPM[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};

int from_hip(hipError_t e) {
  if (e == hipSuccess) return EXA_OK;
  if (e == hipErrorNotSupported) return EXA_E_UNSUPPORTED;
  if (e == hipErrorOutOfMemory) return EXA_E_NOMEM;
  return EXA_E_HIP;
}

#define EXA_TRY(expr)                     \
  do {                                    \
    const hipError_t e_ = (expr);         \
    if (e_ != hipSuccess) return from_hip(e_); \
  } while (0)

bool to_desc(const exa_solver* s, SolverDesc* d) {
  if (!s) return false;
  d->pde = s->pde;
  d->order = s->order;
  d->nvar = s->num_variables;
  d->npar = s->num_parameters;
  d->terms = s->terms;
  d->no_time_averaging = s->no_time_averaging;
  d->cfl = s->cfl;
  d->precision = s->precision;
  d->linear = s->linear;
  return s->order >= 0 && s->order <= 9 && s->num_variables > 0 && s->num_parameters >= 0;
}

// Device buffer owned for the duration of one synchronous call.
struct DevBuf {
  void* p = nullptr;
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 8); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  double* d() const { return static_cast<double*>(p); }
};

struct Sizes {
  long B, B2, B3, B4, D, V;
  explicit Sizes(const SolverDesc& sd)
      : B(sd.order + 1), B2(B * B), B3(B2 * B), B4(B3 * B), D(sd.nvar + sd.npar), V(sd.nvar) {}
};

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// The formula fields of ICArgs from an exa_euler_ic: the seeded Fourier
// modes draw integer wave vectors in [-3,3]^3 (not all zero) and four phases
// in [0, 2 pi) each from splitmix64.
void fill_euler_ic(const exa_euler_ic* ic, exa::ICArgs* a) {
  for (int d = 0; d < 3; ++d) {
    a->k[d] = ic->k[d];
    a->vel[d] = ic->vel[d];
  }
  a->rho0 = ic->rho0;
  a->amp = ic->amp;
  a->p = ic->p;
  a->nmodes = ic->nmodes;
  a->mode_amp = ic->mode_amp;
  uint64_t state = ic->seed;
  for (int m = 0; m < ic->nmodes; ++m) {
    int kk[3];
    do {
      for (int d = 0; d < 3; ++d) kk[d] = (int)(splitmix64(state) % 7) - 3;
    } while (kk[0] == 0 && kk[1] == 0 && kk[2] == 0);
    for (int d = 0; d < 3; ++d) a->mode_k[m][d] = kk[d];
    for (int c = 0; c < 4; ++c)
      a->mode_phase[m][c] = 6.283185307179586476925286766559 * (double)(splitmix64(state) >> 11) * 0x1.0p-53;
  }
}

}  // namespace
