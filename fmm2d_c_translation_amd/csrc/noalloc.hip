// noalloc.hip -- fmm2dhip_h2d_directcdg_noalloc: h2d_directcdg_ for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The direct sum's device memory -- staging copies of the caller's host
// arrays and the per-chunk partial sums (direct.hip, launch_chunked) -- comes
// from `scratch`, a file-scope device array of kScratchBytes held by every
// GPU the library runs on until the process ends.  Host code never uses the
// array's name as an address: each device's copy is found once with
// hipGetSymbolAddress (with that device current) and the pointer kept per
// device.  Every call on a device shares that device's array, whose previous
// contents are garbage (the kernels clear what they accumulate into), so
// calls that overlap on one GPU race.
#include <mutex>

#include "engine.h"

namespace fmm {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDev = 16;
std::mutex g_mu;
void* g_ptr[kMaxDev] = {};

// the current device's copy of scratch (looked up on the first call there)
int scratch_ptr(void** out) {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDev) return IER_NOGPU;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[d]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return IER_HIP;
    g_ptr[d] = p;
  }
  *out = g_ptr[d];
  return IER_OK;
}

}  // namespace
}  // namespace fmm

extern "C" int fmm2dhip_h2d_directcdg_noalloc(const fmm2d_int* nd, const fmm2d_cplx* zk,
                                              const double* sources, const fmm2d_int* ns,
                                              const fmm2d_cplx* charge, const fmm2d_cplx* dipstr,
                                              const double* dipvec, const double* targ,
                                              const fmm2d_int* nt, fmm2d_cplx* pot,
                                              fmm2d_cplx* grad, const double* thresh,
                                              void* stream) {
  try {
    void* base = nullptr;
    int e = fmm::scratch_ptr(&base);
    if (e) return e;
    fmm::FixedArena arena(base, fmm::kScratchBytes, static_cast<hipStream_t>(stream));
    e = fmm::h2d_direct_cdg(*nd, zk, sources, *ns, charge, dipstr, dipvec, targ, *nt, pot, grad,
                            *thresh);
    if (arena.overflowed()) return fmm::IER_SCRATCH;
    return e;
  } catch (...) {
    return fmm::IER_HIP;
  }
}
