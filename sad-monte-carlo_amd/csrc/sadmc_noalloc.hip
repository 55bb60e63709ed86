// sadmc_mean_entropy_noalloc: the walker-set entropy reduction of
// sadmc_mean_entropy (plotting/parse-binning.py:150-171) for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Its scratch -- the grid summary,
// the per-walker constants, S and counts, the chunks' partial sums -- lives
// in a file-scope device array of kScratchBytes instead of the context's
// growable buffer; the call makes no hipMalloc or hipFree.
//
// `scratch` is a device symbol: in host code its name is only a stand-in,
// never an address.  Each device's address of it is looked up once, with
// that device current, by hipGetSymbolAddress on the entry point's first
// call there, and kept per device; kernels and copies get that address.
// Every GPU that runs the library holds its own copy until the process
// ends.  All calls on one GPU share it and find the previous call's
// contents (the kernels write every entry they read, so nothing is
// cleared); calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "sadmc_internal.h"

namespace {
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// This device's address of `scratch`; `device` must be current.
static hipError_t scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

extern "C" int sadmc_mean_entropy_noalloc(sadmc_ctx *c, int32_t set, double e_a, double e_b, int64_t max_bins,
                                          double *e_first, int64_t *nbins, double *S, int64_t *count) {
  if (!c) return SADMC_E_ARG;
  const int device = ctx_device(c);
  if (device < 0 || device >= kMaxDevices) return ctx_fail(c, SADMC_E_ARG, "device index beyond the scratch table");
  if (hipSetDevice(device) != hipSuccess) return ctx_fail(c, SADMC_E_HIP, "hipSetDevice failed");
  void *base = nullptr;
  const hipError_t e = scratch_address(device, &base);
  if (e != hipSuccess || !base) return ctx_fail(c, SADMC_E_HIP, "hipGetSymbolAddress(scratch) failed");
  return mean_entropy_impl(c, set, e_a, e_b, max_bins, e_first, nbins, S, count, (char *)base, kScratchBytes);
}
