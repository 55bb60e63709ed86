// LJ cluster kernels, FP32 pair terms accumulated in FP64, and the check of
// the reciprocal those pair terms use.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"
namespace sadmc {
hipError_t launch_lj32(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  return launch_km<SYS_LJ, 1>(meth, kmax, kk, a, la, s);
}

// The fast reciprocal against the division for every binary32 pattern x in
// [lo, hi): what = 0 counts the x whose recip_f32_fast(x) and 1.0f / x differ
// (two NaN results count as equal); what = 1 counts the x whose
// lj_pair_f32_fast(x) is neither NaN nor equal to lj_pair_f32(x), i.e. the
// inputs on which the pair loops' NaN test would miss a difference.
// *bad += the count.  Grid-stride over 64-bit indices, hi <= 2^32.
__global__ __launch_bounds__(256) void k_check_recip(int what, uint64_t lo, uint64_t hi, unsigned long long *bad) {
  unsigned long long local = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += stride) {
    const float x = __builtin_bit_cast(float, (uint32_t)i);
    bool ok;
    if (what == 0) {
      const float q = 1.0f / x, r = recip_f32_fast(x);
      ok = __builtin_bit_cast(uint32_t, q) == __builtin_bit_cast(uint32_t, r) || (q != q && r != r);
    } else {
      const float q = lj_pair_f32(x), r = lj_pair_f32_fast(x);
      ok = r != r || __builtin_bit_cast(uint32_t, q) == __builtin_bit_cast(uint32_t, r);
    }
    local += ok ? 0ULL : 1ULL;
  }
  if (local) atomicAdd(bad, local);
}

hipError_t launch_check_recip(int what, uint64_t lo, uint64_t hi, unsigned long long *bad, hipStream_t s) {
  hipLaunchKernelGGL(k_check_recip, dim3(16384), dim3(256), 0, s, what, lo, hi, bad);
  return hipGetLastError();
}
}  // namespace sadmc
