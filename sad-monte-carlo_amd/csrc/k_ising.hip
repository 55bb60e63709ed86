// Ising kernels (one lane per walker, bit-packed spins).
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"
namespace sadmc {
hipError_t launch_ising(int meth, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  return launch_m<SYS_ISING, 1, 0>(meth, kk, a, la, s);
}
}  // namespace sadmc
