// LJ cluster kernels, FP32 pair terms accumulated in FP64.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"
namespace sadmc {
hipError_t launch_lj32(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  return launch_km<SYS_LJ, 1>(meth, kmax, kk, a, la, s);
}
}  // namespace sadmc
