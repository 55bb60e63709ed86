// LJ cluster kernels, FP64 pair terms.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"
namespace sadmc {
hipError_t launch_lj64(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  return launch_km<SYS_LJ, 0>(meth, kmax, kk, a, la, s);
}
}  // namespace sadmc
