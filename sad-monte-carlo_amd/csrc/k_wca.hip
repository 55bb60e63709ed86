// WCA fluid kernels (periodic cube, all-pairs minimum image).
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"
namespace sadmc {
hipError_t launch_wca(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  return launch_km<SYS_WCA, 0>(meth, kmax, kk, a, la, s);
}
}  // namespace sadmc
