// LJ kernels with the learned Gaussian move, FP64 pair terms.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"
namespace sadmc {
hipError_t launch_learned64(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la,
                            hipStream_t s) {
  return launch_learned_km<0>(meth, kmax, kk, a, la, s);
}
}  // namespace sadmc
