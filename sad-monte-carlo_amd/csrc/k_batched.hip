// One persistent launch over walker sets of different systems and sizes.
//
// The host lays the sets out as tiles: each workgroup (one wavefront) runs
// 64/G walkers of one set, and a block -> tile table says which.  Tiles are
// ordered heaviest first (pair work per walker x walkers per tile), so the
// expensive WCA-256 waves start first and the light LJ-13 waves fill the
// tail; the hardware hands consecutive workgroups to the 8 XCDs round-robin,
// which spreads every size class over all XCDs.  The kernel switches once per
// workgroup on the set's variant and then runs the same engine as the
// per-set kernels, so results are bit-identical to them.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"

namespace sadmc {

template <int SYS, int KMAX>
__device__ __forceinline__ void batched_walker(const SetArgs &a, int walker, int lane, double *lds, uint64_t moves) {
  Engine<SYS, KMAX, 0, M_SAD> e;
  e.load(a, walker, lane, lds);
  e.run(moves);
  e.store(walker);
}

// Register budget of two waves per SIMD: the kernel holds every variant, and
// without the bound the widest one would cap the whole batch at one wave.
// A workgroup of blockDim.x / 64 wavefronts takes consecutive tiles, each
// wave with `wave_lds` doubles of LDS.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_batched(
    const SetArgs *sets, const BatchTile *tiles, int ntiles, int wave_lds, uint64_t moves) {
  extern __shared__ double lds[];
  const int wave = threadIdx.x / 64;
  const int ti = blockIdx.x * (blockDim.x / 64) + wave;
  if (ti >= ntiles) return;
  const BatchTile t = tiles[ti];
  const SetArgs &a = sets[t.set];
  const int G = a.G;
  const int wl = threadIdx.x & 63;
  const int group = wl / G;
  const int lane = wl & (G - 1);
  const int walker = t.first_walker + group;
  if (walker >= a.W) return;
  double *mine = lds + (size_t)wave * wave_lds + group * 3 * a.npad;
  const int kmax = a.K <= 1 ? 1 : a.K <= 2 ? 2 : a.K <= 4 ? 4 : a.K <= 8 ? 8 : 16;
  if (t.sys == SYS_LJ) {
    switch (kmax) {
      case 1: batched_walker<SYS_LJ, 1>(a, walker, lane, mine, moves); break;
      case 2: batched_walker<SYS_LJ, 2>(a, walker, lane, mine, moves); break;
      case 4: batched_walker<SYS_LJ, 4>(a, walker, lane, mine, moves); break;
      case 8: batched_walker<SYS_LJ, 8>(a, walker, lane, mine, moves); break;
      default: batched_walker<SYS_LJ, 16>(a, walker, lane, mine, moves); break;
    }
  } else {
    switch (kmax) {
      case 1: batched_walker<SYS_WCA, 1>(a, walker, lane, mine, moves); break;
      case 2: batched_walker<SYS_WCA, 2>(a, walker, lane, mine, moves); break;
      case 4: batched_walker<SYS_WCA, 4>(a, walker, lane, mine, moves); break;
      case 8: batched_walker<SYS_WCA, 8>(a, walker, lane, mine, moves); break;
      default: batched_walker<SYS_WCA, 16>(a, walker, lane, mine, moves); break;
    }
  }
}

hipError_t launch_batched(const SetArgs *d_sets, const BatchTile *d_tiles, int ntiles, size_t wave_lds, int waves,
                          uint64_t moves, hipStream_t s) {
  const int wv = waves > 1 ? waves : 1;
  hipLaunchKernelGGL(k_batched, dim3((ntiles + wv - 1) / wv), dim3(64 * wv), wave_lds * wv, s, d_sets, d_tiles,
                     ntiles, (int)(wave_lds / sizeof(double)), moves);
  return hipGetLastError();
}

}  // namespace sadmc
