// Kernel templates and per-system launch entry points.
//
// Launch shape: one 64-lane wavefront per workgroup; a workgroup holds 64/G
// walkers of one set, G lanes each (G = 1 for Ising), and the LDS mirror of
// their positions: (64/G) * 3 * npad doubles.  The run kernel is persistent
// over the K moves of a batch: walker state is read from HBM once, kept in
// registers for the whole batch, and written back once.
#pragma once

#include "sadmc_engine.h"

namespace sadmc {

enum KernelKind { KK_RUN = 0, KK_INIT = 1, KK_PLAN = 2, KK_CONFIRM = 3, KK_PROBE = 4, KK_ENERGY = 5 };

struct LaunchArgs {
  uint64_t moves;
  double e_lo;
  int walker;
  double scale;
  Pending *pending;
  uint64_t seed;  // probe: walker w uses the stream seeded by seed + w
  int count;      // probe: records per walker
  float *out;     // probe: [W][count][24]
  double *eout;   // energy: the recomputed energy of `walker`
  int waves;      // run: wavefronts per workgroup (1, 2 or 4)
};

// One workgroup of the batched launch: walkers [first_walker, first_walker +
// 64/G) of set `set`, whose system is `sys`.
struct BatchTile {
  int32_t set, first_walker, sys, pad;
};

// One entry point per system kind; each lives in its own translation unit.
// `waves` wavefronts per workgroup, each running its own tile with
// `wave_lds` bytes of LDS.
hipError_t launch_batched(const SetArgs *d_sets, const BatchTile *d_tiles, int ntiles, size_t wave_lds, int waves,
                          uint64_t moves, hipStream_t s);
hipError_t launch_ising(int meth, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s);
hipError_t launch_lj64(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s);
hipError_t launch_lj32(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s);
// The fast reciprocal against the division over binary32 patterns [lo, hi) (k_lj32.hip).
hipError_t launch_check_recip(int what, uint64_t lo, uint64_t hi, unsigned long long *bad, hipStream_t s);
// Learned-move LJ kernels (run, plan, probe), FP64 or FP32 pair terms.
hipError_t launch_learned64(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la,
                            hipStream_t s);
hipError_t launch_learned32(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la,
                            hipStream_t s);
// Training of the proposal network (k_train.hip).
hipError_t launch_prop_grad(const float *rec, int64_t nrec, int64_t first, int count, const uint16_t *wbf,
                            float *partial, hipStream_t s);
// On-device mean entropy of a walker set (k_entropy.hip); `ew` holds
// entropy_walker_bytes() per walker.
// The bin pass splits the walkers into `chunks` (entropy_chunks) whose
// partial sums (psum, pcnt: chunks * ng each) are combined in chunk order.
hipError_t launch_entropy(const sadmc_walker_state *ws, const HotBin *hot, int W, int64_t cap, double width,
                          double gmin, int64_t ga, int64_t gb, int64_t ng, void *ew, int chunks, double *psum,
                          int64_t *pcnt, double *S, int64_t *count, hipStream_t s);
int entropy_chunks(int W, int64_t ng);
size_t entropy_walker_bytes();
// The common grid of a walker set (k_entropy_range).
struct EntropyRange {
  double gmin, lo, hi;
  int64_t ng;
  int32_t any, pad;
};
// `part` holds kEntropyRangeBlocks records of scratch.
constexpr int kEntropyRangeBlocks = 64;
hipError_t launch_entropy_range(const sadmc_walker_state *ws, int W, double width, EntropyRange *out,
                                EntropyRange *part, hipStream_t s);
hipError_t launch_prop_update(const float *partial, int nblk, int count, float *master, float *m, float *v,
                              uint16_t *wbf, float lr, int step, float *grad_out, float *loss_out, hipStream_t s);
hipError_t launch_wca(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s);

// Parallel-tempering ladders (k_pt.hip).  The systems stay in their slots
// (slot s of ladder l is walker l * R + s of the SetArgs); a swap exchanges
// which replica (temperature, generator, counters) a slot's system belongs
// to, which is the reference's exchange of systems seen from the systems.
struct PtArgs {
  sadmc_replica *rep;  // [L][R] by replica
  double *slot_e;      // [L][R][2] energy and error estimate of each slot's system
  int32_t *slot_of;    // [L][R] the slot holding replica r's system
  uint64_t *lrng;      // [L][2] the ladders' generators
  const double *geom;  // [L]
  int32_t R, steps;
  double d, sc;        // cube half-width and its Uniform scale
};
// Energy-ceiling replica runs (k_zmc.hip): one wave per run.  Slots
// hold systems (slot s of run i is walker i * RMAX + s of the SetArgs); the
// replica -> slot table follows the reference's exchanges of systems.
struct ZSim {
  uint64_t rng[2];  // the run's own generator
  uint64_t moves;
  int32_t R, status, med_len, pad;
};
struct ZArgs {
  ZSim *sim;              // [S]
  sadmc_zreplica *rep;    // [S][RMAX] by replica
  double *slot;           // [S][RMAX][3]: E, error, lowest_max_energy of each slot's system
  int32_t *slot_of;       // [S][RMAX]
  double *med;            // [S][4096] median estimator
  const double *geom;     // [S] sphere radius (LJ)
  int32_t RMAX, steps;
  uint64_t isbnb;
  double min_T, shrink;   // shrink: 0.5^(1/dimensionality)
};
size_t zmc_lds_bytes(int RMAX, int npad);
hipError_t launch_zmc(int sys, int kmax, const ZArgs &z, const SetArgs &a, int sims, int threads, uint64_t rounds,
                      hipStream_t s);

// Walker sets of the binning path (k_binning.hip); the system part uses
// the SetArgs of the set (positions, spins, lanes), the rest is here.
struct BinArgs {
  sadmc_bwalker_state *ws;  // [W]
  sadmc_bbin *bins;         // [W][cap]
  int32_t *flag;            // set to 1 when a walker stops with a status
  int64_t cap;
  double e_lo;              // energy of physical bin 0's lower edge, for placement
  int32_t move_plan, has_min_e, has_max_e, collect_pressure;
  double move_param, min_e, max_e;
  int32_t linear;           // binning/linear.rs instead of binning/histogram.rs
  // the high_resolution side histogram (null when off): per-walker
  // statistics and [W][hr_cap] counts, bins of width hr_de
  sadmc_hires *hrs;
  uint64_t *hr;
  int64_t hr_cap;
  double hr_de;
};
// `init` runs from_params (energy, descent, method state) instead of moves.
hipError_t launch_binning(int sys, int kmax, const BinArgs &ba, const SetArgs &sa, uint64_t moves, int init,
                          hipStream_t s);

// One workgroup of `threads` lanes per ladder; `init` computes each slot's
// energy instead of running rounds.
size_t pt_lds_bytes(int R, int npad);
hipError_t launch_pt(int sys, int kmax, const SetArgs &a, const PtArgs &p, int ladders, int threads, uint64_t rounds,
                     int init, hipStream_t s);

#ifdef SADMC_KERNEL_TU

// GC > 0: the set's lanes per walker as a compile-time constant.  A
// workgroup of one to four wavefronts holds blockDim.x / G walkers.
template <int SYS, int KMAX, int PREC, int METH, int GC = 0>
__global__ __launch_bounds__(256) void k_run(SetArgs a, uint64_t moves) {
  extern __shared__ double lds[];
  const int G = GC ? GC : a.G;
  const int group = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  const int walker = blockIdx.x * (blockDim.x / G) + group;
  if (walker >= a.W) return;
  Engine<SYS, KMAX, PREC, METH, 0, GC> e;
  e.load(a, walker, lane, lds + group * 3 * a.npad);
  e.run(moves);
  e.store(walker);
}

template <int SYS, int KMAX, int PREC, int METH>
__global__ __launch_bounds__(64) void k_init(SetArgs a, double e_lo) {
  extern __shared__ double lds[];
  const int G = a.G;
  const int group = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  const int walker = blockIdx.x * (64 / G) + group;
  if (walker >= a.W) return;
  Engine<SYS, KMAX, PREC, METH> e;
  e.load_system(a, walker, lane, lds + group * 3 * a.npad);
  e.from_params(walker, e_lo);
  e.store(walker, true);
}

template <int SYS, int KMAX, int PREC, int METH>
__global__ __launch_bounds__(64) void k_plan(SetArgs a, int walker, double scale, Pending *out) {
  extern __shared__ double lds[];
  if ((int)threadIdx.x >= a.G) return;
  Engine<SYS, KMAX, PREC, METH> e;
  e.load(a, walker, threadIdx.x, lds);
  double e2 = 0.0;
  const double sc = SYS == SYS_ISING ? 0.0 : uniform_scale(-scale, scale);
  bool some = e.plan_move(scale, sc, e2);
  if (threadIdx.x == 0) {
    Pending p;
    p.some = some ? 1 : 0;
    p.which = e.pw;
    p.tx = e.ptx;
    p.ty = e.pty;
    p.tz = e.ptz;
    p.e = e2;
    p.dabse = e.pdabse;
    *out = p;
  }
  e.store(walker);
}

template <int SYS, int KMAX, int PREC, int METH>
__global__ __launch_bounds__(64) void k_confirm(SetArgs a, int walker, const Pending *pend) {
  extern __shared__ double lds[];
  if ((int)threadIdx.x >= a.G) return;
  Engine<SYS, KMAX, PREC, METH> e;
  e.load(a, walker, threadIdx.x, lds);
  Pending p = *pend;
  if (p.some) {
    e.pw = p.which;
    e.ptx = p.tx;
    e.pty = p.ty;
    e.ptz = p.tz;
    e.pe = p.e;
    e.pdabse = p.dabse;
    e.confirm();
  }
  e.store(walker);
}

// System::compute_energy (ising.rs:59-75, lj.rs:236-244, wca.rs:222-230) of
// one walker's current configuration, FP64 in lane-group order.
template <int SYS, int KMAX, int PREC, int METH>
__global__ __launch_bounds__(64) void k_energy(SetArgs a, int walker, double *out) {
  extern __shared__ double lds[];
  if ((int)threadIdx.x >= a.G) return;
  Engine<SYS, KMAX, PREC, METH> e;
  e.load_system(a, walker, threadIdx.x, lds);
  e.L = a.ws[walker].geom;
  e.halfL = 0.5 * e.L;
  const double en = e.full_energy();
  if (threadIdx.x == 0) *out = en;
}

template <int SYS, int KMAX, int PREC, int METH>
static hipError_t launch_t(KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  const int G = a.G;
  const int per_block = 64 / G;
  const size_t lds = SYS == SYS_ISING ? 0 : (size_t)per_block * 3 * a.npad * sizeof(double);
  const size_t lds1 = SYS == SYS_ISING ? 0 : (size_t)3 * a.npad * sizeof(double);
  const dim3 grid((a.W + per_block - 1) / per_block), block(64);
  // the run kernel: la.waves wavefronts per workgroup
  const int wv = la.waves > 1 ? la.waves : 1;
  const int run_per_block = wv * per_block;
  const size_t run_lds = SYS == SYS_ISING ? 0 : (size_t)run_per_block * 3 * a.npad * sizeof(double);
  const dim3 run_grid((a.W + run_per_block - 1) / run_per_block), run_block(64 * wv);
  switch (kk) {
    case KK_RUN:
      // SAD sets of 2-8 atoms per lane at 4, 8 or 16 lanes run a kernel with
      // the lanes fixed at compile time
      if constexpr (SYS != SYS_ISING && METH == M_SAD && KMAX >= 2 && KMAX <= 8) {
        if (a.G == 4) {
          hipLaunchKernelGGL((k_run<SYS, KMAX, PREC, METH, 4>), run_grid, run_block, run_lds, s, a, la.moves);
          break;
        }
        if (a.G == 8) {
          hipLaunchKernelGGL((k_run<SYS, KMAX, PREC, METH, 8>), run_grid, run_block, run_lds, s, a, la.moves);
          break;
        }
        if (a.G == 16) {
          hipLaunchKernelGGL((k_run<SYS, KMAX, PREC, METH, 16>), run_grid, run_block, run_lds, s, a, la.moves);
          break;
        }
      }
      hipLaunchKernelGGL((k_run<SYS, KMAX, PREC, METH>), run_grid, run_block, run_lds, s, a, la.moves);
      break;
    case KK_INIT:
      hipLaunchKernelGGL((k_init<SYS, KMAX, PREC, METH>), grid, block, lds, s, a, la.e_lo);
      break;
    case KK_PLAN:
      hipLaunchKernelGGL((k_plan<SYS, KMAX, PREC, METH>), dim3(1), block, lds1, s, a, la.walker, la.scale,
                         la.pending);
      break;
    case KK_CONFIRM:
      hipLaunchKernelGGL((k_confirm<SYS, KMAX, PREC, METH>), dim3(1), block, lds1, s, a, la.walker, la.pending);
      break;
    case KK_ENERGY:
      hipLaunchKernelGGL((k_energy<SYS, KMAX, PREC, METH>), dim3(1), block, lds1, s, a, la.walker, la.eout);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int SYS, int KMAX, int PREC>
static hipError_t launch_m(int meth, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  switch (meth) {
    case M_SAD: return launch_t<SYS, KMAX, PREC, M_SAD>(kk, a, la, s);
    case M_SAMC: return launch_t<SYS, KMAX, PREC, M_SAMC>(kk, a, la, s);
    case M_WL: return launch_t<SYS, KMAX, PREC, M_WL>(kk, a, la, s);
    default: return launch_t<SYS, KMAX, PREC, M_CAN>(kk, a, la, s);
  }
}

template <int SYS, int PREC>
static hipError_t launch_km(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la,
                            hipStream_t s) {
  switch (kmax) {
    case 1: return launch_m<SYS, 1, PREC>(meth, kk, a, la, s);
    case 2: return launch_m<SYS, 2, PREC>(meth, kk, a, la, s);
    case 4: return launch_m<SYS, 4, PREC>(meth, kk, a, la, s);
    case 8: return launch_m<SYS, 8, PREC>(meth, kk, a, la, s);
    default: return launch_m<SYS, 16, PREC>(meth, kk, a, la, s);
  }
}

// ---------------------------------------------------------- learned move
// LDS: the positions of the 64/G walkers, then the network weights as BF16
// (in the first half of a region of kPropLds floats), then 32 floats per
// walker for the hidden-layer exchange of the non-matrix-core form.
// Two waves per SIMD: the register budget the 16384-walker sets need.  GC
// as in k_run (the SAD kernels of 2 to 8 atoms per lane at 4, 8 or 16 lanes
// have it fixed at compile time: fewer scalar spills, +10 % for LJ38 at 8).
template <int KMAX, int PREC, int METH, int GC = 0>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_run_learned(SetArgs a,
                                                                                             uint64_t moves) {
  extern __shared__ double lds[];
  const int G = GC ? GC : a.G, per = 64 / G;
  float *lw = (float *)(lds + per * 3 * a.npad);
  load_weights_bf16(a.mlp, (uint16_t *)lw);
  const int group = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  const int walker = blockIdx.x * per + group;
  if (walker >= a.W) return;
  Engine<SYS_LJ, KMAX, PREC, METH, 1, GC> e;
  e.PW = (const uint16_t *)lw;
  e.XC = lw + kPropLds + group * kPropH;
  e.load(a, walker, lane, lds + group * 3 * a.npad);
  e.run(moves);
  e.store(walker);
}

template <int KMAX, int PREC, int METH>
__global__ __launch_bounds__(64) void k_plan_learned(SetArgs a, int walker, Pending *out) {
  extern __shared__ double lds[];
  float *lw = (float *)(lds + 3 * a.npad);
  load_weights_bf16(a.mlp, (uint16_t *)lw);
  if ((int)threadIdx.x >= a.G) return;
  Engine<SYS_LJ, KMAX, PREC, METH, 1> e;
  e.PW = (const uint16_t *)lw;
  e.XC = lw + kPropLds;
  e.load(a, walker, threadIdx.x, lds);
  double e2 = 0.0;
  bool some = e.plan_move(0.0, 0.0, e2);
  if (threadIdx.x == 0) {
    Pending p;
    p.some = some ? 1 : 0;
    p.which = e.pw;
    p.tx = e.ptx;
    p.ty = e.pty;
    p.tz = e.ptz;
    p.e = e2;
    p.dabse = some ? e.lnh : 0.0;
    *out = p;
  }
  e.store(walker);
}

// Training records of every walker (reads the walkers, changes nothing).
template <int KMAX, int PREC>
__global__ __launch_bounds__(64) void k_probe(SetArgs a, uint64_t seed, int count, float *out) {
  extern __shared__ double lds[];
  const int G = a.G;
  const int group = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  const int walker = blockIdx.x * (64 / G) + group;
  if (walker >= a.W) return;
  Engine<SYS_LJ, KMAX, PREC, M_SAD, 1> e;
  e.load(a, walker, lane, lds + group * 3 * a.npad);
  e.probe(seed + (uint64_t)walker, count, out + (int64_t)walker * count * kPropRec);
}

template <int KMAX, int PREC, int METH>
static hipError_t launch_learned_t(KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  const int G = a.G;
  const int per_block = 64 / G;
  const size_t wbytes = sizeof(float) * (kPropLds + (size_t)per_block * kPropH);
  const size_t lds = (size_t)per_block * 3 * a.npad * sizeof(double) + wbytes;
  const size_t lds1 = (size_t)3 * a.npad * sizeof(double) + sizeof(float) * (kPropLds + kPropH);
  const dim3 grid((a.W + per_block - 1) / per_block), block(64);
  switch (kk) {
    case KK_RUN:
      if constexpr (METH == M_SAD && KMAX >= 2 && KMAX <= 8) {
        if (G == 4 || G == 8 || G == 16) {
          if (G == 4) hipLaunchKernelGGL((k_run_learned<KMAX, PREC, METH, 4>), grid, block, lds, s, a, la.moves);
          if (G == 8) hipLaunchKernelGGL((k_run_learned<KMAX, PREC, METH, 8>), grid, block, lds, s, a, la.moves);
          if (G == 16) hipLaunchKernelGGL((k_run_learned<KMAX, PREC, METH, 16>), grid, block, lds, s, a, la.moves);
          break;
        }
      }
      hipLaunchKernelGGL((k_run_learned<KMAX, PREC, METH>), grid, block, lds, s, a, la.moves);
      break;
    case KK_PLAN:
      hipLaunchKernelGGL((k_plan_learned<KMAX, PREC, METH>), dim3(1), block, lds1, s, a, la.walker, la.pending);
      break;
    case KK_PROBE:
      hipLaunchKernelGGL((k_probe<KMAX, PREC>), grid, block, (size_t)per_block * 3 * a.npad * sizeof(double), s, a,
                         la.seed, la.count, la.out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int KMAX, int PREC>
static hipError_t launch_learned_m(int meth, KernelKind kk, const SetArgs &a, const LaunchArgs &la, hipStream_t s) {
  if (meth == M_SAD) return launch_learned_t<KMAX, PREC, M_SAD>(kk, a, la, s);
  if (meth == M_CAN) return launch_learned_t<KMAX, PREC, M_CAN>(kk, a, la, s);
  return hipErrorInvalidValue;
}

template <int PREC>
static hipError_t launch_learned_km(int meth, int kmax, KernelKind kk, const SetArgs &a, const LaunchArgs &la,
                                    hipStream_t s) {
  switch (kmax) {
    case 1: return launch_learned_m<1, PREC>(meth, kk, a, la, s);
    case 2: return launch_learned_m<2, PREC>(meth, kk, a, la, s);
    case 4: return launch_learned_m<4, PREC>(meth, kk, a, la, s);
    case 8: return launch_learned_m<8, PREC>(meth, kk, a, la, s);
    default: return launch_learned_m<16, PREC>(meth, kk, a, la, s);
  }
}

#endif  // SADMC_KERNEL_TU

}  // namespace sadmc
