// Parallel tempering (src/mc/tempering.rs): one workgroup per ladder, one
// lane group per replica slot.
//
// A round (MC::run_once, tempering.rs:272-341):
// - every slot's group makes `steps` canonical moves of its system at the
//   temperature of the replica that owns it (Replica::run_once,
//   tempering.rs:96-112), with that replica's generator and counters, all
//   groups at once;
// - lane 0 of the workgroup then draws the ladder generator's bool and walks
//   the odd or even neighbour pairs (tempering.rs:288-318) in order, deciding
//   each swap with the second replica's generator and collecting both
//   replicas' energies.  A swap exchanges the owners of two slots, so no
//   positions move; the replica records, the owner tables and the slot
//   energies live in LDS for the whole launch.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"

namespace sadmc {

namespace {

// Replica state in LDS during a launch.
struct PtMeta {
  uint64_t rng0, rng1;
  double T, te, tes;
  uint64_t acc, rej, ign, as, rs;
};

// Replica::collect (tempering.rs:86-94).
__device__ __forceinline__ void pt_collect(PtMeta &m, double e) {
  m.te += e;
  m.tes += e * e;
  if (e >= 0.0) m.ign += 1;
}

}  // namespace

size_t pt_lds_bytes(int R, int npad) {
  return sizeof(double) * 3 * (size_t)R * npad + sizeof(PtMeta) * R + sizeof(double) * 2 * R + sizeof(int32_t) * 2 * R;
}

// GC > 0: the lanes per replica fixed at compile time (2 to 8 atoms per lane
// at 4, 8 or 16 lanes), as for the walker kernels.
template <int SYS, int KMAX, int GC = 0>
__global__ __launch_bounds__(512) void k_pt(SetArgs a, PtArgs p, uint64_t rounds, int init) {
  extern __shared__ double lds[];
  const int R = p.R, G = GC ? GC : a.G, ladder = blockIdx.x;
  const int tid = threadIdx.x;
  const int slot = tid / G, lane = tid & (G - 1);
  const bool active = slot < R;
  const int npad = SYS == SYS_ISING ? 0 : a.npad;
  double *pos = lds;
  PtMeta *meta = (PtMeta *)(lds + 3 * R * npad);
  double *se = (double *)(meta + R);
  int32_t *slot_of = (int32_t *)(se + 2 * R);
  int32_t *rep_of = slot_of + R;
  const int64_t L0 = (int64_t)ladder * R;

  for (int r = tid; r < R; r += blockDim.x) {
    const sadmc_replica &q = p.rep[L0 + r];
    PtMeta m;
    m.rng0 = q.rng[0];
    m.rng1 = q.rng[1];
    m.T = q.T;
    m.te = q.total_energy;
    m.tes = q.total_energy_squared;
    m.acc = q.accepted;
    m.rej = q.rejected;
    m.ign = q.ignored;
    m.as = q.accepted_swaps;
    m.rs = q.rejected_swaps;
    meta[r] = m;
    slot_of[r] = p.slot_of[L0 + r];
    se[2 * r] = p.slot_e[2 * (L0 + r)];
    se[2 * r + 1] = p.slot_e[2 * (L0 + r) + 1];
  }
  __syncthreads();
  for (int r = tid; r < R; r += blockDim.x) rep_of[slot_of[r]] = r;

  Engine<SYS, KMAX, 0, M_CAN, 0, GC> e;
  const int64_t w = L0 + slot;
  if (active) {
    e.load_system(a, (int)w, lane, pos + (int64_t)slot * 3 * npad);
    e.E = se[2 * slot];
    e.err = se[2 * slot + 1];
    if (SYS != SYS_ISING) {
      const double g = p.geom[ladder];
      e.R2 = g * g;
      e.L = g;
      e.halfL = 0.5 * g;
    }
  }
  __syncthreads();

  if (init) {
    // the energy of each slot's initial configuration, in lane-group order
    if (active) {
      const double E = e.full_energy();
      if (lane == 0) {
        p.slot_e[2 * w] = E;
        p.slot_e[2 * w + 1] = 0.0;
      }
    }
    return;
  }

  uint64_t l0 = 0, l1 = 0;
  if (tid == 0) {
    l0 = p.lrng[2 * ladder];
    l1 = p.lrng[2 * ladder + 1];
  }
  for (uint64_t k = 0; k < rounds; k++) {
    if (active) {
      const int r = rep_of[slot];
      PtMeta m = meta[r];
      e.s0 = m.rng0;
      e.s1 = m.rng1;
      for (int i = 0; i < p.steps; i++) {
        double e2;
        if (!e.cube_plan(p.d, p.sc, e2)) continue;
        const double bde = (e2 - e.E) / m.T;
        if (bde < 0.0 || gen_f64(e.s0, e.s1) < exp(-bde)) {
          e.confirm();
          m.acc += 1;
        } else {
          m.rej += 1;
        }
        pt_collect(m, e.E);
      }
      if (lane == 0) {
        m.rng0 = e.s0;
        m.rng1 = e.s1;
        meta[r] = m;
        se[2 * slot] = e.E;
        se[2 * slot + 1] = e.err;
      }
    }
    __syncthreads();
    if (tid == 0) {
      // rand 0.7's bool is the top bit of next_u32, which rand_xoshiro takes
      // from the high half of next_u64
      const int first = (rng_next(l0, l1) >> 63) ? 0 : 1;
      for (int i = first; i + 1 < R; i += 2) {
        PtMeta &m0 = meta[i], &m1 = meta[i + 1];
        const int a0 = slot_of[i], a1 = slot_of[i + 1];
        const double E0 = se[2 * a0], E1 = se[2 * a1];
        const double de_db = (E0 - E1) * (1.0 / m0.T - 1.0 / m1.T);
        bool swap = de_db >= 0.0;
        if (!swap) {
          uint64_t u0 = m1.rng0, u1 = m1.rng1;
          swap = gen_f64(u0, u1) < exp(de_db);
          m1.rng0 = u0;
          m1.rng1 = u1;
        }
        if (swap) {
          m0.as += 1;
          m1.as += 1;
          slot_of[i] = a1;
          slot_of[i + 1] = a0;
          rep_of[a1] = i;
          rep_of[a0] = i + 1;
          pt_collect(m0, E1);
          pt_collect(m1, E0);
        } else {
          m0.rs += 1;
          m1.rs += 1;
          pt_collect(m0, E0);
          pt_collect(m1, E1);
        }
      }
    }
    __syncthreads();
  }

  // write back: the slots' systems, the replica records, the owner table
  if (active) {
    if (SYS == SYS_ISING) {
      if (lane == 0) a.spins[w] = e.spins;
    } else {
      double *P = a.pos + w * 3 * npad;
      const double *S = pos + (int64_t)slot * 3 * npad;
      for (int j = lane; j < npad; j += G) {
        P[j] = S[j];
        P[npad + j] = S[npad + j];
        P[2 * npad + j] = S[2 * npad + j];
      }
    }
  }
  for (int r = tid; r < R; r += blockDim.x) {
    const PtMeta m = meta[r];
    sadmc_replica &q = p.rep[L0 + r];
    q.rng[0] = m.rng0;
    q.rng[1] = m.rng1;
    q.total_energy = m.te;
    q.total_energy_squared = m.tes;
    q.accepted = m.acc;
    q.rejected = m.rej;
    q.ignored = m.ign;
    q.accepted_swaps = m.as;
    q.rejected_swaps = m.rs;
    const int s = slot_of[r];
    q.E = se[2 * s];
    q.error = se[2 * s + 1];
    p.slot_of[L0 + r] = s;
    p.slot_e[2 * (L0 + r)] = se[2 * r];
    p.slot_e[2 * (L0 + r) + 1] = se[2 * r + 1];
  }
  if (tid == 0) {
    p.lrng[2 * ladder] = l0;
    p.lrng[2 * ladder + 1] = l1;
  }
}

template <int SYS, int KMAX>
static hipError_t launch_pt_t(const SetArgs &a, const PtArgs &p, int ladders, int threads, uint64_t rounds, int init,
                              hipStream_t s) {
  const size_t lds = pt_lds_bytes(p.R, SYS == SYS_ISING ? 0 : a.npad);
  if constexpr (SYS != SYS_ISING && KMAX >= 2 && KMAX <= 8) {
    if (a.G == 4 || a.G == 8 || a.G == 16) {
      const dim3 grid(ladders), block(threads);
      if (a.G == 4) hipLaunchKernelGGL((k_pt<SYS, KMAX, 4>), grid, block, lds, s, a, p, rounds, init);
      if (a.G == 8) hipLaunchKernelGGL((k_pt<SYS, KMAX, 8>), grid, block, lds, s, a, p, rounds, init);
      if (a.G == 16) hipLaunchKernelGGL((k_pt<SYS, KMAX, 16>), grid, block, lds, s, a, p, rounds, init);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_pt<SYS, KMAX>), dim3(ladders), dim3(threads), lds, s, a, p, rounds, init);
  return hipGetLastError();
}

template <int SYS>
static hipError_t launch_pt_s(int kmax, const SetArgs &a, const PtArgs &p, int ladders, int threads, uint64_t rounds,
                              int init, hipStream_t s) {
  switch (kmax) {
    case 1: return launch_pt_t<SYS, 1>(a, p, ladders, threads, rounds, init, s);
    case 2: return launch_pt_t<SYS, 2>(a, p, ladders, threads, rounds, init, s);
    case 4: return launch_pt_t<SYS, 4>(a, p, ladders, threads, rounds, init, s);
    case 8: return launch_pt_t<SYS, 8>(a, p, ladders, threads, rounds, init, s);
    case 16: return launch_pt_t<SYS, 16>(a, p, ladders, threads, rounds, init, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pt(int sys, int kmax, const SetArgs &a, const PtArgs &p, int ladders, int threads, uint64_t rounds,
                     int init, hipStream_t s) {
  if (sys == SYS_ISING) return launch_pt_t<SYS_ISING, 1>(a, p, ladders, threads, rounds, init, s);
  if (sys == SYS_LJ) return launch_pt_s<SYS_LJ>(kmax, a, p, ladders, threads, rounds, init, s);
  if (sys == SYS_WCA) return launch_pt_s<SYS_WCA>(kmax, a, p, ladders, threads, rounds, init, s);
  return hipErrorInvalidValue;
}

}  // namespace sadmc
