// Energy-ceiling replicas (src/mc/energy_replicas.rs): one wave per run;
// its lane groups of G lanes take the system slots in turn (group g runs
// slots g, g + 64/G, ... that hold replicas), so slots without a replica
// cost nothing.
//
// A round (MC::run_once, energy_replicas.rs:504-641, without plugins):
// - every replica's group makes its moves: N (LJ, WCA) or N^2
//   (Ising) moves accepted iff E' < max_energy for a bounded replica
//   (Replica::run_once, energy_replicas.rs:200-245), one randomization for
//   the unbounded one, with the replica's generator and counters; a WCA
//   replica above its cutoff also collects the pressure on iterations that
//   are multiples of N^2 (System::data_to_collect, wca.rs:202-218);
// - lane 0 then walks the neighbour exchanges, feeds the median estimator
//   and decides whether the lowest replica splits; the split's median is a
//   selection over the estimator's energies by the whole workgroup; lane 0
//   makes the new replica, the workgroup copies its system, and lane 0 runs
//   the occasional translation-scale updates.
// Systems stay in their slots: an exchange swaps the owners of two slots,
// and lowest_max_energy, which the reference swaps with the systems, lives
// with the slot.  Replica records, slot data and tables are in LDS for the
// whole launch.
#define SADMC_KERNEL_TU
#include "sadmc_kernels.h"

namespace sadmc {

namespace {

struct ZMeta {
  uint64_t rng0, rng1;
  double max_e, cutoff, ts;
  uint64_t rej, acc, above_c, below_c, upw, uniq;
  double above_t, below_t, above_sq, below_sq;
  double p_t;    // above_extra["pressure"]: total and count
  uint64_t p_c;
  int32_t collecting, pad;
};

// the data collection of Replica::run_once (energy_replicas.rs:222-244)
__device__ __forceinline__ void z_collect(ZMeta &m, double E, double low, double very_lowest) {
  if (!m.collecting) return;
  if (E > m.cutoff) {
    m.above_c += 1;
    m.above_t += E;
    m.above_sq += E * E;
  } else {
    m.below_c += 1;
    m.below_t += E;
    m.below_sq += E * E;
  }
  if (low == very_lowest) m.upw += 1;
}

// The WCA energy of atoms [0, m) in lane-group order: lane l sums the pairs
// (i, j < i) of its atoms i = l, l + G, ... < m.
template <int SYS, int KMAX, int GC>
__device__ double wca_energy_first(const Engine<SYS, KMAX, 0, M_CAN, 0, GC> &e, int m, double rc2) {
  double acc = 0.0;
  for (int i = e.lane; i < m; i += e.G)
    for (int j = 0; j < i; j++)
      acc += wca_pair(norm2(min_image(e.lx[i] - e.lx[j], e.L, e.halfL), min_image(e.ly[i] - e.ly[j], e.L, e.halfL),
                            min_image(e.lz[i] - e.lz[j], e.L, e.halfL)),
                      rc2);
  return group_sum(acc, e.G);
}

}  // namespace

size_t zmc_lds_bytes(int RMAX, int npad) {
  return sizeof(double) * 3 * (size_t)RMAX * npad + sizeof(uint64_t) * RMAX + sizeof(ZMeta) * RMAX +
         sizeof(double) * 3 * RMAX + sizeof(int32_t) * 2 * RMAX + sizeof(double) * 40 + sizeof(int32_t) * 8;
}

// GC > 0: the lanes per replica fixed at compile time (2 to 8 atoms per lane
// at 4, 8 or 16 lanes), as for the walker kernels.
template <int SYS, int KMAX, int GC = 0>
__global__ __launch_bounds__(64) void k_zmc(ZArgs z, SetArgs a, uint64_t rounds) {
  extern __shared__ double lds[];
  const int RM = z.RMAX, G = GC ? GC : a.G, sim = blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  // one wave per run: its groups take slots group, group + per, ... in
  // turn, so only slots holding replicas cost anything
  const int per = nthr / G, group = tid / G, lane = tid & (G - 1);
  const int npad = SYS == SYS_ISING ? 0 : a.npad;
  double *pos = lds;
  uint64_t *spn = (uint64_t *)(pos + 3 * RM * npad);
  ZMeta *meta = (ZMeta *)(spn + RM);
  double *sE = (double *)(meta + RM);
  double *sErr = sE + RM;
  double *sLow = sErr + RM;
  int32_t *slot_of = (int32_t *)(sLow + RM);
  int32_t *rep_of = slot_of + RM;
  double *red = (double *)(rep_of + RM);  // 40 doubles: selection results and per-wave partials
  int32_t *ctl = (int32_t *)(red + 40);   // R, split, source slot, status
  uint64_t *smoves = (uint64_t *)(ctl + 4);  // the run's move count at the round's start
  const int64_t S0 = (int64_t)sim * RM;
  double *M = z.med + (int64_t)sim * 4096;

  for (int r = tid; r < RM; r += nthr) {
    const sadmc_zreplica &q = z.rep[S0 + r];
    ZMeta m;
    m.rng0 = q.rng[0];
    m.rng1 = q.rng[1];
    m.max_e = q.max_energy;
    m.cutoff = q.cutoff_energy;
    m.ts = q.translation_scale;
    m.rej = q.rejected_count;
    m.acc = q.accepted_count;
    m.above_c = q.above_count;
    m.below_c = q.below_count;
    m.upw = q.upwelling_count;
    m.uniq = q.unique_visitors;
    m.above_t = q.above_total;
    m.below_t = q.below_total;
    m.above_sq = q.above_total_squared;
    m.below_sq = q.below_total_squared;
    m.p_t = q.pressure_total;
    m.p_c = q.pressure_count;
    m.collecting = q.collecting_data;
    m.pad = 0;
    meta[r] = m;
    slot_of[r] = z.slot_of[S0 + r];
    sE[r] = z.slot[3 * (S0 + r)];
    sErr[r] = z.slot[3 * (S0 + r) + 1];
    sLow[r] = z.slot[3 * (S0 + r) + 2];
    if (SYS == SYS_ISING) spn[r] = a.spins[S0 + r];
  }
  if (tid == 0) {
    ctl[0] = z.sim[sim].R;
    ctl[1] = 0;
    ctl[3] = z.sim[sim].status;
    *smoves = z.sim[sim].moves;
  }
  __syncthreads();
  for (int r = tid; r < RM; r += nthr) rep_of[slot_of[r]] = r;

  Engine<SYS, KMAX, 0, M_CAN, 0, GC> e;
  for (int slot = group; slot < RM; slot += per)
    e.load_system(a, (int)(S0 + slot), lane, pos + (int64_t)slot * 3 * npad);
  {
    if (SYS == SYS_LJ) {
      const double g = z.geom[sim];
      e.R2 = g * g;
    } else if (SYS == SYS_WCA) {
      e.L = z.geom[sim];
      e.halfL = 0.5 * e.L;
    }
    e.status = 0;
  }
  uint64_t l0 = 0, l1 = 0, moves = 0;
  int32_t med_len = 0;
  if (tid == 0) {
    l0 = z.sim[sim].rng[0];
    l1 = z.sim[sim].rng[1];
    moves = z.sim[sim].moves;
    med_len = z.sim[sim].med_len;
  }
  __syncthreads();

  const double sc1 = uniform_scale(-1.0, 1.0);
  const uint64_t nn2 = (uint64_t)a.n * (uint64_t)a.n;
  for (uint64_t round = 0; round < rounds; round++) {
    const int R = ctl[0];
    if (ctl[3]) break;
    const double very_lowest = meta[R - 1].max_e;
    const uint64_t mv0 = *smoves;
    // ---- every replica's moves
    for (int slot = group; slot < R; slot += per) {
      if (SYS != SYS_ISING) {
        e.lx = pos + (int64_t)slot * 3 * npad;
        e.ly = e.lx + npad;
        e.lz = e.ly + npad;
      }
      e.E = sE[slot];
      e.err = sErr[slot];
      if (SYS == SYS_ISING) {
        e.spins = spn[slot];
      } else if (!e.kLdsPos) {
#pragma unroll
        for (int k = 0; k < KMAX; k++)
          if (k < e.K) {
            const int j = lane + G * k;
            e.x[k] = e.lx[j];
            e.y[k] = e.ly[j];
            e.z[k] = e.lz[j];
          }
      }
      const int r = rep_of[slot];
      ZMeta m = meta[r];
      double low = sLow[slot];
      e.s0 = m.rng0;
      e.s1 = m.rng1;
      if (__builtin_isfinite(m.max_e)) {
        const double usc = SYS == SYS_ISING ? 0.0 : uniform_scale(-m.ts, m.ts);
        for (int i = 0; i < z.steps; i++) {
          double e2;
          if (e.cube_plan(m.ts, usc, e2)) {
            if (e2 < m.max_e) {
              e.confirm();
              m.acc += 1;
            } else {
              m.rej += 1;
            }
          } else {
            m.rej += 1;
          }
          z_collect(m, e.E, low, very_lowest);
          if (SYS == SYS_WCA && m.collecting && e.E > m.cutoff && (mv0 + (uint64_t)i) % nn2 == 0) {
            m.p_t += e.pressure();
            m.p_c += 1;
          }
        }
      } else {
        // System::randomize: ising.rs:79-85, lj.rs:262-279
        if (SYS == SYS_ISING) {
          uint64_t bits = 0;
          const int nn = e.n * e.n;
          for (int k = 0; k < nn; k++)
            if (rng_next(e.s0, e.s1) & 1ULL) bits |= 1ULL << k;
          e.spins = bits;
        } else if (SYS == SYS_WCA) {
          // wca.rs:251-267: the atoms are added back one at a time at
          // put_in_cell(Uniform(0, L)^3); each addition goes through
          // set_energy (wca.rs:164-180) with the atom count after it, so
          // the error estimate runs on from the old system's; dabse sums
          // the new atom's pairs with the atoms already placed, lane-group
          // order
          const double Lb = e.L, scL = uniform_scale(0.0, Lb);
          for (int i = 0; i < e.n; i++) {
            const double x = put_in_cell(uniform_sample(e.s0, e.s1, 0.0, scL), Lb);
            const double y = put_in_cell(uniform_sample(e.s0, e.s1, 0.0, scL), Lb);
            const double zz = put_in_cell(uniform_sample(e.s0, e.s1, 0.0, scL), Lb);
            double acc = 0.0;
            for (int j = lane; j < i; j += G)
              acc += wca_pair(norm2(min_image(e.lx[j] - x, Lb, e.halfL), min_image(e.ly[j] - y, Lb, e.halfL),
                                    min_image(e.lz[j] - zz, Lb, e.halfL)),
                              a.rc2);
            const double dabse = group_sum(acc, G);
            if ((i & (G - 1)) == lane) {
              e.lx[i] = x;
              e.ly[i] = y;
              e.lz[i] = zz;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const double new_e = e.E + dabse, cnt = (double)(i + 1);
            const double single = dabse > fabs(new_e) ? 1e-14 * dabse * cnt : 1e-14 * fabs(new_e) * cnt;
            e.err += single * cnt;
            if (e.err > fabs(new_e) * 1e-13 * cnt * cnt) {
              e.E = wca_energy_first(e, i + 1, a.rc2);
              e.err = 1e-15 * e.E * cnt;
            } else {
              e.E = new_e;
            }
          }
        } else {
          const double Rr = z.geom[sim];
          for (int i = 0; i < e.n; i++) {
            double x, y, zz;
            for (;;) {
              x = uniform_sample(e.s0, e.s1, -1.0, sc1);
              y = uniform_sample(e.s0, e.s1, -1.0, sc1);
              zz = uniform_sample(e.s0, e.s1, -1.0, sc1);
              if (norm2(x, y, zz) < 1.0) break;
            }
            if ((i & (G - 1)) == lane) {
              e.lx[i] = x * Rr;
              e.ly[i] = y * Rr;
              e.lz[i] = zz * Rr;
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          if (!e.kLdsPos) {
#pragma unroll
            for (int k = 0; k < KMAX; k++)
              if (k < e.K) {
                const int j = lane + G * k;
                e.x[k] = e.lx[j];
                e.y[k] = e.ly[j];
                e.z[k] = e.lz[j];
              }
          }
        }
        e.E = e.full_energy();
        low = m.max_e;
        z_collect(m, e.E, low, very_lowest);
        if (SYS == SYS_WCA && m.collecting && e.E > m.cutoff && mv0 % nn2 == 0) {
          m.p_t += e.pressure();
          m.p_c += 1;
        }
      }
      if (lane == 0) {
        m.rng0 = e.s0;
        m.rng1 = e.s1;
        meta[r] = m;
        sE[slot] = e.E;
        sErr[slot] = e.err;
        sLow[slot] = low;
        if (SYS == SYS_ISING) spn[slot] = e.spins;
      }
    }
    __syncthreads();
    // ---- exchanges, the median estimator, the split decision (lane 0)
    if (tid == 0) {
      // rand 0.7's bool: the top bit of next_u32, the high half of next_u64
      const int first = (rng_next(l0, l1) >> 63) ? 0 : 1;
      for (int i = first; i + 1 < R; i += 2) {
        const int s0 = slot_of[i], s1 = slot_of[i + 1];
        if (sE[s0] < meta[i + 1].max_e) {
          slot_of[i] = s1;
          slot_of[i + 1] = s0;
          rep_of[s1] = i;
          rep_of[s0] = i + 1;
          meta[i].collecting = 1;
          meta[i + 1].collecting = 1;
          if (sLow[s0] > meta[i + 1].max_e) {
            meta[i + 1].uniq += 1;
            sLow[s0] = meta[i + 1].max_e;
          }
        }
      }
      const int sl = slot_of[R - 1];
      const double le = sE[sl];
      if (le < meta[R - 1].cutoff) {  // MedianEstimator::add_energy
        if (med_len < 4096) {
          M[med_len++] = le;
        } else if (gen_f64(l0, l1) < 1.0 / ((double)med_len + 1.0)) {
          M[gen_range(l0, l1, 0, (uint64_t)med_len)] = le;
        }
      }
      const ZMeta &lr = meta[R - 1];
      int split = 0;
      if (lr.uniq >= z.isbnb && sLow[sl] == lr.max_e) {
        const double mean_below = lr.below_t / (double)lr.below_c;
        if (mean_below + z.min_T < lr.cutoff && le < lr.cutoff) {
          if (R < RM) {
            split = 1;
          } else {
            ctl[3] = SADMC_WS_CAPACITY;
          }
        }
      }
      ctl[1] = split;
      ctl[2] = med_len;
    }
    __syncthreads();
    if (ctl[1]) {
      // MedianEstimator::median: the middle element of the sorted energies
      // by counting, then the nearest different energy above (or below)
      const int n = ctl[2], kmid = n / 2;
      for (int i = tid; i < n; i += nthr) {
        const double vi = M[i];
        int less = 0, leq = 0;
        for (int j = 0; j < n; j++) {
          const double v = M[j];
          less += v < vi;
          leq += v <= vi;
        }
        if (less <= kmid && kmid < leq) red[0] = vi;
      }
      __syncthreads();
      const double em = red[0];
      double nx = __builtin_inf(), pv = -__builtin_inf();
      for (int i = tid; i < n; i += nthr) {
        const double v = M[i];
        if (v > em && v < nx) nx = v;
        if (v < em && v > pv) pv = v;
      }
      for (int d = 1; d < 64; d <<= 1) {
        const double onx = __builtin_bit_cast(double, (uint64_t)__shfl_xor(
                                                           (long long)__builtin_bit_cast(uint64_t, nx), d, 64));
        const double opv = __builtin_bit_cast(double, (uint64_t)__shfl_xor(
                                                           (long long)__builtin_bit_cast(uint64_t, pv), d, 64));
        nx = onx < nx ? onx : nx;
        pv = opv > pv ? opv : pv;
      }
      if ((tid & 63) == 0) {
        red[2 + (tid >> 6)] = nx;
        red[20 + (tid >> 6)] = pv;
      }
      __syncthreads();
      if (tid == 0) {
        double next = __builtin_inf(), prev = -__builtin_inf();
        for (int w = 0; w < (nthr + 63) / 64; w++) {
          next = red[2 + w] < next ? red[2 + w] : next;
          prev = red[20 + w] > prev ? red[20 + w] : prev;
        }
        const double median_below =
            next < __builtin_inf() ? 0.5 * (em + next) : (prev > -__builtin_inf() ? 0.5 * (em + prev) : em);
        M[0] = median_below;
        med_len = 1;
        // the new replica: a clone of the lowest, decimated
        // (energy_replicas.rs:593-620, Replica::decimate :169-192)
        const int src = slot_of[R - 1];
        ZMeta nr = meta[R - 1];
        nr.max_e = meta[R - 1].cutoff;
        nr.cutoff = median_below;
        nr.upw = 0;
        if (nr.above_c > 1) {
          nr.above_t /= (double)nr.above_c;
          nr.above_sq /= (double)nr.above_c;
          nr.above_c = 1;
        }
        if (nr.below_c > 1) {
          nr.below_t /= (double)nr.below_c;
          nr.below_sq /= (double)nr.below_c;
          nr.below_c = 1;
        }
        if (nr.p_c > 1) {
          nr.p_t /= (double)nr.p_c;
          nr.p_c = 1;
        }
        nr.acc = 1;
        nr.rej = 1;
        nr.uniq = 1;
        nr.ts = meta[R - 1].ts * z.shrink;
        rng_jump(nr.rng0, nr.rng1);
        meta[R] = nr;
        sE[R] = sE[src];
        sErr[R] = sErr[src];
        sLow[R] = -__builtin_inf();
        slot_of[R] = R;
        rep_of[R] = R;
        ctl[2] = src;
        ctl[0] = R + 1;
      }
      __syncthreads();
      const int src = ctl[2];
      if (SYS == SYS_ISING) {
        if (tid == 0) spn[R] = spn[src];
      } else {
        for (int i = tid; i < 3 * npad; i += nthr) pos[(int64_t)R * 3 * npad + i] = pos[(int64_t)src * 3 * npad + i];
      }
    }
    if (tid == 0) {
      // Replica::occasional_update (energy_replicas.rs:246-290)
      const double mx = (double)z.steps;
      for (int k = 0; k < ctl[0]; k++) {
        ZMeta &m = meta[k];
        if (m.rej > 128 && m.acc > 128 && __builtin_isfinite(m.max_e)) {
          const double ratio = (double)m.acc / (double)m.rej;
          if (ratio < 0.5 || ratio > 2.0 * mx) {
            double adj = ratio < 0.5 ? ratio / sqrt(mx) : ratio * sqrt(mx);
            adj = adj > 2.0 ? 2.0 : (adj < 0.5 ? 0.5 : adj);
            m.ts *= adj;
            m.acc = 0;
            m.rej = 0;
          }
        }
      }
      moves += (uint64_t)R * (uint64_t)z.steps;
      *smoves = moves;
      ctl[1] = 0;
    }
    __syncthreads();
  }

  // write back: replica records (with their systems' energies), slot data,
  // the tables, the systems, the run's scalars
  for (int r = tid; r < RM; r += nthr) {
    const ZMeta m = meta[r];
    sadmc_zreplica &q = z.rep[S0 + r];
    const int s = slot_of[r];
    q.max_energy = m.max_e;
    q.cutoff_energy = m.cutoff;
    q.rejected_count = m.rej;
    q.accepted_count = m.acc;
    q.above_count = m.above_c;
    q.below_count = m.below_c;
    q.upwelling_count = m.upw;
    q.above_total = m.above_t;
    q.below_total = m.below_t;
    q.above_total_squared = m.above_sq;
    q.below_total_squared = m.below_sq;
    q.lowest_max_energy = sLow[s];
    q.unique_visitors = m.uniq;
    q.collecting_data = m.collecting;
    q.reserved = 0;
    q.rng[0] = m.rng0;
    q.rng[1] = m.rng1;
    q.translation_scale = m.ts;
    q.E = sE[s];
    q.error = sErr[s];
    q.pressure_total = m.p_t;
    q.pressure_count = m.p_c;
    z.slot_of[S0 + r] = s;
    z.slot[3 * (S0 + r)] = sE[r];
    z.slot[3 * (S0 + r) + 1] = sErr[r];
    z.slot[3 * (S0 + r) + 2] = sLow[r];
    if (SYS == SYS_ISING) a.spins[S0 + r] = spn[r];
  }
  if (SYS != SYS_ISING)
    for (int i = tid; i < 3 * RM * npad; i += nthr) a.pos[S0 * 3 * npad + i] = pos[i];
  if (tid == 0) {
    z.sim[sim].rng[0] = l0;
    z.sim[sim].rng[1] = l1;
    z.sim[sim].moves = moves;
    z.sim[sim].R = ctl[0];
    z.sim[sim].status = ctl[3];
    z.sim[sim].med_len = med_len;
  }
}

template <int SYS, int KMAX>
static hipError_t launch_zmc_t(const ZArgs &z, const SetArgs &a, int sims, int threads, uint64_t rounds,
                               hipStream_t s) {
  const size_t lds = zmc_lds_bytes(z.RMAX, SYS == SYS_ISING ? 0 : a.npad);
  if constexpr (SYS != SYS_ISING && KMAX >= 2 && KMAX <= 8) {
    if (a.G == 4 || a.G == 8 || a.G == 16) {
      if (a.G == 4) hipLaunchKernelGGL((k_zmc<SYS, KMAX, 4>), dim3(sims), dim3(threads), lds, s, z, a, rounds);
      if (a.G == 8) hipLaunchKernelGGL((k_zmc<SYS, KMAX, 8>), dim3(sims), dim3(threads), lds, s, z, a, rounds);
      if (a.G == 16) hipLaunchKernelGGL((k_zmc<SYS, KMAX, 16>), dim3(sims), dim3(threads), lds, s, z, a, rounds);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_zmc<SYS, KMAX>), dim3(sims), dim3(threads), lds, s, z, a, rounds);
  return hipGetLastError();
}

hipError_t launch_zmc(int sys, int kmax, const ZArgs &z, const SetArgs &a, int sims, int threads, uint64_t rounds,
                      hipStream_t s) {
  if (sys == SYS_ISING) return launch_zmc_t<SYS_ISING, 1>(z, a, sims, threads, rounds, s);
  if (sys == SYS_WCA) switch (kmax) {
      case 1: return launch_zmc_t<SYS_WCA, 1>(z, a, sims, threads, rounds, s);
      case 2: return launch_zmc_t<SYS_WCA, 2>(z, a, sims, threads, rounds, s);
      case 4: return launch_zmc_t<SYS_WCA, 4>(z, a, sims, threads, rounds, s);
      case 8: return launch_zmc_t<SYS_WCA, 8>(z, a, sims, threads, rounds, s);
      case 16: return launch_zmc_t<SYS_WCA, 16>(z, a, sims, threads, rounds, s);
      default: return hipErrorInvalidValue;
    }
  if (sys != SYS_LJ) return hipErrorInvalidValue;
  switch (kmax) {
    case 1: return launch_zmc_t<SYS_LJ, 1>(z, a, sims, threads, rounds, s);
    case 2: return launch_zmc_t<SYS_LJ, 2>(z, a, sims, threads, rounds, s);
    case 4: return launch_zmc_t<SYS_LJ, 4>(z, a, sims, threads, rounds, s);
    case 8: return launch_zmc_t<SYS_LJ, 8>(z, a, sims, threads, rounds, s);
    case 16: return launch_zmc_t<SYS_LJ, 16>(z, a, sims, threads, rounds, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace sadmc
