// EnergyMC over a histogram Binning (src/mc/energy_binning.rs with
// src/mc/binning/histogram.rs): the walker of the reference's `binning`
// binary, one walker per lane group like the main engine.
//
// The system part (plan_move, confirm, energies, pressure) is the main
// engine's; the bookkeeping is this file's.  Every lane of a group holds the
// walker's scalars (identical values), the bins live in HBM as one 80-byte
// record per bin.  The O(bins) scans the reference makes are either replaced
// by exact trackers of the extremes (BinCounts' min/max recomputations) or
// split over the group's lanes and combined with shuffles (set_lnw,
// count_states, the trackers' own rescans), so that every result is the
// sequential one.
#define SADMC_KERNEL_TU
#include <type_traits>

#include "sadmc_kernels.h"

namespace sadmc {

namespace {

enum { BC_LNW = 0, BC_EN = 1, BC_TF = 2, BC_HIST = 3, BC_PR = 4 };

// BinCounts' running statistics (histogram.rs:11-33, linear.rs:11-33):
// counts are integers in the histogram, fractions in the linear binning.
template <typename CT>
struct BC {
  double min_total, max_total, e_max_total, e_max_count;
  CT min_count, max_count;
  uint64_t total_count;
  int32_t exists;
};

// Counts live in the 64-bit count fields of the bin records: as integers,
// or as the bits of doubles.
template <typename CT>
__device__ __forceinline__ uint64_t ct_bits(CT v) {
  if constexpr (std::is_floating_point<CT>::value) {
    return __builtin_bit_cast(uint64_t, v);
  } else {
    return (uint64_t)v;
  }
}
template <typename CT>
__device__ __forceinline__ CT ct_from(uint64_t u) {
  if constexpr (std::is_floating_point<CT>::value) {
    return __builtin_bit_cast(double, u);
  } else {
    return (CT)u;
  }
}

__device__ __forceinline__ double shfl_f64(double v, int d) {
  return __builtin_bit_cast(double, (uint64_t)__shfl_xor((long long)__builtin_bit_cast(uint64_t, v), d, 64));
}

}  // namespace

// LIN: the linear interpolation of binning/linear.rs instead of the
// histogram of binning/histogram.rs.
template <int SYS, int KMAX, bool LIN>
struct BinWalker {
  using CT = typename std::conditional<LIN, double, uint64_t>::type;
  Engine<SYS, KMAX, 0, M_CAN> e;
  const BinArgs *Bp;
  sadmc_bbin *B;  // this walker's window of bins
  int lane, G;
  uint64_t moves, accepted;
  double acc_rate, tscale, usc;
  double min, width, min_e, max_e;
  int64_t len, base, cap;
  int32_t method, status;
  uint64_t num_states, tL;
  double min_T, too_lo, too_hi, tF, latest, t0, gamma;
  int32_t inv_t, has_min_gamma;
  double min_gamma;
  BC<CT> bc[5];
  // Exact extremes of each array kept beside the reference's statistics,
  // so that its min_of / max_of / min() rescans (histogram.rs:64-80,
  // 226-244) become lookups: the smallest total and how many bins hold it,
  // the smallest count and its multiplicity, and the largest ln w total.
  // A full rescan is needed only when the last bin holding an extreme
  // leaves it, after set_lnw / zero_out, and at the start of a launch.
  double tmin[5], lmax;
  uint64_t nmin[5], ncmin[5];
  CT cmin[5];
  // an extreme lost its last bin: rescan once the move's stores are all in
  // (a rescan between the two bins of a linear update would count the
  // second bin's new value twice)
  bool stale_t[5], stale_c[5];
  // the high_resolution side histogram (energy_binning.rs:125, 328-330):
  // counts in HR[hs.base + i], its statistics in *HS (read and written
  // back by each increment, so they hold no registers across the move loop
  // when the histogram is off), and the smallest count's multiplicity
  uint64_t *HR;
  sadmc_hires *HS;
  uint64_t hr_ncmin;

  SADMC_DF const BinArgs &args() const { return *Bp; }

  // ------------------------------------------------------------ state I/O
  SADMC_DF static void bc_load(BC<CT> &c, const sadmc_bcounts &s) {
    c.min_total = s.min_total;
    c.max_total = s.max_total;
    c.e_max_total = s.e_max_total;
    c.e_max_count = s.e_max_count;
    c.min_count = ct_from<CT>(s.min_count);
    c.max_count = ct_from<CT>(s.max_count);
    c.total_count = s.total_count;
    c.exists = s.exists;
  }
  SADMC_DF static void bc_store(const BC<CT> &c, sadmc_bcounts &s) {
    s.min_total = c.min_total;
    s.max_total = c.max_total;
    s.e_max_total = c.e_max_total;
    s.e_max_count = c.e_max_count;
    s.min_count = ct_bits<CT>(c.min_count);
    s.max_count = ct_bits<CT>(c.max_count);
    s.total_count = c.total_count;
    s.exists = c.exists;
    s.reserved = 0;
  }

  SADMC_DF void load(const BinArgs &ba, const SetArgs &sa, int walker, int lane_in_group, double *lds_group) {
    Bp = &ba;
    lane = lane_in_group;
    G = sa.G;
    e.load_system(sa, walker, lane_in_group, lds_group);
    const sadmc_bwalker_state &w = ba.ws[walker];
    e.s0 = w.rng[0];
    e.s1 = w.rng[1];
    e.E = w.E;
    e.err = w.error;
    if (SYS != SYS_ISING) {
      if (SYS == SYS_LJ) e.R2 = w.geom * w.geom;
      e.L = w.geom;
      e.halfL = 0.5 * w.geom;
    }
    e.status = 0;
    moves = w.moves;
    accepted = w.accepted_moves;
    acc_rate = w.acceptance_rate;
    tscale = w.translation_scale;
    usc = SYS == SYS_ISING ? 0.0 : uniform_scale(-tscale, tscale);
    min = w.min;
    width = w.width;
    min_e = w.min_e;
    max_e = w.max_e;
    len = w.len;
    base = w.base;
    cap = w.cap;
    method = w.method;
    status = w.status;
    num_states = w.num_states;
    tL = w.tL;
    min_T = w.min_T;
    too_lo = w.too_lo;
    too_hi = w.too_hi;
    tF = w.tF;
    latest = w.latest_parameter;
    t0 = w.t0;
    gamma = w.gamma;
    inv_t = w.inv_t;
    has_min_gamma = w.has_min_gamma;
    min_gamma = w.min_gamma;
    bc_load(bc[BC_LNW], w.lnw);
    bc_load(bc[BC_EN], w.energy);
    bc_load(bc[BC_TF], w.t_found);
    bc_load(bc[BC_HIST], w.hist);
    bc_load(bc[BC_PR], w.pressure);
    B = ba.bins + (int64_t)walker * ba.cap;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      stale_t[k] = false;
      stale_c[k] = false;
      if (bc[k].exists) rescan_all(k);
    }
    HR = nullptr;
    HS = nullptr;
    if (ba.hr) {
      HR = ba.hr + (int64_t)walker * ba.hr_cap;
      HS = ba.hrs + walker;
      sadmc_hires hs = *HS;
      hr_rescan(hs);
      *HS = hs;
    }
  }

  SADMC_DF void store(int walker) {
    if (SYS != SYS_ISING) {
      // the positions of the group's atoms, from the LDS mirror
      const SetArgs &sa = e.args();
      double *P = sa.pos + (int64_t)walker * 3 * sa.npad;
      for (int j = lane; j < sa.npad; j += G) {
        P[j] = e.lx[j];
        P[sa.npad + j] = e.ly[j];
        P[2 * sa.npad + j] = e.lz[j];
      }
    }
    if (lane != 0) return;
    if (SYS == SYS_ISING) e.args().spins[walker] = e.spins;
    sadmc_bwalker_state &w = args().ws[walker];
    w.rng[0] = e.s0;
    w.rng[1] = e.s1;
    w.E = e.E;
    w.error = e.err;
    w.moves = moves;
    w.accepted_moves = accepted;
    w.acceptance_rate = acc_rate;
    w.translation_scale = tscale;
    w.min = min;
    w.width = width;
    w.min_e = min_e;
    w.max_e = max_e;
    w.len = len;
    w.base = base;
    w.method = method;
    w.status = status;
    w.num_states = num_states;
    w.tL = tL;
    w.min_T = min_T;
    w.too_lo = too_lo;
    w.too_hi = too_hi;
    w.tF = tF;
    w.latest_parameter = latest;
    w.t0 = t0;
    w.gamma = gamma;
    bc_store(bc[BC_LNW], w.lnw);
    bc_store(bc[BC_EN], w.energy);
    bc_store(bc[BC_TF], w.t_found);
    bc_store(bc[BC_HIST], w.hist);
    bc_store(bc[BC_PR], w.pressure);
    if (status) *args().flag = 1;
  }

  // ------------------------------------------------------- bin records
  SADMC_DF static double &tot(sadmc_bbin &b, int k) {
    return k == BC_LNW ? b.lnw_total : k == BC_EN ? b.energy_total : k == BC_TF ? b.t_found_total : b.pressure_total;
  }
  SADMC_DF static uint64_t &cnt(sadmc_bbin &b, int k) {
    return k == BC_LNW   ? b.lnw_count
           : k == BC_EN  ? b.energy_count
           : k == BC_TF  ? b.t_found_count
           : k == BC_HIST ? b.hist_count
                          : b.pressure_count;
  }
  SADMC_DF static CT cget(sadmc_bbin &b, int k) { return ct_from<CT>(cnt(b, k)); }
  SADMC_DF static void cset(sadmc_bbin &b, int k, CT v) { cnt(b, k) = ct_bits<CT>(v); }
  // the "hist" extra only ever adds 0.0 to totals that start at 0.0
  SADMC_DF double total_at(int k, int64_t i) const { return k == BC_HIST ? 0.0 : tot(B[base + i], k); }
  SADMC_DF CT count_at(int k, int64_t i) const { return cget(B[base + i], k); }

  // histogram.rs:129-140; -1 for the out-of-range cases
  SADMC_DF int64_t index(double en) const {
    if (en < min) return -1;
    const double i = (en - min) / width;
    if (i == (double)len) return len == 0 ? -1 : (int64_t)i - 1;
    if (!(i < 9.2e18)) return INT64_MAX;
    return (int64_t)i;
  }
  SADMC_DF double fidx(double en) const { return (en - min) / width; }  // linear.rs:201-203
  SADMC_DF double energy_of(int64_t i) const { return min + ((double)i + 0.5) * width; }
  // BinCounts::get_total / get_count of linear.rs:139-152 over the
  // interpretation of linear.rs:84-98, in the iterator's order
  SADMC_DF double lin_interp(bool counts, double f) const {
    double t = 0.0;
    if (len == 0 || f < -1.0) return t;
    auto val = [&](int64_t i) -> double {
      return counts ? (double)count_at(BC_LNW, i) : B[base + i].lnw_total;
    };
    if (f < 0.0) {
      t += val(0) * (1.0 - (-f));
    } else if (f < (double)len - 1.0) {
      const int64_t i = (int64_t)f;
      const double o = f - (double)i;
      t += val(i + 1) * o;
      t += val(i) * (1.0 - o);
    } else if (f < (double)len) {
      t += val(len - 1) * (1.0 - (f - ((double)len - 1.0)));
    }
    return t;
  }
  SADMC_DF double get_lnw(double en) const {
    if (LIN) return lin_interp(false, fidx(en));
    const int64_t i = index(en);
    return (i >= 0 && i < len) ? B[base + i].lnw_total : 0.0;
  }
  // get_count as PerEnergy: count / width
  SADMC_DF double get_count(double en) const {
    if (LIN) return lin_interp(true, fidx(en)) / width;
    const int64_t i = index(en);
    return (double)((i >= 0 && i < len) ? (uint64_t)count_at(BC_LNW, i) : 0) / width;
  }

  // ---------------------------------------------------- group reductions
  SADMC_DF uint64_t group_sum_u64(uint64_t v) const {
    for (int d = 1; d < G; d <<= 1) v += (uint64_t)__shfl_xor((long long)v, d, 64);
    return v;
  }
  // Rescans, four bins per lane per step so that their loads overlap.
  // Smallest total of array k and its multiplicity (+inf, 0 when empty).
  SADMC_DF void rescan_total(int k) {
    double m = __builtin_inf();
    uint64_t c = 0;
    if (k != BC_HIST) {
      for (int64_t j0 = lane; j0 < len; j0 += 4 * G) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int64_t j = j0 + u * G;
          v[u] = j < len ? total_at(k, j) : __builtin_inf();
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (j0 + u * G >= len) continue;
          if (v[u] < m) {
            m = v[u];
            c = 1;
          } else if (v[u] == m) {
            c += 1;
          }
        }
      }
    } else if (lane < len) {
      m = 0.0;
      c = (uint64_t)((len - lane + G - 1) / G);
    }
    for (int d = 1; d < G; d <<= 1) {
      const double om = shfl_f64(m, d);
      const uint64_t oc = (uint64_t)__shfl_xor((long long)c, d, 64);
      if (om < m) {
        m = om;
        c = oc;
      } else if (om == m) {
        c += oc;
      }
    }
    tmin[k] = m;
    nmin[k] = c;
  }
  SADMC_DF void rescan_count(int k) {
    CT big;
    if constexpr (LIN) {
      big = __builtin_inf();
    } else {
      big = ~0ULL;
    }
    CT m = big;
    uint64_t c = 0;
    for (int64_t j0 = lane; j0 < len; j0 += 4 * G) {
      CT v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = j0 + u * G;
        v[u] = j < len ? count_at(k, j) : big;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (j0 + u * G >= len) continue;
        if (v[u] < m) {
          m = v[u];
          c = 1;
        } else if (v[u] == m) {
          c += 1;
        }
      }
    }
    for (int d = 1; d < G; d <<= 1) {
      const CT om = ct_from<CT>((uint64_t)__shfl_xor((long long)ct_bits<CT>(m), d, 64));
      const uint64_t oc = (uint64_t)__shfl_xor((long long)c, d, 64);
      if (om < m) {
        m = om;
        c = oc;
      } else if (om == m) {
        c += oc;
      }
    }
    cmin[k] = m;
    ncmin[k] = c;
  }
  SADMC_DF void rescan_lmax() {
    double m = -__builtin_inf();
    for (int64_t j0 = lane; j0 < len; j0 += 4 * G) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = j0 + u * G;
        v[u] = j < len ? B[base + j].lnw_total : -__builtin_inf();
      }
#pragma unroll
      for (int u = 0; u < 4; u++) m = v[u] > m ? v[u] : m;
    }
    for (int d = 1; d < G; d <<= 1) {
      const double om = shfl_f64(m, d);
      m = om > m ? om : m;
    }
    lmax = m;
  }
  SADMC_DF void rescan_all(int k) {
    rescan_total(k);
    rescan_count(k);
    if (k == BC_LNW) rescan_lmax();
  }
  // One bin of array k goes from v to w (totals) / from a to b (counts);
  // settle() afterwards, when every bin of the update is stored.
  SADMC_DF void track_total(int k, double v, double w) {
    if (stale_t[k]) return;
    if (w < tmin[k]) {
      tmin[k] = w;
      nmin[k] = 1;
    } else if (w == tmin[k]) {
      if (v != w) nmin[k] += 1;
    } else if (v == tmin[k]) {
      if (--nmin[k] == 0) stale_t[k] = true;
    }
  }
  SADMC_DF void track_count(int k, CT a, CT b) {
    if (stale_c[k]) return;
    if (b < cmin[k]) {
      cmin[k] = b;
      ncmin[k] = 1;
    } else if (b == cmin[k]) {
      if (a != b) ncmin[k] += 1;
    } else if (a == cmin[k]) {
      if (--ncmin[k] == 0) stale_c[k] = true;
    }
  }
  SADMC_DF void settle(int k) {
    if (stale_t[k]) {
      rescan_total(k);
      stale_t[k] = false;
    }
    if (stale_c[k]) {
      rescan_count(k);
      stale_c[k] = false;
    }
  }
  // a zero bin joins array k
  SADMC_DF void track_new_zero(int k) {
    if (0.0 < tmin[k]) {
      tmin[k] = 0.0;
      nmin[k] = 1;
    } else if (0.0 == tmin[k]) {
      nmin[k] += 1;
    }
    if ((CT)0 < cmin[k]) {
      cmin[k] = 0;
      ncmin[k] = 1;
    } else if ((CT)0 == cmin[k]) {
      ncmin[k] += 1;
    }
    if (k == BC_LNW && !(lmax >= 0.0)) lmax = 0.0;
  }

  // ---------------------------------------- the high_resolution histogram
  // The smallest count and its multiplicity (a count never exceeds the
  // number of moves, so ~0 stands for an empty histogram).
  SADMC_DF void hr_rescan(sadmc_hires &hs) {
    uint64_t m = ~0ULL, c = 0;
    for (int64_t j0 = lane; j0 < hs.len; j0 += 4 * G) {
      uint64_t v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = j0 + u * G;
        v[u] = j < hs.len ? HR[hs.base + j] : ~0ULL;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (j0 + u * G >= hs.len) continue;
        if (v[u] < m) {
          m = v[u];
          c = 1;
        } else if (v[u] == m) {
          c += 1;
        }
      }
    }
    for (int d = 1; d < G; d <<= 1) {
      const uint64_t om = (uint64_t)__shfl_xor((long long)m, d, 64);
      const uint64_t oc = (uint64_t)__shfl_xor((long long)c, d, 64);
      if (om < m) {
        m = om;
        c = oc;
      } else if (om == m) {
        c += oc;
      }
    }
    if (hs.len > 0) hs.min_count = m;
    hr_ncmin = c;
  }
  // insert_zero / push_zero (histogram.rs:47-58) on the side histogram
  SADMC_DF void hr_new_zero(sadmc_hires &hs, int64_t p) {
    HR[p] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (hs.min_count == 0) {
      hr_ncmin += 1;
    } else {
      hs.min_count = 0;
      hr_ncmin = 1;
    }
    hs.len += 1;
  }
  // Bins::increment_count(e, 0.0) (histogram.rs:182-192, 147-167, 59-81):
  // min_e / max_e, the window grown to hold e, then the bin's count and the
  // count statistics (the totals stay 0)
  SADMC_DF bool hr_increment(double en) {
    sadmc_hires hs = *HS;
    const bool ok = hr_increment(hs, en);
    *HS = hs;
    return ok;
  }
  SADMC_DF bool hr_increment(sadmc_hires &hs, double en) {
    if (en > hs.max_e) hs.max_e = en;
    if (en < hs.min_e) hs.min_e = en;
    if (hs.len == 0) {
      hs.min = floor(en / hs.width) * hs.width;
      const int64_t b0 = (int64_t)floor((hs.min - args().e_lo) / hs.width) + 4;
      hs.base = b0 < 0 ? 0 : (b0 >= args().hr_cap ? args().hr_cap - 1 : b0);
    }
    while (en < hs.min) {
      if (hs.base == 0) {
        status = SADMC_WS_CAPACITY;
        return false;
      }
      hs.base -= 1;
      hr_new_zero(hs, hs.base);
      hs.min -= hs.width;
    }
    while (en >= hs.min + hs.width * (double)hs.len) {
      if (hs.base + hs.len >= args().hr_cap) {
        status = SADMC_WS_CAPACITY;
        return false;
      }
      hr_new_zero(hs, hs.base + hs.len);
    }
    // histogram.rs:129-140
    const double fi = (en - hs.min) / hs.width;
    const int64_t idx = fi == (double)hs.len ? (int64_t)fi - 1 : (int64_t)fi;
    if (!(en >= hs.min) || idx < 0 || idx >= hs.len) {  // an index panic in the reference
      status = SADMC_WS_INDEX;
      return false;
    }
    uint64_t &r = HR[hs.base + idx];
    const uint64_t old_count = r, new_count = old_count + 1;
    r = new_count;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    hs.total_count += 1;
    if (old_count == hs.min_count && --hr_ncmin == 0) hr_rescan(hs);
    if (new_count > hs.max_count) {
      hs.max_count = new_count;
      hs.e_max_count = en;
    }
    return true;
  }

  // ------------------------------------------------------------ growth
  SADMC_DF void zero_bin(int64_t p) {
    {
      sadmc_bbin z;
      z.lnw_total = 0.0;
      z.lnw_count = 0;
      z.energy_total = 0.0;
      z.energy_count = 0;
      z.t_found_total = 0.0;
      z.t_found_count = 0;
      z.hist_count = 0;
      z.pressure_total = 0.0;
      z.pressure_count = 0;
      z.reserved = 0;
      B[p] = z;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  SADMC_DF void grown() {
#pragma unroll
    for (int k = 0; k < 5; k++)
      if (bc[k].exists) {
        bc[k].min_total = 0.0;
        bc[k].min_count = 0;
        track_new_zero(k);
      }
    len += 1;
  }
  // histogram.rs:141-158 / linear.rs:204-222 (insert_zero / push_zero on
  // lnw and every extra; the linear binning keeps one more bin above)
  SADMC_DF bool prep(double en) {
    if (len == 0) {
      min = floor(en / width) * width;
      int64_t b0 = (int64_t)floor((min - args().e_lo) / width) + 4;
      base = b0 < 0 ? 0 : (b0 >= cap ? cap - 1 : b0);
    }
    while (en < min) {
      if (base == 0) {
        status = SADMC_WS_CAPACITY;
        return false;
      }
      base -= 1;
      zero_bin(base);
      grown();
      min -= width;
    }
    while (LIN ? en >= min + width * ((double)len - 1.0) : en >= min + width * (double)len) {
      if (base + len >= cap) {
        status = SADMC_WS_CAPACITY;
        return false;
      }
      zero_bin(base + len);
      grown();
    }
    return true;
  }

  // BinCounts::increment_count (histogram.rs:59-81) on lnw
  SADMC_DF void lnw_increment(double en, int64_t idx, double value) {
    BC<CT> &c = bc[BC_LNW];
    sadmc_bbin &r = B[base + idx];
    c.total_count += 1;
    const CT old_count = cget(r, BC_LNW);
    const CT new_count = old_count + 1;
    const double old_total = r.lnw_total;
    const double new_total = old_total + value;
    cset(r, BC_LNW, new_count);
    r.lnw_total = new_total;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    track_total(BC_LNW, old_total, new_total);
    track_count(BC_LNW, old_count, new_count);
    settle(BC_LNW);
    if (new_total > lmax) lmax = new_total;
    // max_of(total) == total[idx], min_of(total) and min(count) after the
    // update are the tracked extremes
    if (new_total > c.max_total && lmax == new_total) {
      c.max_total = new_total;
      c.e_max_total = en;
    }
    if (old_total == c.min_total) c.min_total = tmin[BC_LNW];
    if (old_count == c.min_count) c.min_count = cmin[BC_LNW];
    if (new_count > c.max_count) {
      c.max_count = new_count;
      c.e_max_count = en;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }

  // BinCounts::increment_count of linear.rs:99-137 on lnw: the weight
  // splits over bins idx and idx + 1
  SADMC_DF void lnw_increment_lin(double elo, double ehi, double f, double value) {
    BC<CT> &c = bc[BC_LNW];
    const int64_t idx = (int64_t)f;
    const double offset = f - (double)idx;
    sadmc_bbin &r0 = B[base + idx];
    sadmc_bbin &r1 = B[base + idx + 1];
    c.total_count += 1;
    const CT oc = cget(r0, BC_LNW), opc = cget(r1, BC_LNW);
    const CT c0 = oc + (1.0 - offset), c1 = opc + offset;
    const double ot = r0.lnw_total, opt = r1.lnw_total;
    const double t0v = ot + value * (1.0 - offset), t1v = opt + value * offset;
    cset(r0, BC_LNW, c0);
    cset(r1, BC_LNW, c1);
    r0.lnw_total = t0v;
    r1.lnw_total = t1v;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    track_total(BC_LNW, ot, t0v);
    track_total(BC_LNW, opt, t1v);
    track_count(BC_LNW, oc, c0);
    track_count(BC_LNW, opc, c1);
    settle(BC_LNW);
    if (t0v > lmax) lmax = t0v;
    if (t1v > lmax) lmax = t1v;
    if (t0v > c.max_total && lmax == t0v) {
      c.max_total = t0v;
      c.e_max_total = elo;
    }
    if (t1v > c.max_total && lmax == t1v) {
      c.max_total = t1v;
      c.e_max_total = ehi;
    }
    if (ot == c.min_total || opt == c.min_total) c.min_total = tmin[BC_LNW];
    if (oc == c.min_count || opc == c.min_count) c.min_count = cmin[BC_LNW];
    if (c0 > c.max_count) {
      c.max_count = c0;
      c.e_max_count = elo;
    }
    if (c1 > c.max_count) {
      c.max_count = c1;
      c.e_max_count = ehi;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }

  // Bins::accumulate_extra (histogram.rs:219-245, linear.rs:267-310)
  SADMC_DF bool accumulate(int k, double en, double value) {
    if (!prep(en)) return false;
    int64_t idx;
    double f = 0.0;
    if (LIN) {
      f = fidx(en);
      idx = (int64_t)f;
      if (!(f >= 0.0) || idx + 1 >= len) {  // the reference's assert / index panic
        status = SADMC_WS_INDEX;
        return false;
      }
    } else {
      idx = index(en);
      if (idx < 0 || idx >= len) {  // the reference's assert (histogram.rs:222)
        status = SADMC_WS_INDEX;
        return false;
      }
    }
    BC<CT> &c = bc[k];
    if (!c.exists) {  // BinCounts::new over the current bins
      tmin[k] = 0.0;
      nmin[k] = (uint64_t)len;
      cmin[k] = 0;
      ncmin[k] = (uint64_t)len;
      c.exists = 1;
      c.min_total = 0.0;
      c.max_total = 0.0;
      c.e_max_total = -__builtin_inf();
      c.e_max_count = -__builtin_inf();
      c.min_count = 0;
      c.max_count = 0;
      c.total_count = 0;
    }
    c.total_count += 1;
    if constexpr (!LIN) {
      sadmc_bbin &r = B[base + idx];
      const CT old_count = cget(r, k);
      const CT new_count = old_count + 1;
      const double old_total = k == BC_HIST ? 0.0 : tot(r, k);
      const double new_total = old_total + value;
      cset(r, k, new_count);
      if (k != BC_HIST) tot(r, k) = new_total;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      track_count(k, old_count, new_count);
      if (k != BC_HIST) track_total(k, old_total, new_total);
      settle(k);
      if (new_count == c.min_count + 1) c.min_count = cmin[k];
      if (new_count > c.max_count) c.max_count = new_count;
      if (new_total > c.max_total) c.max_total = new_total;
      if (old_total == c.min_total) c.min_total = k == BC_HIST ? 0.0 : tmin[k];
    } else {
      const double off = f - (double)idx;
      sadmc_bbin &r0 = B[base + idx];
      sadmc_bbin &r1 = B[base + idx + 1];
      const CT oc = cget(r0, k), opc = cget(r1, k);
      const CT c0 = oc + (1.0 - off), c1 = opc + off;
      const double ot = k == BC_HIST ? 0.0 : tot(r0, k), opt = k == BC_HIST ? 0.0 : tot(r1, k);
      const double t0v = ot + value * (1.0 - off), t1v = opt + value * off;
      cset(r0, k, c0);
      cset(r1, k, c1);
      if (k != BC_HIST) {
        tot(r0, k) = t0v;
        tot(r1, k) = t1v;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      track_count(k, oc, c0);
      track_count(k, opc, c1);
      if (k != BC_HIST) {
        track_total(k, ot, t0v);
        track_total(k, opt, t1v);
      }
      settle(k);
      if (oc == c.min_count || opc == c.min_count) c.min_count = cmin[k];
      if (c0 > c.max_count) c.max_count = c0;
      if (c1 > c.max_count) c.max_count = c1;
      if (t0v > c.max_total) c.max_total = t0v;
      if (t1v > c.max_total) c.max_total = t1v;
      if (ot == c.min_total || opt == c.min_total) c.min_total = k == BC_HIST ? 0.0 : tmin[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return true;
  }

  // Bins::zero_out_extra (histogram.rs:246-259) on "hist"
  SADMC_DF void zero_out_hist() {
    BC<CT> &c = bc[BC_HIST];
    if (!c.exists) return;
    for (int64_t j = lane; j < len; j += G) cset(B[base + j], BC_HIST, (CT)0);
    c.min_total = 0.0;
    c.max_total = -__builtin_inf();
    c.min_count = 0;
    c.max_count = 0;
    c.total_count = 0;
    cmin[BC_HIST] = 0;
    ncmin[BC_HIST] = (uint64_t)len;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }

  // count_states with e >= too_lo && e <= too_hi (histogram.rs:191-200)
  SADMC_DF uint64_t count_states() const {
    uint64_t t = 0;
    for (int64_t j = lane; j < len; j += G) {
      const double en = energy_of(j);
      if (en >= too_lo && en <= too_hi) t += 1;
    }
    return group_sum_u64(t);
  }

  // -------------------------------------------------------------- method
  // energy_binning.rs:503-530
  SADMC_DF double gamma_now() const {
    if (method == SADMC_SAD) {
      const double ns = (double)num_states;
      if (latest * tF * ns == 0.0) return 0.0;
      const double t = (double)moves;
      return (latest + t / tF) / (latest + t / ns * (t / tF));
    }
    if (method == SADMC_SAMC) {
      const double t = (double)moves;
      return t > t0 ? t0 / t : 1.0;
    }
    return gamma;
  }

  // energy_binning.rs:276-319
  SADMC_DF bool reject(double e1, double e2) {
    double lnw1 = get_lnw(e1), lnw2 = get_lnw(e2);
    if (method == SADMC_SAD) {
      if (e1 < too_lo)
        lnw1 = get_lnw(too_lo) + (e1 - too_lo) / min_T;
      else if (e1 > too_hi)
        lnw1 = get_lnw(too_hi);
      if (e2 < too_lo)
        lnw2 = get_lnw(too_lo) + (e2 - too_lo) / min_T;
      else if (e2 > too_hi)
        lnw2 = get_lnw(too_hi);
      const bool rejected = lnw2 > lnw1 && gen_f64(e.s0, e.s1) > exp(lnw1 - lnw2);
      if (!rejected && get_count(e2) == 0.0 && e2 < too_hi && e2 > too_lo) tL = moves;
      return rejected;
    }
    return lnw2 > lnw1 && gen_f64(e.s0, e.s1) > exp(lnw1 - lnw2);
  }

  // set_lnw (histogram.rs:182-190, linear.rs:254-262) for SAD's two range
  // extensions (in the linear binning a bin's count is read at its centre,
  // from it and the next bin, before that next bin may be cleared: each
  // step's loads come before its stores); bins
  // above `edge` (up) or below it (down) with a nonzero count take `v` (up)
  // or v + (e - edge)/min_T (down) and lose their count.
  SADMC_DF void set_lnw_sad(bool up, double edge, double v) {
    for (int64_t j = lane; j < len; j += G) {
      const double en = energy_of(j);
      const bool side = up ? en > edge : en < edge;
      if (side && get_count(en) > 0.0) {
        sadmc_bbin &r = B[base + j];
        r.lnw_total = up ? v : v + (en - edge) / min_T;
        cset(r, BC_LNW, (CT)0);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    rescan_all(BC_LNW);
  }

  // energy_binning.rs:320-500
  SADMC_DF bool update(double en) {
    const double g = gamma_now();
    const double old_highest = (double)bc[BC_LNW].max_count / width;
    const double old_here = get_count(en);
    if (en > max_e) max_e = en;
    if (en < min_e) min_e = en;
    if (!prep(en)) return false;
    if constexpr (LIN) {  // linear.rs:237-253
      const double f = fidx(en);
      const int64_t ii = (int64_t)f;
      if (!(f >= 0.0) || ii + 1 >= len) {  // an index panic in the reference
        status = SADMC_WS_INDEX;
        return false;
      }
      lnw_increment_lin(energy_of(ii), energy_of(ii + 1), f, g * 1.0 / width);
    } else {
      const int64_t idx = index(en);
      if (idx < 0 || idx >= len) {  // an index panic in the reference
        status = SADMC_WS_INDEX;
        return false;
      }
      lnw_increment(en, idx, g);
    }
    if (HR && !hr_increment(en)) return false;
    if (method == SADMC_SAD) {
      const double here = get_count(en);
      if (old_here == 0.0 && !accumulate(BC_TF, en, (double)moves)) return false;
      if (here > old_highest) {
        if (en > too_hi) {
          set_lnw_sad(true, too_hi, get_lnw(too_hi));
          latest = (en - too_lo) / min_T;
          tL = moves;
          too_hi = en;
          num_states = count_states();
        } else if (en < too_lo) {
          set_lnw_sad(false, too_lo, get_lnw(too_lo));
          latest = (too_hi - en) / min_T;
          tL = moves;
          too_lo = en;
          num_states = count_states();
        }
      }
      if (tL == moves) {
        const double old_tF = tF;
        tF = bc[BC_TF].exists ? bc[BC_TF].max_total : 0.0;
        if (old_tF != tF && args().move_plan == SADMC_ACCEPTANCE_RATE) {
          double s = acc_rate / args().move_param;
          s = s < 0.8 ? 0.8 : (s > 1.2 ? 1.2 : s);
          tscale *= s;
          if (SYS != SYS_ISING) usc = uniform_scale(-tscale, tscale);
        }
      }
    } else if (method == SADMC_WL) {
      const double old_lowest = bc[BC_HIST].exists ? (double)bc[BC_HIST].min_count / width : 0.0;
      if (!accumulate(BC_HIST, en, 0.0)) return false;
      if (has_min_gamma && gamma < min_gamma) return true;
      const double lowest = (double)bc[BC_HIST].min_count / width;
      const bool ok_lo = !args().has_min_e || get_count(args().min_e) > 0.0;
      const bool ok_hi = !args().has_max_e || get_count(args().max_e) > 0.0;
      if (lowest > old_lowest && ok_lo && ok_hi) {
        const double mean = (double)bc[BC_HIST].total_count / (width * (double)len);
        if ((inv_t && lowest > 0.0) || lowest >= 0.8 * mean) {
          gamma *= 0.5;
          zero_out_hist();
          if (has_min_gamma && gamma < min_gamma) gamma = 0.0;
        }
        if (inv_t && gamma < (double)len / (double)moves) {
          method = SADMC_SAMC;
          t0 = (double)len;
        }
      }
    }
    return true;
  }

  // energy_binning.rs:566-611
  SADMC_DF void move_once() {
    moves += 1;
    if (SYS != SYS_ISING && moves % 100000000ULL == 0) {
      e.verify_energy();
      if (e.status) {
        status = e.status;
        return;
      }
    }
    const double e1 = e.E;
    if (!accumulate(BC_EN, e1, e1)) return;
    if (SYS == SYS_WCA && args().collect_pressure && moves % ((uint64_t)e.n * (uint64_t)e.n) == 0) {
      const double pv = e.pressure();
      if (!accumulate(BC_PR, e1, pv)) return;
    }
    const double rs = sqrt(1.0 / (double)moves);
    acc_rate *= 1.0 - rs;
    double e2;
    if (e.cube_plan(tscale, usc, e2)) {
      bool oob = false;
      if (args().has_max_e) oob = e2 > args().max_e && e2 > e1;
      if (args().has_min_e) oob = oob || (e2 < args().min_e && e2 < e1);
      if (!oob && !reject(e1, e2)) {
        accepted += 1;
        acc_rate += rs;
        e.confirm();
      }
    }
    update(e.E);
  }

  // energy_binning.rs:533-564: the greedy descent below max_allowed_energy
  // and the initial method state (the histogram stays empty until the
  // first move).
  SADMC_DF void from_params() {
    e.E = e.full_energy();
    e.err = 0.0;
    if (args().has_max_e) {
      const double sc = SYS == SYS_ISING ? 0.0 : uniform_scale(-0.05, 0.05);
      for (uint64_t k = 0; k < 100000000ULL; k++) {
        double newe;
        if (e.cube_plan(0.05, sc, newe)) {
          if (newe < e.E) e.confirm();
          if (e.E < args().max_e) break;
        }
      }
    }
    const double E = e.E;
    too_lo = E;
    too_hi = E;
    num_states = 0;
    tL = 0;
    tF = 0.0;
    latest = 0.0;
    min = (round(E / width) - 0.5) * width;  // Bins::new (histogram.rs:162-170)
    min_e = E;
    max_e = E;
    len = 0;
    moves = 0;
    accepted = 0;
    acc_rate = 0.5;
    tscale = args().move_plan == SADMC_TRANSLATION_SCALE ? args().move_param : 0.05;
    // ln w's BinCounts::new(0); the extras do not exist yet (all zero)
#pragma unroll
    for (int k = 0; k < 5; k++) {
      bc[k].exists = 0;
      bc[k].min_total = 0.0;
      bc[k].max_total = 0.0;
      bc[k].e_max_total = 0.0;
      bc[k].e_max_count = 0.0;
      bc[k].min_count = 0;
      bc[k].max_count = 0;
      bc[k].total_count = 0;
    }
    bc[BC_LNW].exists = 1;
    bc[BC_LNW].e_max_total = -__builtin_inf();
    bc[BC_LNW].e_max_count = -__builtin_inf();
    if (HR) {  // Bins::new(system.energy(), de) (energy_binning.rs:568-570)
      sadmc_hires hs;
      hs.width = args().hr_de;
      hs.min = (round(E / hs.width) - 0.5) * hs.width;
      hs.min_e = E;
      hs.max_e = E;
      hs.len = 0;
      hs.base = 0;
      hs.total_count = 0;
      hs.min_count = 0;
      hs.max_count = 0;
      hs.e_max_count = -__builtin_inf();
      *HS = hs;
      hr_ncmin = 0;
    }
    status = e.status;
  }
};

template <int SYS, int KMAX, bool LIN>
__global__ __launch_bounds__(64) void k_bin(BinArgs ba, SetArgs sa, uint64_t nmoves, int init) {
  extern __shared__ double lds[];
  const int G = sa.G;
  const int group = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  const int walker = blockIdx.x * (64 / G) + group;
  if (walker >= sa.W) return;
  BinWalker<SYS, KMAX, LIN> b;
  b.load(ba, sa, walker, lane, lds + group * 3 * sa.npad);
  if (init) {
    b.from_params();
  } else {
    for (uint64_t k = 0; k < nmoves && !b.status; k++) b.move_once();
  }
  b.store(walker);
}

template <int SYS, int KMAX>
static hipError_t launch_bin_t(const BinArgs &ba, const SetArgs &sa, uint64_t moves, int init, hipStream_t s) {
  const int per = 64 / sa.G;
  const int blocks = (sa.W + per - 1) / per;
  const size_t lds = SYS == SYS_ISING ? 0 : sizeof(double) * 3 * (size_t)per * sa.npad;
  if (ba.linear)
    hipLaunchKernelGGL((k_bin<SYS, KMAX, true>), dim3(blocks), dim3(64), lds, s, ba, sa, moves, init);
  else
    hipLaunchKernelGGL((k_bin<SYS, KMAX, false>), dim3(blocks), dim3(64), lds, s, ba, sa, moves, init);
  return hipGetLastError();
}

template <int SYS>
static hipError_t launch_bin_s(int kmax, const BinArgs &ba, const SetArgs &sa, uint64_t moves, int init,
                               hipStream_t s) {
  switch (kmax) {
    case 1: return launch_bin_t<SYS, 1>(ba, sa, moves, init, s);
    case 2: return launch_bin_t<SYS, 2>(ba, sa, moves, init, s);
    case 4: return launch_bin_t<SYS, 4>(ba, sa, moves, init, s);
    case 8: return launch_bin_t<SYS, 8>(ba, sa, moves, init, s);
    case 16: return launch_bin_t<SYS, 16>(ba, sa, moves, init, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_binning(int sys, int kmax, const BinArgs &ba, const SetArgs &sa, uint64_t moves, int init,
                          hipStream_t s) {
  if (sys == SYS_ISING) return launch_bin_t<SYS_ISING, 1>(ba, sa, moves, init, s);
  if (sys == SYS_LJ) return launch_bin_s<SYS_LJ>(kmax, ba, sa, moves, init, s);
  if (sys == SYS_WCA) return launch_bin_s<SYS_WCA>(kmax, ba, sa, moves, init, s);
  return hipErrorInvalidValue;
}

}  // namespace sadmc
