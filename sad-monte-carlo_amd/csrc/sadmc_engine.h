// Device walker engine: one reference `EnergyMC::move_once`
// (src/mc/energy.rs:904-974) per call, for a group of G lanes that share one
// walker.
//
// Layout.  Every lane of a group holds the walker's hot scalar state in
// registers (identical values: all lanes execute the same bookkeeping, so no
// broadcast is ever needed).  For LJ/WCA, lane l of the group owns atoms l,
// l+G, l+2G, ... (K slots, SoA registers); the group also keeps an LDS mirror
// of all N positions so any lane can read the moved atom.  The energy
// difference of a move is a per-lane partial sum over the lane's atoms
// followed by an xor butterfly across the G lanes.  Ising uses one lane per
// walker (G = 1) with the N*N <= 64 spins bit-packed in one 64-bit register.
//
// Bins.  A walker's bins are a fixed window of `cap` 64-byte records in HBM.
// Logical bin i lives in physical slot base + i; the reference's front
// insertion (energy.rs:400-420) becomes base -= 1, and the slots outside the
// window are zero from allocation, which is exactly the value the reference
// inserts.  The current bin's record is held in registers; the record of a
// proposed bin is loaded once (one dependent memory round trip per move) and
// becomes the current record if the move lands there.  Indices that only
// change when the window or the SAD range changes (the current bin, the bins
// of too_lo and too_hi, and ln w at those two bins) are cached and refreshed
// exactly when their inputs change, so each move computes one bin index.
// Periodic checks (verify_energy, the fixed refresh, WCA pressure) use
// next-trigger counters instead of a 64-bit modulo per move.
#pragma once

#include "sadmc_common.h"
#include "sadmc_proposal.h"
#include "../../include/sadmc.h"

namespace sadmc {

enum { SYS_ISING = 0, SYS_LJ = 1, SYS_WCA = 2 };
// Method families compiled separately: SAD; SAMC; WL and 1/t-WL (which may
// switch to SAMC, energy.rs:747-756); canonical.
enum { M_SAD = 0, M_SAMC = 1, M_WL = 2, M_CAN = 3 };

struct PExtra {
  uint64_t count;
  double total;
};

// Bin records are split by when they are needed.  The hot part (16 bytes)
// is all the accept/reject decision reads; the cold part is read-modify-
// written one move later, so its load overlaps the next move's pair sums.
// WL's own histogram lives in a third array that only WL kernels touch.
struct HotBin {
  double lnw;
  uint64_t hist;
};
struct ColdBin {
  double etot, esq;
  uint64_t rt, vep;
  uint64_t tfound, pad;
};

// Arguments of one walker set, passed by value to the kernels (they live in
// scalar registers).
struct SetArgs {
  sadmc_walker_state *ws;
  double *pos;       // [W][3][npad]
  uint64_t *spins;   // [W]
  HotBin *hot;       // [W][cap]
  ColdBin *cold;     // [W][cap]
  uint64_t *wlh;     // [W][cap] WL histogram, or null
  PExtra *extra;     // [W][cap] (WCA pressure) or null
  int32_t *flag;     // set to 1 when any walker of the set stops with a status
  int64_t cap;
  uint64_t zone_n;     // Uniform::new(0, n) acceptance zone
  uint64_t zone_site;  // gen_range(0, N) acceptance zone (Ising)
  int32_t W, n, npad, G, K;
  int32_t method, move_plan, has_min_e, has_max_e, has_min_gamma;
  int32_t refresh, collect_pressure, precision, move_kind;
  double method_param, move_param, min_e, max_e, rc2, width, inv_width;
  const uint16_t *mlp;  // learned proposal: 1377 BF16 weights, or null
};

// Pending change of the single-walker plan/confirm entry points (for a
// learned LJ move, `dabse` carries ln H).
struct Pending {
  int32_t some, which;
  double tx, ty, tz, e, dabse;
};

// Cross-lane move of a double with a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  uint64_t b = __builtin_bit_cast(uint64_t, v);
  int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over the G lanes of a group, every lane ending with the same value.
// Steps 1-4 are DPP: quad_perm [1,0,3,2] (xor 1), quad_perm [2,3,0,1]
// (xor 2), row_half_mirror and row_mirror.  After steps 1-2 the four lanes of
// a quad hold one value, after step 3 the eight lanes of a half row do, so
// the mirrors pair each lane with a lane holding exactly the partial sum the
// xor-4 / xor-8 partner would hold: the result is bit-identical to an xor
// butterfly.  Wider groups finish with cross-row shuffles.
__device__ __forceinline__ double group_sum(double v, int G) {
  if (G >= 2) v = v + dpp_f64<0xB1>(v);
  if (G >= 4) v = v + dpp_f64<0x4E>(v);
  if (G >= 8) v = v + dpp_f64<0x141>(v);
  if (G >= 16) v = v + dpp_f64<0x140>(v);
  if (G >= 32) v = v + __shfl_xor(v, 16, 64);
  if (G >= 64) v = v + __shfl_xor(v, 32, 64);
  return v;
}

// rand 0.7 UniformInt sampling with a precomputed zone (lj.rs:368 /
// ising.rs:105 semantics; the zone is the only part that depends on n).
__device__ __forceinline__ uint64_t sample_zone(uint64_t &s0, uint64_t &s1, uint64_t range, uint64_t zone) {
  for (;;) {
    uint64_t v = rng_next(s0, s1);
    uint64_t lo = v * range;
    if (lo <= zone) return __umul64hi(v, range);
  }
}



// MV = 1: the learned Gaussian move with its Hastings term (LJ only).
// GC > 0: the lanes per walker fixed at compile time (must equal the set's
// G), which folds the group sums, lane indices and atom-slot arithmetic.
template <int SYS, int KMAX, int PREC, int METH, int MV = 0, int GC = 0>
struct Engine {
  const SetArgs *Ap;  // kernel argument segment (uniform; read with scalar loads)
  SADMC_DF const SetArgs &args() const { return *Ap; }
  int lane, G, K, n;
  HotBin *H;
  ColdBin *Cb;
  uint64_t *WH;
  PExtra *X;
  // system
  double *lx, *ly, *lz;
  // This lane's atoms in registers, or (kLdsPos) read from the LDS mirror:
  // the learned move, WCA and groups with many atoms per lane keep their
  // registers for the rest of the state.
  static constexpr bool kLdsPos = MV == 1 || SYS == SYS_WCA || KMAX >= 8;
  double x[KMAX], y[KMAX], z[KMAX];
  SADMC_DF double xk(int k) const { return kLdsPos ? lx[lane + G * k] : x[k]; }
  SADMC_DF double yk(int k) const { return kLdsPos ? ly[lane + G * k] : y[k]; }
  SADMC_DF double zk(int k) const { return kLdsPos ? lz[lane + G * k] : z[k]; }
  uint64_t spins;
  double R2, L, halfL;
  // EnergyMC scalars
  uint64_t s0, s1;
  double E, err;
  uint64_t moves, accepted;
  double tm;  // (double)moves, kept exact by +1.0 steps
  double acc_rate, tscale, usc;
  double bmin;
  int32_t base, len;
  double max_S;
  int32_t max_S_index;
  uint64_t epoch;
  int32_t bulk, status, method;
  // SAD
  double too_lo, too_hi, lam, min_T, tF_d, Ns_d, lnw_lo, lnw_hi;
  uint64_t tL, tF, Ns, highest, tmax;
  int32_t ilo, ihi;
  // SAMC / WL
  double t0, wl_gamma, wl_Ns, wl_min_e;
  uint64_t wl_lowest, wl_highest, wl_total;
  int32_t wl_len;
  int32_t has_pressure;
  // current bin (logical cur_i, physical cur_p): hot record (and WL count)
  int32_t cur_i, cur_p;
  double c_lnw;
  uint64_t c_hist, c_wlh;
  // cold record of bin kp, and the previous move's cold work on it
  int32_t kp;
  double k_etot, k_esq;
  uint64_t k_rt, k_vep;
  int32_t pend, pend_rt, pend_bulk;
  double pend_E;
  uint64_t pend_epoch;
  // periodic triggers
  uint64_t next_check, next_press;
  // pending move
  int32_t pw;
  double ptx, pty, ptz, pe, pdabse;
  // learned move: weights and this walker's exchange row in LDS, ln H
  const uint16_t *PW;
  float *XC;
  double lnh;
  // this move's SAD gamma, computed while the proposed bin's record loads
  // (valid unless the move changed lambda or N_s), and the next uniform with
  // the generator state after drawing it, computed before it is needed
  double gpre, spec_u;
  bool gdirty;
  uint64_t spec0, spec1;
  // sqrt(1 / moves) for the group's next G moves: lane l holds the value of
  // move rs_first + l; rs_next counts the moves served from it (G = empty)
  double rs_lane;
  int32_t rs_next;

  // ---------------------------------------------------------------- setup
  SADMC_DF void load_system(const SetArgs &a, int walker, int lane_in_group, double *lds_group) {
    Ap = &a;
    lane = lane_in_group;
    G = GC ? GC : a.G;
    K = a.K;
    n = a.n;
    H = a.hot + (int64_t)walker * a.cap;
    Cb = a.cold + (int64_t)walker * a.cap;
    WH = a.wlh ? a.wlh + (int64_t)walker * a.cap : nullptr;
    X = a.extra ? a.extra + (int64_t)walker * a.cap : nullptr;
    kp = -1;
    pend = 0;
    if (SYS == SYS_ISING) {
      spins = a.spins[walker];
    } else {
      lx = lds_group;
      ly = lds_group + a.npad;
      lz = lds_group + 2 * a.npad;
      const double *P = a.pos + (int64_t)walker * 3 * a.npad;
#pragma unroll
      for (int k = 0; k < KMAX; k++) {
        if (k < K) {
          int j = lane + G * k;
          const double vx = P[j], vy = P[a.npad + j], vz = P[2 * a.npad + j];
          if (!kLdsPos) {
            x[k] = vx;
            y[k] = vy;
            z[k] = vz;
          }
          lx[j] = vx;
          ly[j] = vy;
          lz[j] = vz;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }

  SADMC_DF void load(const SetArgs &a, int walker, int lane_in_group, double *lds_group) {
    load_system(a, walker, lane_in_group, lds_group);
    const sadmc_walker_state &w = a.ws[walker];
    s0 = w.rng[0];
    s1 = w.rng[1];
    E = w.E;
    err = w.error;
    moves = w.moves;
    tm = (double)moves;
    rs_next = 1 << 30;
    accepted = w.accepted_moves;
    acc_rate = w.acceptance_rate;
    tscale = w.translation_scale;
    bmin = w.bins_min;
    base = (int32_t)w.base;
    len = (int32_t)w.len;
    max_S = w.max_S;
    max_S_index = (int32_t)w.max_S_index;
    epoch = w.epoch;
    bulk = (int32_t)w.bulk_visited;
    status = w.status;
    method = w.method;
    if (METH == M_SAD) {
      too_lo = w.too_lo;
      too_hi = w.too_hi;
      lam = w.latest_parameter;
      min_T = w.min_T;
      tL = w.tL;
      tF = w.tF;
      Ns = w.num_states;
      highest = w.highest_hist;
      tmax = w.tmax;
    }
    if (METH == M_SAMC || METH == M_WL) t0 = w.t0;
    if (METH == M_WL) {
      wl_gamma = w.wl_gamma;
      wl_Ns = w.wl_num_states;
      wl_min_e = w.wl_min_energy;
      wl_lowest = w.wl_lowest_hist;
      wl_highest = w.wl_highest_hist;
      wl_total = w.wl_total_hist;
      wl_len = (int32_t)w.wl_len;
    }
    if (SYS == SYS_WCA) has_pressure = w.has_pressure;
    if (SYS != SYS_ISING) {
      if (SYS == SYS_LJ) R2 = w.geom * w.geom;
      L = w.geom;
      halfL = 0.5 * w.geom;
    }
    tF_d = (double)tF;
    Ns_d = (double)Ns;
    usc = SYS == SYS_ISING ? 0.0 : uniform_scale(-tscale, tscale);
    cur_p = -1;
    if (!status) {
      refresh_indices();
      reset_checks();
    }
  }

  // Write back the walker state; `all` also writes the fields a method
  // family does not use (only from_params needs that).
  SADMC_DF void store(int walker, bool all = false) {
    apply_pending();
    write_cold();
    if (lane == 0) {
      sadmc_walker_state &w = args().ws[walker];
      w.rng[0] = s0;
      w.rng[1] = s1;
      w.E = E;
      w.error = err;
      w.moves = moves;
      w.accepted_moves = accepted;
      w.acceptance_rate = acc_rate;
      w.translation_scale = tscale;
      w.bins_min = bmin;
      w.bins_width = args().width;
      w.base = base;
      w.len = len;
      w.max_S = max_S;
      w.max_S_index = max_S_index;
      w.epoch = epoch;
      w.bulk_visited = (uint64_t)bulk;
      w.status = status;
      w.method = method;
      if (all || METH == M_SAD) {
        w.too_lo = too_lo;
        w.too_hi = too_hi;
        w.latest_parameter = lam;
        w.min_T = min_T;
        w.tL = tL;
        w.tF = tF;
        w.num_states = Ns;
        w.highest_hist = highest;
        w.tmax = tmax;
      }
      if (all || METH == M_SAMC || METH == M_WL) w.t0 = t0;
      if (all || METH == M_WL) {
        w.wl_gamma = wl_gamma;
        w.wl_num_states = wl_Ns;
        w.wl_min_energy = wl_min_e;
        w.wl_lowest_hist = wl_lowest;
        w.wl_highest_hist = wl_highest;
        w.wl_total_hist = wl_total;
        w.wl_len = wl_len;
      }
      if (all || SYS == SYS_WCA) w.has_pressure = has_pressure;
      if (status) *args().flag = 1;
    }
    if (SYS == SYS_ISING) {
      args().spins[walker] = spins;
    } else {
      double *P = args().pos + (int64_t)walker * 3 * args().npad;
#pragma unroll
      for (int k = 0; k < KMAX; k++) {
        if (k < K) {
          int j = lane + G * k;
          P[j] = xk(k);
          P[args().npad + j] = yk(k);
          P[2 * args().npad + j] = zk(k);
        }
      }
    }
  }

  // --------------------------------------------------------------- bins
  // Bins::state_to_index (energy.rs:371-373); `as usize` saturates.
  // The quotient by the width is first formed with the reciprocal; only
  // when that lands within 1e-9 of an integer (where the two roundings could
  // truncate differently) is the division itself computed.
  SADMC_DF int32_t e_index(double e) const {
    const double d = e - bmin;
    double v = d * args().inv_width;
    const double f = v - trunc(v);
    if (!(f > 1e-9 && f < 1.0 - 1e-9 && v < 1.0e6)) v = d / args().width;
    if (!(v > 0.0)) return 0;
    if (v >= 2.0e9) return 0x7fffffff;
    return (int32_t)v;
  }
  SADMC_DF double i_energy(int32_t i) const { return bmin + ((double)i + 0.5) * args().width; }
  SADMC_DF bool idx_ok(int32_t i) {
    if (i < 0 || i >= len) {
      status = SADMC_WS_INDEX;
      return false;
    }
    return true;
  }
  SADMC_DF double lnw_at(int32_t p) const { return p == cur_p ? c_lnw : H[p].lnw; }
  SADMC_DF uint64_t hist_at(int32_t p) const { return p == cur_p ? c_hist : H[p].hist; }
  SADMC_DF uint64_t wlh_at(int32_t p) const { return p == cur_p ? c_wlh : WH[p]; }

  // Cold record of the current bin: written back when the walker leaves the
  // bin, loaded for the new bin, and first used one move later.
  SADMC_DF void write_cold() {
    if (kp >= 0) {
      Cb[kp].etot = k_etot;
      Cb[kp].esq = k_esq;
      Cb[kp].rt = k_rt;
      Cb[kp].vep = k_vep;
    }
  }
  SADMC_DF void switch_cold(int32_t p) {
    write_cold();
    const ColdBin &q = Cb[p];
    k_etot = q.etot;
    k_esq = q.esq;
    k_rt = q.rt;
    k_vep = q.vep;
    kp = p;
  }
  // energy.rs:941-943 and 950-965 for the previous move: the energy sums and,
  // when the round-trip branch reached have_visited, its test and update.
  SADMC_DF void apply_pending() {
    if (pend) {
      k_etot += pend_E;
      k_esq += pend_E * pend_E;
      if (pend_rt && !(pend_bulk || k_vep == pend_epoch)) {
        k_vep = pend_epoch;
        k_rt += 1;
      }
      pend = 0;
    }
  }
  SADMC_DF void set_current(int32_t p, double lnw, uint64_t hist) {
    c_lnw = lnw;
    c_hist = hist;
    if (METH == M_WL) c_wlh = WH[p];
    cur_p = p;
    switch_cold(p);
  }

  // Re-derive every cached index from its inputs (after init, resume or a
  // window change) and reload the values cached at those indices.  An index
  // past the bins is left for move_once to report where the reference would
  // index out of bounds.
  SADMC_DF void refresh_indices() {
    int32_t i = e_index(E);
    cur_i = i;
    if (i < len) {
      int32_t p = base + i;
      if (p != cur_p) {
        HotBin hb = H[p];
        set_current(p, hb.lnw, hb.hist);
      }
    } else {
      cur_p = -1;
    }
    if (METH == M_SAD) {
      ilo = e_index(too_lo);
      ihi = e_index(too_hi);
      lnw_lo = ilo < len ? lnw_at(base + ilo) : 0.0;
      lnw_hi = ihi < len ? lnw_at(base + ihi) : 0.0;
    }
  }

  // True when every cached index is inside the bins.
  SADMC_DF bool indices_valid() const {
    if (cur_p < 0) return false;
    if (METH == M_SAD && (ilo >= len || ihi >= len)) return false;
    return true;
  }

  SADMC_DF void reset_checks() {
    if (SYS != SYS_ISING) {
      uint64_t P = args().refresh > 0 ? (uint64_t)args().refresh : (uint64_t)len * (uint64_t)len * 1000ULL;
      next_check = (moves / P + 1) * P;
      if (SYS == SYS_WCA) {
        uint64_t Q = (uint64_t)n * (uint64_t)n;
        next_press = (moves / Q + 1) * Q;
      }
    }
  }

  // energy.rs:400-434 (prepare_for_state) for every energy in [lo, hi]: the
  // bins grow to both ends.  New slots get have_visited = true (visit_epoch =
  // epoch) and round_trips = 1; every other field is already zero.
  SADMC_DF bool prepare_for_range(double lo, double hi) {
    bool front = false, grew = false;
    while (lo < bmin) {
      if (base == 0) {
        status = SADMC_WS_CAPACITY;
        return false;
      }
      base -= 1;
      len += 1;
      Cb[base].rt = 1;
      Cb[base].vep = epoch;
      bmin -= args().width;
      front = true;
    }
    while (hi >= bmin + args().width * (double)len) {
      int32_t q = base + len;
      if (q >= args().cap) {
        status = SADMC_WS_CAPACITY;
        return false;
      }
      Cb[q].rt = 1;
      Cb[q].vep = epoch;
      len += 1;
      grew = true;
    }
    if (front || (grew && !indices_valid())) refresh_indices();
    if ((front || grew) && SYS != SYS_ISING && args().refresh <= 0) {
      uint64_t P = (uint64_t)len * (uint64_t)len * 1000ULL;
      next_check = (moves / P + 1) * P;
    }
    return status == 0;
  }

  // ------------------------------------------------------------ systems
  SADMC_DF int spin(int b) const { return ((spins >> b) & 1ULL) ? 1 : -1; }

  // Full FP64 energy in lane-group order: lane l sums pairs (i, j<i) for its
  // atoms i; the G partial sums meet in the butterfly.
  SADMC_DF double full_energy() {
    if (SYS == SYS_ISING) {
      int N = n;
      double e = 0.0;
      for (int i1 = 0; i1 < N; i1++)
        for (int j1 = 0; j1 < N; j1++) {
          int j2 = j1 + 1 == N ? 0 : j1 + 1;
          int i2 = i1 + 1 == N ? 0 : i1 + 1;
          int nt = spin(i1 + j2 * N) + spin(i2 + j1 * N);
          e += (double)nt * (double)spin(i1 + j1 * N);
        }
      return e;
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        int i = lane + G * k;
        if (i < n) {
          for (int j = 0; j < i; j++) {
            if (SYS == SYS_LJ) {
              acc += lj_pair(norm2(xk(k) - lx[j], yk(k) - ly[j], zk(k) - lz[j]));
            } else {
              acc += wca_pair(norm2(min_image(xk(k) - lx[j], L, halfL), min_image(yk(k) - ly[j], L, halfL),
                                    min_image(zk(k) - lz[j], L, halfL)),
                              args().rc2);
            }
          }
        }
      }
    }
    return group_sum(acc, G);
  }

  // wca.rs:202-218 virial, ordered pairs, lane-group order.
  SADMC_DF double pressure() {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        int i = lane + G * k;
        if (i < n) {
          for (int j = 0; j < n; j++) {
            if (j == i) continue;
            acc += wca_pair_pressure(norm2(min_image(xk(k) - lx[j], L, halfL), min_image(yk(k) - ly[j], L, halfL),
                                           min_image(zk(k) - lz[j], L, halfL)),
                                     args().rc2);
          }
        }
      }
    }
    double p = group_sum(acc, G);
    return p / (3.0 * (L * L * L));
  }

  // ising.rs:104-119.
  SADMC_DF bool ising_plan(double &e2) {
    const int N = n;
    int i = (int)sample_zone(s0, s1, (uint64_t)N, args().zone_site);
    int j = (int)sample_zone(s0, s1, (uint64_t)N, args().zone_site);
    int jp = j + 1 == N ? 0 : j + 1;
    int jm = j == 0 ? N - 1 : j - 1;
    int ip = i + 1 == N ? 0 : i + 1;
    int im = i == 0 ? N - 1 : i - 1;
    int nt = spin(i + jp * N) + spin(i + jm * N) + spin(ip + j * N) + spin(im + j * N);
    double e = E - (double)nt * (double)spin(i + j * N) * 2.0;
    pw = i + j * N;
    pe = e;
    e2 = e;
    return true;
  }

  // lj.rs:365-374 / wca.rs:341-353 with the cube move (half-width d,
  // Uniform::new(-d, d) scale sc); lj.rs:86-105 and wca.rs:119-140 for the
  // energy difference.
  SADMC_DF bool particle_plan(double d, double sc, double &e2) {
    int w = (int)sample_zone(s0, s1, (uint64_t)n, args().zone_n);
    double ux = uniform_sample(s0, s1, -d, sc);
    double uy = uniform_sample(s0, s1, -d, sc);
    double uz = uniform_sample(s0, s1, -d, sc);
    double fx = lx[w], fy = ly[w], fz = lz[w];
    double tx, ty, tz;
    if (SYS == SYS_LJ) {
      tx = fx + ux;
      ty = fy + uy;
      tz = fz + uz;
      double r2to = norm2(tx, ty, tz), r2from = norm2(fx, fy, fz);
      if (r2to > R2 && r2to > r2from) return false;
    } else {
      tx = put_in_cell(fx + ux, L);
      ty = put_in_cell(fy + uy, L);
      tz = put_in_cell(fz + uz, L);
    }
    // Every slot computes its pair terms; the moved atom's own slot and the
    // padding slots (whose terms may be inf or NaN) are dropped by a select
    // of the accumulator, which keeps the sums those of the skipped form
    // without a divergent branch per atom.
    double acc = 0.0, accabs = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        const int j = lane + G * k;
        const bool use = j < n && j != w;
        if (SYS == SYS_LJ) {
          if (PREC == 0) {
            const double t = lj_pair(norm2(xk(k) - tx, yk(k) - ty, zk(k) - tz)) -
                             lj_pair(norm2(xk(k) - fx, yk(k) - fy, zk(k) - fz));
            const double na = acc + t;
            acc = use ? na : acc;
          }
        } else {
          double vn = wca_pair(norm2(min_image(xk(k) - tx, L, halfL), min_image(yk(k) - ty, L, halfL),
                                     min_image(zk(k) - tz, L, halfL)),
                               args().rc2);
          double vo = wca_pair(norm2(min_image(xk(k) - fx, L, halfL), min_image(yk(k) - fy, L, halfL),
                                     min_image(zk(k) - fz, L, halfL)),
                               args().rc2);
          const double na = acc + (vn - vo), nb = accabs + (vn + vo);
          acc = use ? na : acc;
          accabs = use ? nb : accabs;
        }
      }
    }
    if (SYS == SYS_LJ && PREC == 1) acc = lj32_lane_diff(w, tx, ty, tz, fx, fy, fz);
    double e = E + group_sum(acc, G);
    if (SYS == SYS_WCA) pdabse = group_sum(accabs, G);
    pw = w;
    ptx = tx;
    pty = ty;
    ptz = tz;
    pe = e;
    e2 = e;
    return true;
  }

  // Features of the group's atoms around (px, py, pz), skipping atom w;
  // positions are read from the LDS mirror to keep registers free.
  SADMC_DF void features(double px, double py, double pz, int w, float *f) {
#pragma unroll
    for (int c = 0; c < kPropNF; c++) f[c] = 0.0f;
#pragma unroll 1
    for (int k = 0; k < K; k++) {
      int j = lane + G * k;
      if (j < n && j != w) {
        float a = (float)(lx[j] - px), b = (float)(ly[j] - py), c = (float)(lz[j] - pz);
        feature_add(a * a + b * b + c * c, f);
      }
    }
#pragma unroll
    for (int c = 0; c < kPropNF; c++) f[c] = group_sum_f32(f[c], G);
  }

  // This lane's sum of the binary32 pair-term differences V(|x_j - t|^2) -
  // V(|x_j - f|^2) over its atoms j != w (lj.rs:86-105 in FP32 pair terms,
  // accumulated in FP64), 1/r^2 from recip_f32_fast when FAST, else from the
  // division.
  template <bool FAST>
  SADMC_DF double lj32_lane_sum(int w, double tx, double ty, double tz, double fx, double fy, double fz) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        const int j = lane + G * k;
        const bool use = j < n && j != w;
        const float an = (float)(xk(k) - tx), bn = (float)(yk(k) - ty), cn = (float)(zk(k) - tz);
        const float ao = (float)(xk(k) - fx), bo = (float)(yk(k) - fy), co = (float)(zk(k) - fz);
        const float r2n = an * an + bn * bn + cn * cn;
        const float r2o = ao * ao + bo * bo + co * co;
        const double t = FAST ? (double)lj_pair_f32_fast(r2n) - (double)lj_pair_f32_fast(r2o)
                              : (double)lj_pair_f32(r2n) - (double)lj_pair_f32(r2o);
        const double na = acc + t;
        acc = use ? na : acc;
      }
    }
    return acc;
  }

  // lj32_lane_sum with the fast reciprocal, equal to the division's sum bit
  // for bit: the two reciprocals agree on [2^-126, 2^126); below 2^-126
  // 1/r^2 exceeds 2^126, so s^6 overflows and the pair term is NaN with
  // either; from 2^126 up s^2 underflows and both terms are +0; at +inf
  // only the fast one is NaN.  So the fast sum differs only where it is
  // NaN, and then the lanes still running this code recompute with the
  // division.
  SADMC_DF double lj32_lane_diff(int w, double tx, double ty, double tz, double fx, double fy, double fz) {
    double acc = lj32_lane_sum<true>(w, tx, ty, tz, fx, fy, fz);
    if (__builtin_expect(__any(acc != acc), 0)) acc = lj32_lane_sum<false>(w, tx, ty, tz, fx, fy, fz);
    return acc;
  }

  // Sum over the group of the pair-term differences for atom w moving from
  // f to t (lj.rs:86-105).
  SADMC_DF double lj_delta(int w, double tx, double ty, double tz, double fx, double fy, double fz) {
    if (PREC == 1) return group_sum(lj32_lane_diff(w, tx, ty, tz, fx, fy, fz), G);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        int j = lane + G * k;
        if (j < n && j != w)
          acc += lj_pair(norm2(xk(k) - tx, yk(k) - ty, zk(k) - tz)) -
                 lj_pair(norm2(xk(k) - fx, yk(k) - fy, zk(k) - fz));
      }
    }
    return group_sum(acc, G);
  }

  // The learned move: sigma from the network at the old position, a
  // Gaussian step, sigma' at the new position for the Hastings term.
  SADMC_DF bool learned_plan(double &e2) {
    int w = (int)sample_zone(s0, s1, (uint64_t)n, args().zone_n);
    const double fx = lx[w], fy = ly[w], fz = lz[w];
    float f[kPropNF], xi[3];
    features(fx, fy, fz, w, f);
    const float s = mlp_eval(PW, XC, f, lane, G);
    gaussian3(s0, s1, xi);
    const float sig = expf(s);
    const double tx = fx + (double)(sig * xi[0]), ty = fy + (double)(sig * xi[1]), tz = fz + (double)(sig * xi[2]);
    const double r2to = norm2(tx, ty, tz), r2from = norm2(fx, fy, fz);
    // the wall's verdict waits until after the second evaluation, so that
    // the whole wave reaches it together (its result is then unused)
    const bool outside = r2to > R2 && r2to > r2from;
    features(tx, ty, tz, w, f);
    const float s2 = mlp_eval(PW, XC, f, lane, G);
    if (outside) return false;
    const double e = E + lj_delta(w, tx, ty, tz, fx, fy, fz);
    lnh = prop_lnh(s, s2, xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
    pw = w;
    ptx = tx;
    pty = ty;
    ptz = tz;
    pe = e;
    pdabse = lnh;
    e2 = e;
    return true;
  }

  // SAD ln w(e) with the clamped extension (energy.rs:445-463); false if e
  // has no bin.
  SADMC_DF bool sad_lnw(double e, double &out) {
    if (ilo < 0 || ilo >= len || ihi < 0 || ihi >= len) return false;
    if (e < too_lo) {
      out = lnw_lo + (e - too_lo) / min_T;
      return true;
    }
    if (e > too_hi) {
      out = lnw_hi;
      return true;
    }
    const int32_t i = e_index(e);
    if (i < 0 || i >= len) return false;
    out = lnw_at(base + i);
    return true;
  }

  // `count` training records: the features of a random atom and, per scale
  // of the grid, the mean of alpha |d|^2 over kPropR Gaussian steps, from the
  // probe stream seeded by `seed`.  Lane 0 writes out[count][24].
  SADMC_DF void probe(uint64_t seed, int count, float *out) {
    uint64_t q0, q1;
    rng_seed(seed, q0, q1);
    double lnw1 = 0.0;
    const bool have1 = status == 0 && sad_lnw(E, lnw1);
    for (int c = 0; c < count; c++) {
      float rec[kPropRec];
      const int w = (int)uniform_index(q0, q1, (uint64_t)n);
      const double fx = lx[w], fy = ly[w], fz = lz[w];
      features(fx, fy, fz, w, rec);
#pragma unroll
      for (int k = 0; k < kPropNS; k++) {
        const float sig = expf(kPropSMin + (float)k * ((kPropSMax - kPropSMin) / (float)(kPropNS - 1)));
        float jsum = 0.0f;
        for (int r = 0; r < kPropR; r++) {
          float xi[3];
          gaussian3(q0, q1, xi);
          const float dx = sig * xi[0], dy = sig * xi[1], dz = sig * xi[2];
          const double tx = fx + (double)dx, ty = fy + (double)dy, tz = fz + (double)dz;
          const double r2to = norm2(tx, ty, tz), r2from = norm2(fx, fy, fz);
          if (r2to > R2 && r2to > r2from) continue;
          const double e2 = E + lj_delta(w, tx, ty, tz, fx, fy, fz);
          if (args().has_max_e && e2 > args().max_e && e2 > E) continue;
          double lnw2;
          if (!have1 || !sad_lnw(e2, lnw2)) continue;
          const double lr = lnw1 - lnw2;
          const float al = lr >= 0.0 ? 1.0f : (float)exp(lr);
          jsum += al * (dx * dx + dy * dy + dz * dz);
        }
        rec[kPropNF + k] = jsum / (float)kPropR;
      }
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < kPropRec; i++) out[(int64_t)c * kPropRec + i] = rec[i];
      }
    }
  }

  SADMC_DF bool plan_move(double d, double sc, double &e2) {
    if (SYS == SYS_ISING) return ising_plan(e2);
    if (MV == 1 && SYS == SYS_LJ) return learned_plan(e2);
    return particle_plan(d, sc, e2);
  }

  // The initial descent of from_params always uses the cube move.
  SADMC_DF bool cube_plan(double d, double sc, double &e2) {
    if (SYS == SYS_ISING) return ising_plan(e2);
    return particle_plan(d, sc, e2);
  }

  // lj.rs:106-123.
  SADMC_DF void lj_set_energy(double new_e) {
    double nn = (double)n;
    double new_error = fabs(new_e) > fabs(E) ? fabs(new_e) * 1e-15 * nn : fabs(E) * 1e-15 * nn;
    err = new_error + err;
    if (err > fabs(new_e) * 1e-14 * nn * nn) {
      err *= 0.0;
      E = full_energy();
    } else {
      E = new_e;
    }
  }

  // wca.rs:164-180.
  SADMC_DF void wca_set_energy(double new_e, double dabse) {
    double nn = (double)n;
    double single = dabse > fabs(new_e) ? 1e-14 * dabse * nn : 1e-14 * fabs(new_e) * nn;
    err += single * nn;
    if (err > fabs(new_e) * 1e-13 * nn * nn) {
      E = full_energy();
      err = 1e-15 * E * nn;
    } else {
      E = new_e;
    }
  }

  // ConfirmSystem::confirm: ising.rs:94-101, lj.rs:338-358, wca.rs:276-301.
  SADMC_DF void confirm() {
    if (SYS == SYS_ISING) {
      spins ^= (1ULL << pw);
      E = pe;
      return;
    }
    int owner = pw & (G - 1);
    int slot = pw / G;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (!kLdsPos && lane == owner && k == slot) {
        x[k] = ptx;
        y[k] = pty;
        z[k] = ptz;
      }
    }
    if (lane == owner) {
      lx[pw] = ptx;
      ly[pw] = pty;
      lz[pw] = ptz;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (args().refresh > 0) {
      E = pe;
    } else if (SYS == SYS_LJ) {
      lj_set_energy(pe);
    } else {
      wca_set_energy(pe, pdabse);
    }
  }

  // lj.rs:249-261, wca.rs:237-250.
  SADMC_DF void verify_energy() {
    double egood = full_energy();
    double nn = (double)n;
    double expected = SYS == SYS_LJ ? fabs(E) * 1e-14 * nn * nn : fabs(E) * 1e-13 * nn * nn;
    if (fabs(egood - E) > expected) status = SADMC_WS_VERIFY;
  }

  // ------------------------------------------------------------- method
  // energy.rs:22-38 with t = moves, tF, N_s (energy.rs:799-813).
  SADMC_DF double sad_gamma() const {
    if (lam * tF_d * Ns_d == 0.0) return 0.0;
    return (lam + tm / tF_d) / (lam + tm / Ns_d * (tm / tF_d));
  }

  // Draw the next uniform ahead of the test that may use it.
  SADMC_DF void spec_draw() {
    spec0 = s0;
    spec1 = s1;
    spec_u = gen_f64(spec0, spec1);
  }
  // `need && u > exp(x)`, consuming the uniform only when `need`.
  // The comparison is settled from a single-precision 2^(x log2 e) whenever
  // u lies outside that estimate's error band; only inside the band is the
  // double exp computed.  Band: t = x log2 e (double, relative error 2^-52)
  // rounds to float with |dt| <= |t| 2^-24, v_exp_f32 is within a few ulp
  // (2^-21), so the estimate is within |t| 2^-24 ln 2 + 2^-21 < 1e-7 |t| +
  // 1e-6 of exp(x), and of the correctly rounded double exp (2^-53 more);
  // the band is ten times that.  Below t = -125 exp(x) < 2^-125 lies under
  // every nonzero u (multiples of 2^-53), so the test is u > 0.  NaN and
  // x >= 0 take the double path.
  SADMC_DF bool mh_test(bool need, double x) {
    if (!need) return false;
    s0 = spec0;
    s1 = spec1;
    const double t = x * 1.4426950408889634;
    if (t < -125.0) return spec_u > 0.0;
    if (t <= 0.0) {
      const double est = (double)__builtin_amdgcn_exp2f((float)t);
      const double band = est * (1e-5 - 1e-6 * t);
      if (spec_u > est + band) return true;
      if (spec_u < est - band) return false;
    }
    return spec_u > exp(x);
  }
  // Metropolis-Hastings test on the log acceptance ratio.
  SADMC_DF bool reject_lr(double lr) { return mh_test(lr < 0.0, lr); }
  // ln w(e) with the SAD clamped extension (energy.rs:445-463); `inside` is
  // the bin's ln w.
  SADMC_DF double sad_lnw_ext(double e, double inside) const {
    if (e < too_lo) return lnw_lo + (e - too_lo) / min_T;
    if (e > too_hi) return lnw_hi;
    return inside;
  }

  // energy.rs:440-512.  `r2` is the record of the proposed bin.  The
  // learned move adds ln H to the exponent (the tests are the reference's
  // when ln H = 0).
  SADMC_DF bool reject_move(double e1, double e2, double r2_lnw, uint64_t r2_hist) {
    if (MV == 1) {
      if (METH == M_SAD) {
        const double lnw1 = sad_lnw_ext(e1, c_lnw), lnw2 = sad_lnw_ext(e2, r2_lnw);
        bool rejected = reject_lr(lnw1 - lnw2 + lnh);
        if (!rejected && r2_hist == 0 && e2 < too_hi && e2 > too_lo) {
          lam = (too_hi - too_lo) / min_T;
          Ns += 1;
          Ns_d = (double)Ns;
          tL = moves;
          gdirty = true;
        }
        return rejected;
      } else if (METH == M_SAMC) {
        return reject_lr(c_lnw - r2_lnw + lnh);
      } else if (METH == M_WL) {
        bool rejected = reject_lr(c_lnw - r2_lnw + lnh);
        if (method != SADMC_SAMC && !rejected && r2_hist == 0 && wl_lowest > 0) wl_Ns += 1.0;
        return rejected;
      } else {
        return reject_lr((e1 - e2) / args().method_param + lnh);
      }
    }
    if (METH == M_SAD) {
      const double lnw1 = sad_lnw_ext(e1, c_lnw), lnw2 = sad_lnw_ext(e2, r2_lnw);
      bool rejected = mh_test(lnw2 > lnw1, lnw1 - lnw2);
      if (!rejected && r2_hist == 0 && e2 < too_hi && e2 > too_lo) {
        lam = (too_hi - too_lo) / min_T;
        Ns += 1;
        Ns_d = (double)Ns;
        tL = moves;
        gdirty = true;
      }
      return rejected;
    } else if (METH == M_SAMC) {
      return mh_test(r2_lnw > c_lnw, c_lnw - r2_lnw);
    } else if (METH == M_WL) {
      bool rejected = mh_test(r2_lnw > c_lnw, c_lnw - r2_lnw);
      if (method != SADMC_SAMC && !rejected && r2_hist == 0 && wl_lowest > 0) wl_Ns += 1.0;
      return rejected;
    } else {
      return mh_test(e1 < e2, (e1 - e2) / args().method_param);
    }
  }

  // Running max of t_found over logical bins [a, b].
  // (The scans over bins that SAD's range changes need are split over the
  // walker's G lanes: each lane takes every G-th bin, so the loads of one
  // scan are in flight together; integer sums and maxima are then combined
  // across the group, which leaves every result as the sequential one.)
  SADMC_DF uint64_t group_max_u64(uint64_t v) const {
    for (int d = 1; d < G; d <<= 1) {
      const uint64_t o = __shfl_xor(v, d, 64);
      v = o > v ? o : v;
    }
    return v;
  }
  SADMC_DF uint32_t group_sum_u32(uint32_t v) const {
    for (int d = 1; d < G; d <<= 1) v += __shfl_xor(v, d, 64);
    return v;
  }
  SADMC_DF void tmax_scan(int32_t a, int32_t b) {
    uint64_t m = 0;
    for (int32_t j = a + lane; j <= b; j += G) {
      const uint64_t t = Cb[base + j].tfound;
      m = t > m ? t : m;
    }
    m = group_max_u64(m);
    if (m > tmax) tmax = m;
  }

  SADMC_DF void set_lnw(int32_t p, double v) {
    H[p].lnw = v;
    if (p == cur_p) c_lnw = v;
  }

  // energy.rs:514-636 for the current bin (logical i = cur_i).
  SADMC_DF void update_sad(double E_, int32_t i) {
    const double E = E_;
    if (!(too_lo > too_hi || E < too_lo || E > too_hi)) {
      c_lnw += gdirty ? sad_gamma() : gpre;
      if (cur_p == base + ilo) lnw_lo = c_lnw;
      if (cur_p == base + ihi) lnw_hi = c_lnw;
    }
    if (c_hist > highest) {
      highest = c_hist;
      // the cached ln w at too_lo / too_hi and N_s change here only when
      // the range grows; otherwise ln w at those bins changes only through
      // the current bin, which the increment above already tracks
      bool grew = false;
      if (E > too_hi) {
        grew = true;
        if (ihi >= len) {
          status = SADMC_WS_INDEX;
          return;
        }
        const double vhi = lnw_hi;
        int32_t jlo = ihi > 0 ? ihi - 1 : 0, jhi = i + 1 < len ? i + 1 : len - 1;
        uint32_t cnt = 0;
        for (int32_t j = jlo + lane; j <= jhi; j += G) {
          const double ej = i_energy(j);
          const int32_t pj = base + j;
          if (ej > too_hi && ej <= E && pj != cur_p) {
            const bool nz = H[pj].hist != 0;
            H[pj].lnw = nz ? vhi : 0.0;
            cnt += nz ? 1u : 0u;
          }
        }
        cnt = group_sum_u32(cnt);
        if (i >= jlo && i <= jhi) {  // the current bin, whose record every lane holds
          const double ec = i_energy(i);
          if (ec > too_hi && ec <= E) {
            c_lnw = c_hist != 0 ? vhi : 0.0;
            cnt += c_hist != 0 ? 1u : 0u;
          }
        }
        Ns += cnt;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        lam = (E - too_lo) / min_T;
        tL = moves;
        double nh = i_energy(i);
        int32_t ihi2 = e_index(nh);
        if (ihi2 > ihi) tmax_scan(ihi + 1, ihi2 < len ? ihi2 : len - 1);
        too_hi = nh;
        ihi = ihi2;
        if (ihi >= len) {
          status = SADMC_WS_INDEX;
          return;
        }
      } else if (E < too_lo) {
        grew = true;
        if (ilo >= len) {
          status = SADMC_WS_INDEX;
          return;
        }
        // every bin below too_lo extends from the old ln w(too_lo): the bin
        // of too_lo itself comes last in energy order, and ln w(too_lo) is
        // reloaded from the new range below
        int32_t jlo = i > 0 ? i - 1 : 0, jhi = ilo + 1 < len ? ilo + 1 : len - 1;
        const double vlo = lnw_lo;
        uint32_t cnt = 0;
        for (int32_t j = jlo + lane; j <= jhi; j += G) {
          const double ej = i_energy(j);
          const int32_t pj = base + j;
          if (ej < too_lo && ej >= E && pj != cur_p) {
            const bool nz = H[pj].hist != 0;
            double v = 0.0;
            if (nz) {
              v = vlo + (ej - too_lo) / min_T;
              if (v < 0.0) v = 0.0;
            }
            H[pj].lnw = v;
            cnt += nz ? 1u : 0u;
          }
        }
        cnt = group_sum_u32(cnt);
        if (i >= jlo && i <= jhi) {  // the current bin, whose record every lane holds
          const double ec = i_energy(i);
          if (ec < too_lo && ec >= E) {
            double v = 0.0;
            if (c_hist != 0) {
              v = vlo + (ec - too_lo) / min_T;
              if (v < 0.0) v = 0.0;
            }
            c_lnw = v;
            cnt += c_hist != 0 ? 1u : 0u;
          }
        }
        Ns += cnt;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        lam = (too_hi - E) / min_T;
        tL = moves;
        double nl = i_energy(i);
        int32_t ilo2 = e_index(nl);
        if (ilo2 < ilo) tmax_scan(ilo2, ilo - 1);
        too_lo = nl;
        ilo = ilo2;
      }
      if (grew) {
        Ns_d = (double)Ns;
        lnw_lo = lnw_at(base + ilo);
        lnw_hi = lnw_at(base + ihi);
      }
    }
    if (tL == moves) {
      if (ilo >= len || ihi >= len) {
        status = SADMC_WS_INDEX;
        return;
      }
      uint64_t old_tF = tF;
      tF = tmax;
      tF_d = (double)tF;
      if (old_tF != tF && args().move_plan == SADMC_ACCEPTANCE_RATE) {
        double s = acc_rate / args().move_param;
        s = s < 0.8 ? 0.8 : (s > 1.2 ? 1.2 : s);
        tscale *= s;
        if (SYS != SYS_ISING) usc = uniform_scale(-tscale, tscale);
      }
    }
  }

  // energy.rs:638-756 (WL and 1/t-WL); SAMC after the 1/t switch.
  SADMC_DF void update_wl(int32_t i) {
    if (method == SADMC_SAMC) {
      double t = tm;
      c_lnw += t > t0 ? t0 / t : 1.0;
      return;
    }
    c_lnw += wl_gamma;
    const bool inv_t = method == SADMC_INV_T_WL;
    const bool has_mg = !inv_t && args().has_min_gamma;
    if (has_mg && wl_gamma < args().method_param) {
      if (i >= wl_len) {
        status = SADMC_WS_INDEX;
        return;
      }
      c_wlh += 1;
      WH[cur_p] = c_wlh;
      return;
    }
    if (wl_len != len) {
      if (wl_len == 0 || (wl_gamma != 1.0 && wl_lowest > 0)) {
        wl_gamma = 1.0;
        wl_lowest = 0;
        wl_highest = 0;
        wl_total = 0;
        for (int32_t q = 0; q < len; q++) WH[base + q] = 0;
        c_wlh = 0;
        wl_len = len;
        wl_min_e = bmin;
      } else {
        while (wl_min_e > bmin) {
          wl_min_e -= args().width;
          wl_len += 1;
        }
        while (wl_len < len) wl_len += 1;
        uint64_t mn = ~0ULL;
        for (int32_t q = 0; q < wl_len; q++) {
          uint64_t h = wlh_at(base + q);
          if (h < mn) mn = h;
        }
        wl_lowest = mn;
      }
    }
    if (i >= wl_len) {
      status = SADMC_WS_INDEX;
      return;
    }
    c_wlh += 1;
    WH[cur_p] = c_wlh;
    if (c_wlh > wl_highest) wl_highest = c_wlh;
    wl_total += 1;
    double max_energy = wl_min_e + (double)wl_len * args().width;
    if (c_wlh == wl_lowest + 1 && wl_len > 1 && (!args().has_min_e || args().min_e >= wl_min_e) &&
        (!args().has_max_e || args().max_e <= max_energy)) {
      bool found = false;
      uint64_t mn = ~0ULL;
      for (int32_t q = 0; q < wl_len; q++) {
        if (hist_at(base + q) != 0) {
          found = true;
          uint64_t h = wlh_at(base + q);
          if (h < mn) mn = h;
        }
      }
      if (found && mn == wl_lowest + 1) {
        wl_lowest = c_wlh;
        if ((inv_t && wl_lowest > 0) || (double)wl_lowest >= 0.8 * (double)wl_total / wl_Ns) {
          wl_gamma *= 0.5;
          for (int32_t q = 0; q < wl_len; q++) WH[base + q] = 0;
          c_wlh = 0;
          wl_total = 0;
          wl_lowest = 0;
          wl_highest = 0;
          if (has_mg && wl_gamma < args().method_param) wl_gamma = 0.0;
        }
        if (inv_t && wl_gamma < wl_Ns / tm) {
          method = SADMC_SAMC;
          t0 = wl_Ns;
        }
      }
    }
  }

  // energy.rs:904-974.
  SADMC_DF void move_once() {
    bool regrow = false;
    moves += 1;
    tm += 1.0;
    gdirty = true;
    if (SYS != SYS_ISING && moves == next_check) {
      if (args().refresh > 0) {
        E = full_energy();
        next_check += (uint64_t)args().refresh;
        // the refreshed energy may lie in another bin than the tracked one
        // (move_once indexes the bins by the energy itself, energy.rs:904-974),
        // or past the window's edge by the FP32 drift: the window then grows
        // to it as for a proposed energy (energy.rs:400-434), together with
        // the proposal's growth below (the order of the two growths does not
        // change the bins they add); the previous move's pending sums belong
        // to the bin it left
        if (e_index(E) != cur_i || E < bmin) {
          apply_pending();
          refresh_indices();
          regrow = cur_p < 0 || E < bmin;
        }
      } else {
        verify_energy();
        next_check += (uint64_t)len * (uint64_t)len * 1000ULL;
      }
    }
    const double e1 = E;
    int32_t i1 = cur_i;
    double rs;
    if (SYS == SYS_ISING) {
      rs = sqrt(1.0 / tm);
    } else {
      // each lane of the group computes one of the next G values at once
      if (rs_next >= G) {
        rs_lane = sqrt(1.0 / (tm + (double)lane));
        rs_next = 0;
      }
      rs = __shfl(rs_lane, rs_next, G);
      rs_next += 1;
    }
    acc_rate *= 1.0 - rs;
    double e2;
    const bool some = plan_move(tscale, usc, e2);
    apply_pending();
    bool go = false;
    if (some) {
      bool oob = false;
      if (args().has_max_e) oob = e2 > args().max_e && e2 > e1;
      if (args().has_min_e) oob = oob || (e2 < args().min_e && e2 < e1);
      go = !oob;
    }
    if (go || regrow) {
      double lo = go ? e2 : E, hi = lo;
      if (regrow) {
        lo = E < lo ? E : lo;
        hi = E > hi ? E : hi;
      }
      if (!prepare_for_range(lo, hi)) return;
    }
    if (go) {
      i1 = cur_i;
      int32_t i2 = e_index(e2);
      if (!indices_valid() || !idx_ok(i2)) {
        status = SADMC_WS_INDEX;
        return;
      }
      const int32_t p2 = base + i2;
      double w2;
      uint64_t h2;
      if (p2 == cur_p) {
        w2 = c_lnw;
        h2 = c_hist;
      } else {
        HotBin hb = H[p2];
        w2 = hb.lnw;
        h2 = hb.hist;
      }
      // independent of the record being loaded
      if (METH == M_SAD) {
        gpre = sad_gamma();
        gdirty = false;
      }
      spec_draw();
      if (!reject_move(e1, e2, w2, h2)) {
        accepted += 1;
        acc_rate += rs;
        confirm();
        int32_t i = (E == e2) ? i2 : e_index(E);
        if (!idx_ok(i)) return;
        const int32_t p = base + i;
        if (p != cur_p) {
          if (p == p2) {
            set_current(p, w2, h2);
          } else {
            HotBin hb = H[p];
            set_current(p, hb.lnw, hb.hist);
          }
        }
        cur_i = i;
      }
    }
    if (status) return;
    if (cur_p < 0) {
      status = SADMC_WS_INDEX;
      return;
    }
    const double En = E;
    const int32_t i = cur_i;
    const int32_t p = cur_p;
    if (c_hist == 0) {
      Cb[p].tfound = moves;
      if (METH == M_SAD && i >= ilo && i <= ihi) tmax = moves;
    }
    c_hist += 1;
    if (SYS == SYS_WCA && args().collect_pressure && moves == next_press) {
      double pv = pressure();
      has_pressure = 1;
      X[p].count += 1;
      X[p].total += pv;
      next_press += (uint64_t)n * (uint64_t)n;
    }
    if (METH == M_SAD) {
      update_sad(En, i);
    } else if (METH == M_SAMC) {
      c_lnw += tm > t0 ? t0 / tm : 1.0;
    } else if (METH == M_WL) {
      update_wl(i);
    }
    {
      HotBin hb;
      hb.lnw = c_lnw;
      hb.hist = c_hist;
      H[p] = hb;
    }
    if (status) return;
    int32_t rt_cand = 0;
    if (c_lnw > max_S) {
      max_S = c_lnw;
      max_S_index = i;
      epoch += 1;
      bulk = 1;
    } else if (i == max_S_index) {
      if (i1 != i) {
        epoch += 1;
        bulk = 0;
      }
    } else {
      rt_cand = 1;
    }
    pend = 1;
    pend_E = En;
    pend_rt = rt_cand;
    pend_epoch = epoch;
    pend_bulk = bulk;
  }

  SADMC_DF void run(uint64_t nmoves) {
    for (uint64_t k = 0; k < nmoves; k++) {
      if (status) break;
      move_once();
    }
  }

  // energy.rs:830-898 after the system exists (initial energy, greedy
  // descent below max_e, one-bin histogram, method state).
  SADMC_DF void from_params(int walker, double e_lo) {
    const sadmc_walker_state &w = args().ws[walker];
    s0 = w.rng[0];
    s1 = w.rng[1];
    if (SYS != SYS_ISING) {
      if (SYS == SYS_LJ) R2 = w.geom * w.geom;
      L = w.geom;
      halfL = 0.5 * w.geom;
    }
    status = 0;
    E = full_energy();
    err = 0.0;
    if (args().has_max_e) {
      const double sc = SYS == SYS_ISING ? 0.0 : uniform_scale(-0.05, 0.05);
      for (uint64_t k = 0; k < 100000000ULL; k++) {
        double newe;
        if (cube_plan(0.05, sc, newe)) {
          if (newe < E) confirm();
          if (E < args().max_e) break;
        }
      }
    }
    const double width = args().width;
    double emin = (round(E / width) - 0.5) * width;
    bmin = emin;
    len = 1;
    double b = floor((emin - e_lo) / width) + 2.0;
    if (!(b >= 0.0) || b >= (double)(args().cap - 2)) {
      status = SADMC_WS_CAPACITY;
      base = 0;
    } else {
      base = (int32_t)b;
      HotBin hb;
      hb.lnw = 0.0;
      hb.hist = 1;
      H[base] = hb;
      ColdBin cb;
      cb.etot = E;
      cb.esq = E * E;
      cb.rt = 1;
      cb.vep = 0;
      cb.tfound = 0;
      cb.pad = 0;
      Cb[base] = cb;
    }
    max_S = 0.0;
    max_S_index = 0;
    epoch = 1;
    bulk = 0;
    moves = 0;
    accepted = 0;
    acc_rate = 0.5;
    tscale = args().move_plan == SADMC_TRANSLATION_SCALE ? args().move_param : 0.05;
    method = args().method;
    min_T = args().method == SADMC_SAD ? args().method_param : 0.0;
    t0 = args().method == SADMC_SAMC ? args().method_param : 0.0;
    too_lo = E;
    too_hi = E;
    tL = 0;
    tF = 0;
    Ns = 1;
    highest = 1;
    lam = 0.0;
    tmax = 0;
    wl_gamma = 1.0;
    wl_lowest = (args().has_min_e && args().has_max_e) ? 0 : 1;
    wl_highest = 1;
    wl_total = 0;
    wl_Ns = (args().has_min_e && args().has_max_e) ? (args().max_e - args().min_e) / width : 1.0;
    wl_len = 0;
    wl_min_e = E;
    has_pressure = 0;
  }
};

}  // namespace sadmc
