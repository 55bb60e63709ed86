// Device walker engine: one reference `EnergyMC::move_once`
// (src/mc/energy.rs:904-974) per call, for a group of G lanes that share one
// walker.
//
// Layout.  Every lane of a group holds the walker's scalar state in registers
// (identical values; all lanes execute the same bookkeeping, so no broadcast
// is ever needed).  For LJ/WCA, lane l of the group owns atoms l, l+G, l+2G,
// ... (K slots, SoA registers); the group also keeps an LDS mirror of all N
// positions so any lane can read the moved atom.  The energy difference of a
// move is a per-lane partial sum over the lane's atoms followed by an xor
// butterfly across the G lanes.  Ising uses one lane per walker (G = 1) with
// the N*N <= 64 spins bit-packed in one 64-bit register.
//
// Bins.  A walker's bins are a fixed window of `cap` 64-byte records in HBM.
// Logical bin i lives in physical slot base + i; the reference's front
// insertion (energy.rs:400-420) becomes base -= 1, and the slots outside the
// window are zero from allocation, which is exactly the value the reference
// inserts.  The record of the current bin is cached in registers and written
// through, so a move that stays in its bin touches no bin memory.
#pragma once

#include "sadmc_common.h"
#include "../../include/sadmc.h"

namespace sadmc {

enum { SYS_ISING = 0, SYS_LJ = 1, SYS_WCA = 2 };

struct PExtra {
  uint64_t count;
  double total;
};

// Arguments of one walker set, passed by value to the kernels.
struct SetArgs {
  sadmc_walker_state *ws;
  double *pos;       // [W][3][npad]
  uint64_t *spins;   // [W]
  sadmc_bin *bins;   // [W][cap]
  PExtra *extra;     // [W][cap] (WCA pressure) or null
  int64_t cap;
  int32_t W, n, npad, G, K;
  int32_t method, move_plan, has_min_e, has_max_e, has_min_gamma;
  int32_t refresh, collect_pressure, precision, move_kind;
  double method_param, move_param, min_e, max_e, rc2;
  int32_t *flag;     // set to 1 when any walker of the set stops with a status
};

// Pending change of the single-walker plan/confirm entry points.
struct Pending {
  int32_t some, which;
  double tx, ty, tz, e, dabse;
};

__device__ __forceinline__ double group_sum(double v, int G) {
  for (int off = 1; off < G; off <<= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

template <int SYS, int KMAX, int PREC>
struct Engine {
  SetArgs A;
  int lane, G, K, n;
  sadmc_walker_state st;
  sadmc_bin *B;
  PExtra *X;
  double *lx, *ly, *lz;
  double x[KMAX], y[KMAX], z[KMAX];
  uint64_t spins;
  double R2, L, halfL;
  int64_t cur_p;
  sadmc_bin cur;
  // pending move
  int32_t pw;
  double ptx, pty, ptz, pe, pdabse;

  // ---------------------------------------------------------------- setup
  __device__ void load(const SetArgs &a, int walker, int lane_in_group, double *lds_group) {
    A = a;
    lane = lane_in_group;
    G = a.G;
    K = a.K;
    n = a.n;
    st = a.ws[walker];
    B = a.bins + (int64_t)walker * a.cap;
    X = a.extra ? a.extra + (int64_t)walker * a.cap : nullptr;
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
          x[k] = P[j];
          y[k] = P[a.npad + j];
          z[k] = P[2 * a.npad + j];
          lx[j] = x[k];
          ly[j] = y[k];
          lz[j] = z[k];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (SYS == SYS_LJ) R2 = st.geom * st.geom;
      L = st.geom;
      halfL = 0.5 * st.geom;
    }
    cur_p = -1;
  }

  __device__ void store(int walker) {
    if (lane == 0) {
      A.ws[walker] = st;
      if (st.status) *A.flag = 1;
    }
    if (SYS == SYS_ISING) {
      A.spins[walker] = spins;
    } else {
      double *P = A.pos + (int64_t)walker * 3 * A.npad;
#pragma unroll
      for (int k = 0; k < KMAX; k++) {
        if (k < K) {
          int j = lane + G * k;
          P[j] = x[k];
          P[A.npad + j] = y[k];
          P[2 * A.npad + j] = z[k];
        }
      }
    }
  }

  // --------------------------------------------------------------- bins
  __device__ int64_t e_index(double e) const {
    double v = (e - st.bins_min) / st.bins_width;
    if (!(v > 0.0)) return 0;
    if (v >= 9.2e18) return INT64_MAX;
    return (int64_t)v;
  }
  __device__ double i_energy(int64_t i) const { return st.bins_min + ((double)i + 0.5) * st.bins_width; }
  __device__ bool idx_ok(int64_t i) {
    if (i < 0 || i >= st.len) {
      st.status = SADMC_WS_INDEX;
      return false;
    }
    return true;
  }
  __device__ double bin_lnw(int64_t p) const { return p == cur_p ? cur.lnw : B[p].lnw; }
  __device__ uint64_t bin_hist(int64_t p) const { return p == cur_p ? cur.histogram : B[p].histogram; }

  // energy.rs:400-434.  New slots get have_visited = true (visit_epoch =
  // epoch) and round_trips = 1; every other field is already zero.
  __device__ bool prepare_for_state(double e) {
    while (e < st.bins_min) {
      if (st.base == 0) {
        st.status = SADMC_WS_CAPACITY;
        return false;
      }
      st.base -= 1;
      st.len += 1;
      B[st.base].round_trips = 1;
      B[st.base].visit_epoch = st.epoch;
      st.bins_min -= st.bins_width;
    }
    while (e >= st.bins_min + st.bins_width * (double)st.len) {
      int64_t q = st.base + st.len;
      if (q >= A.cap) {
        st.status = SADMC_WS_CAPACITY;
        return false;
      }
      B[q].round_trips = 1;
      B[q].visit_epoch = st.epoch;
      st.len += 1;
    }
    return true;
  }

  // ------------------------------------------------------------ systems
  // Full FP64 energy in lane-group order: lane l sums pairs (i, j<i) for its
  // atoms i; the G partial sums meet in the butterfly.
  __device__ double full_energy() {
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
              acc += lj_pair(norm2(x[k] - lx[j], y[k] - ly[j], z[k] - lz[j]));
            } else {
              acc += wca_pair(norm2(min_image(x[k] - lx[j], L, halfL), min_image(y[k] - ly[j], L, halfL),
                                    min_image(z[k] - lz[j], L, halfL)),
                              A.rc2);
            }
          }
        }
      }
    }
    return group_sum(acc, G);
  }

  // wca.rs:202-218 virial, ordered pairs, lane-group order.
  __device__ double pressure() {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        int i = lane + G * k;
        if (i < n) {
          for (int j = 0; j < n; j++) {
            if (j == i) continue;
            acc += wca_pair_pressure(norm2(min_image(x[k] - lx[j], L, halfL), min_image(y[k] - ly[j], L, halfL),
                                           min_image(z[k] - lz[j], L, halfL)),
                                     A.rc2);
          }
        }
      }
    }
    double p = group_sum(acc, G);
    return p / (3.0 * (L * L * L));
  }

  __device__ int spin(int b) const { return ((spins >> b) & 1ULL) ? 1 : -1; }

  // ising.rs:104-119.
  __device__ bool ising_plan(double &e2) {
    int N = n;
    int i = (int)gen_range(st.rng[0], st.rng[1], 0, (uint64_t)N);
    int j = (int)gen_range(st.rng[0], st.rng[1], 0, (uint64_t)N);
    int jp = j + 1 == N ? 0 : j + 1;
    int jm = j == 0 ? N - 1 : j - 1;
    int ip = i + 1 == N ? 0 : i + 1;
    int im = i == 0 ? N - 1 : i - 1;
    int nt = spin(i + jp * N) + spin(i + jm * N) + spin(ip + j * N) + spin(im + j * N);
    double e = st.E - (double)nt * (double)spin(i + j * N) * 2.0;
    pw = i + j * N;
    pe = e;
    e2 = e;
    return true;
  }

  // lj.rs:365-374 / wca.rs:341-353 with the cube move; lj.rs:86-105 and
  // wca.rs:119-140 for the energy difference.
  __device__ bool particle_plan(double d, double &e2) {
    int w = (int)uniform_index(st.rng[0], st.rng[1], (uint64_t)n);
    double sc = uniform_scale(-d, d);
    double ux = uniform_sample(st.rng[0], st.rng[1], -d, sc);
    double uy = uniform_sample(st.rng[0], st.rng[1], -d, sc);
    double uz = uniform_sample(st.rng[0], st.rng[1], -d, sc);
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
    double acc = 0.0, accabs = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < K) {
        int j = lane + G * k;
        if (j < n && j != w) {
          if (SYS == SYS_LJ) {
            if (PREC == 1) {
              float an = (float)(x[k] - tx), bn = (float)(y[k] - ty), cn = (float)(z[k] - tz);
              float ao = (float)(x[k] - fx), bo = (float)(y[k] - fy), co = (float)(z[k] - fz);
              float r2n = an * an + bn * bn + cn * cn;
              float r2o = ao * ao + bo * bo + co * co;
              acc += (double)lj_pair_f32(r2n) - (double)lj_pair_f32(r2o);
            } else {
              acc += lj_pair(norm2(x[k] - tx, y[k] - ty, z[k] - tz)) - lj_pair(norm2(x[k] - fx, y[k] - fy, z[k] - fz));
            }
          } else {
            double vn = wca_pair(norm2(min_image(x[k] - tx, L, halfL), min_image(y[k] - ty, L, halfL),
                                       min_image(z[k] - tz, L, halfL)),
                                 A.rc2);
            double vo = wca_pair(norm2(min_image(x[k] - fx, L, halfL), min_image(y[k] - fy, L, halfL),
                                       min_image(z[k] - fz, L, halfL)),
                                 A.rc2);
            acc += vn - vo;
            accabs += vn + vo;
          }
        }
      }
    }
    double e = st.E + group_sum(acc, G);
    if (SYS == SYS_WCA) pdabse = group_sum(accabs, G);
    pw = w;
    ptx = tx;
    pty = ty;
    ptz = tz;
    pe = e;
    e2 = e;
    return true;
  }

  __device__ bool plan_move(double d, double &e2) {
    if (SYS == SYS_ISING) return ising_plan(e2);
    return particle_plan(d, e2);
  }

  // lj.rs:106-123.
  __device__ void lj_set_energy(double new_e) {
    double nn = (double)n;
    double new_error = fabs(new_e) > fabs(st.E) ? fabs(new_e) * 1e-15 * nn : fabs(st.E) * 1e-15 * nn;
    st.error = new_error + st.error;
    if (st.error > fabs(new_e) * 1e-14 * nn * nn) {
      st.error *= 0.0;
      st.E = full_energy();
    } else {
      st.E = new_e;
    }
  }

  // wca.rs:164-180.
  __device__ void wca_set_energy(double new_e, double dabse) {
    double nn = (double)n;
    double single = dabse > fabs(new_e) ? 1e-14 * dabse * nn : 1e-14 * fabs(new_e) * nn;
    st.error += single * nn;
    if (st.error > fabs(new_e) * 1e-13 * nn * nn) {
      st.E = full_energy();
      st.error = 1e-15 * st.E * nn;
    } else {
      st.E = new_e;
    }
  }

  // ConfirmSystem::confirm: ising.rs:94-101, lj.rs:338-358, wca.rs:276-301.
  __device__ void confirm() {
    if (SYS == SYS_ISING) {
      spins ^= (1ULL << pw);
      st.E = pe;
      return;
    }
    int owner = pw & (G - 1);
    int slot = pw / G;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (lane == owner && k == slot) {
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
    if (A.refresh > 0) {
      st.E = pe;
    } else if (SYS == SYS_LJ) {
      lj_set_energy(pe);
    } else {
      wca_set_energy(pe, pdabse);
    }
  }

  // lj.rs:249-261, wca.rs:237-250.
  __device__ void verify_energy() {
    double egood = full_energy();
    double nn = (double)n;
    double expected = SYS == SYS_LJ ? fabs(st.E) * 1e-14 * nn * nn : fabs(st.E) * 1e-13 * nn * nn;
    if (fabs(egood - st.E) > expected) st.status = SADMC_WS_VERIFY;
  }

  // ------------------------------------------------------------- method
  // energy.rs:22-38, 799-824.
  __device__ double gamma() const {
    switch (st.method) {
      case SADMC_CANONICAL:
        return 0.0;
      case SADMC_SAD: {
        double lam = st.latest_parameter, t = (double)st.moves, tF = (double)st.tF, ns = (double)st.num_states;
        if (lam * tF * ns == 0.0) return 0.0;
        return (lam + t / tF) / (lam + t / ns * (t / tF));
      }
      case SADMC_SAMC: {
        double t = (double)st.moves;
        return t > st.t0 ? st.t0 / t : 1.0;
      }
      default:
        return st.wl_gamma;
    }
  }

  // energy.rs:440-512.
  __device__ bool reject_move(double e1, double e2) {
    int64_t i1 = e_index(e1), i2 = e_index(e2);
    if (!idx_ok(i1) || !idx_ok(i2)) return true;
    int64_t p1 = st.base + i1, p2 = st.base + i2;
    switch (st.method) {
      case SADMC_SAD: {
        int64_t ilo = e_index(st.too_lo), ihi = e_index(st.too_hi);
        if (!idx_ok(ilo) || !idx_ok(ihi)) return true;
        int64_t q1 = e1 < st.too_lo ? st.base + ilo : (e1 > st.too_hi ? st.base + ihi : p1);
        int64_t q2 = e2 < st.too_lo ? st.base + ilo : (e2 > st.too_hi ? st.base + ihi : p2);
        double w1 = bin_lnw(q1), w2 = bin_lnw(q2);
        uint64_t h2 = bin_hist(p2);
        double lnw1 = e1 < st.too_lo ? w1 + (e1 - st.too_lo) / st.min_T : w1;
        double lnw2 = e2 < st.too_lo ? w2 + (e2 - st.too_lo) / st.min_T : w2;
        bool rejected = lnw2 > lnw1 && gen_f64(st.rng[0], st.rng[1]) > exp(lnw1 - lnw2);
        if (!rejected && h2 == 0 && e2 < st.too_hi && e2 > st.too_lo) {
          st.latest_parameter = (st.too_hi - st.too_lo) / st.min_T;
          st.num_states += 1;
          st.tL = st.moves;
        }
        return rejected;
      }
      case SADMC_SAMC: {
        double lnw1 = bin_lnw(p1), lnw2 = bin_lnw(p2);
        return lnw2 > lnw1 && gen_f64(st.rng[0], st.rng[1]) > exp(lnw1 - lnw2);
      }
      case SADMC_WL:
      case SADMC_INV_T_WL: {
        double lnw1 = bin_lnw(p1), lnw2 = bin_lnw(p2);
        uint64_t h2 = bin_hist(p2);
        bool rejected = lnw2 > lnw1 && gen_f64(st.rng[0], st.rng[1]) > exp(lnw1 - lnw2);
        if (!rejected && h2 == 0 && st.wl_lowest_hist > 0) st.wl_num_states += 1.0;
        return rejected;
      }
      default:
        if (e1 >= e2) return false;
        return gen_f64(st.rng[0], st.rng[1]) > exp((e1 - e2) / A.method_param);
    }
  }

  // Running max of t_found over logical bins [a, b] (physical base + a ..).
  __device__ void tmax_scan(int64_t a, int64_t b) {
    for (int64_t j = a; j <= b; j++) {
      uint64_t t = B[st.base + j].t_found;
      if (t > st.tmax) st.tmax = t;
    }
  }

  // energy.rs:514-761 for the current bin (cur, physical cur_p, logical i).
  __device__ void update_weights(double E, int64_t i) {
    double gam = gamma();
    double old = cur.lnw;
    cur.lnw += gam;
    const int64_t p = cur_p;
    bool to_samc = false;
    double samc_t0 = 0.0;
    if (st.method == SADMC_SAD) {
      if (st.too_lo > st.too_hi || E < st.too_lo || E > st.too_hi) cur.lnw = old;
      B[p].lnw = cur.lnw;
      if (cur.histogram > st.highest_hist) {
        st.highest_hist = cur.histogram;
        if (E > st.too_hi) {
          int64_t ihi = e_index(st.too_hi);
          if (!idx_ok(ihi)) return;
          double vhi = bin_lnw(st.base + ihi);
          int64_t jlo = ihi > 0 ? ihi - 1 : 0, jhi = i + 1 < st.len ? i + 1 : st.len - 1;
          for (int64_t j = jlo; j <= jhi; j++) {
            double ej = i_energy(j);
            if (ej > st.too_hi && ej <= E) {
              int64_t pj = st.base + j;
              double v;
              if (bin_hist(pj) != 0) {
                v = vhi;
                st.num_states += 1;
              } else {
                v = 0.0;
              }
              B[pj].lnw = v;
              if (pj == p) cur.lnw = v;
            }
          }
          st.latest_parameter = (E - st.too_lo) / st.min_T;
          st.tL = st.moves;
          double nh = i_energy(e_index(E));
          int64_t ihi2 = e_index(nh);
          if (ihi2 > ihi) tmax_scan(ihi + 1, ihi2 < st.len ? ihi2 : st.len - 1);
          st.too_hi = nh;
        } else if (E < st.too_lo) {
          int64_t ilo = e_index(st.too_lo);
          if (!idx_ok(ilo)) return;
          int64_t jlo = i > 0 ? i - 1 : 0, jhi = ilo + 1 < st.len ? ilo + 1 : st.len - 1;
          for (int64_t j = jlo; j <= jhi; j++) {
            double ej = i_energy(j);
            if (ej < st.too_lo && ej >= E) {
              int64_t pj = st.base + j;
              double v;
              if (bin_hist(pj) != 0) {
                v = bin_lnw(st.base + ilo) + (ej - st.too_lo) / st.min_T;
                if (v < 0.0) v = 0.0;
                st.num_states += 1;
              } else {
                v = 0.0;
              }
              B[pj].lnw = v;
              if (pj == p) cur.lnw = v;
            }
          }
          st.latest_parameter = (st.too_hi - E) / st.min_T;
          st.tL = st.moves;
          double nl = i_energy(e_index(E));
          int64_t ilo2 = e_index(nl);
          if (ilo2 < ilo) tmax_scan(ilo2, ilo - 1);
          st.too_lo = nl;
        }
      }
      if (st.tL == st.moves) {
        uint64_t old_tF = st.tF;
        st.tF = st.tmax;
        if (old_tF != st.tF && A.move_plan == SADMC_ACCEPTANCE_RATE) {
          double s = st.acceptance_rate / A.move_param;
          s = s < 0.8 ? 0.8 : (s > 1.2 ? 1.2 : s);
          st.translation_scale *= s;
        }
      }
    } else if (st.method == SADMC_WL || st.method == SADMC_INV_T_WL) {
      B[p].lnw = cur.lnw;
      const bool inv_t = st.method == SADMC_INV_T_WL;
      const bool has_mg = !inv_t && A.has_min_gamma;
      if (has_mg && st.wl_gamma < A.method_param) {
        if (i >= st.wl_len) {
          st.status = SADMC_WS_INDEX;
          return;
        }
        cur.wl_hist += 1;
        B[p].wl_hist = cur.wl_hist;
        return;
      }
      if (st.wl_len != st.len) {
        if (st.wl_len == 0 || (st.wl_gamma != 1.0 && st.wl_lowest_hist > 0)) {
          st.wl_gamma = 1.0;
          st.wl_lowest_hist = 0;
          st.wl_highest_hist = 0;
          st.wl_total_hist = 0;
          for (int64_t q = 0; q < st.len; q++) B[st.base + q].wl_hist = 0;
          cur.wl_hist = 0;
          st.wl_len = st.len;
          st.wl_min_energy = st.bins_min;
        } else {
          while (st.wl_min_energy > st.bins_min) {
            st.wl_min_energy -= st.bins_width;
            st.wl_len += 1;
          }
          while (st.wl_len < st.len) st.wl_len += 1;
          uint64_t mn = ~0ULL;
          for (int64_t q = 0; q < st.wl_len; q++) {
            uint64_t h = B[st.base + q].wl_hist;
            if (h < mn) mn = h;
          }
          st.wl_lowest_hist = mn;
        }
      }
      if (i >= st.wl_len) {
        st.status = SADMC_WS_INDEX;
        return;
      }
      cur.wl_hist += 1;
      B[p].wl_hist = cur.wl_hist;
      if (cur.wl_hist > st.wl_highest_hist) st.wl_highest_hist = cur.wl_hist;
      st.wl_total_hist += 1;
      double max_energy = st.wl_min_energy + (double)st.wl_len * st.bins_width;
      if (cur.wl_hist == st.wl_lowest_hist + 1 && st.wl_len > 1 && (!A.has_min_e || A.min_e >= st.wl_min_energy) &&
          (!A.has_max_e || A.max_e <= max_energy)) {
        bool found = false;
        uint64_t mn = ~0ULL;
        for (int64_t q = 0; q < st.wl_len; q++) {
          if (B[st.base + q].histogram != 0) {
            found = true;
            uint64_t h = B[st.base + q].wl_hist;
            if (h < mn) mn = h;
          }
        }
        if (found && mn == st.wl_lowest_hist + 1) {
          st.wl_lowest_hist = cur.wl_hist;
          if ((inv_t && st.wl_lowest_hist > 0) ||
              (double)st.wl_lowest_hist >= 0.8 * (double)st.wl_total_hist / st.wl_num_states) {
            st.wl_gamma *= 0.5;
            for (int64_t q = 0; q < st.wl_len; q++) B[st.base + q].wl_hist = 0;
            cur.wl_hist = 0;
            st.wl_total_hist = 0;
            st.wl_lowest_hist = 0;
            st.wl_highest_hist = 0;
            if (has_mg && st.wl_gamma < A.method_param) st.wl_gamma = 0.0;
          }
          if (inv_t && st.wl_gamma < st.wl_num_states / (double)st.moves) {
            to_samc = true;
            samc_t0 = st.wl_num_states;
          }
        }
      }
    } else {
      B[p].lnw = cur.lnw;
    }
    if (to_samc) {
      st.method = SADMC_SAMC;
      st.t0 = samc_t0;
    }
  }

  // energy.rs:904-974.
  __device__ void move_once() {
    st.moves += 1;
    if (SYS != SYS_ISING) {
      if (A.refresh > 0) {
        if (st.moves % (uint64_t)A.refresh == 0) st.E = full_energy();
      } else {
        uint64_t Lb = (uint64_t)st.len;
        if (st.moves % (Lb * Lb * 1000ULL) == 0) verify_energy();
      }
    }
    double e1 = st.E;
    double rs = sqrt(1.0 / (double)st.moves);
    st.acceptance_rate *= 1.0 - rs;
    double e2;
    if (plan_move(st.translation_scale, e2)) {
      bool oob = false;
      if (A.has_max_e) oob = e2 > A.max_e && e2 > e1;
      if (A.has_min_e) oob = oob || (e2 < A.min_e && e2 < e1);
      if (!oob) {
        if (!prepare_for_state(e2)) return;
        if (!reject_move(e1, e2)) {
          st.accepted_moves += 1;
          st.acceptance_rate += rs;
          confirm();
        }
      }
    }
    if (st.status) return;
    double E = st.E;
    int64_t i = e_index(E);
    if (!idx_ok(i)) return;
    int64_t p = st.base + i;
    if (p != cur_p) {
      cur = B[p];
      cur_p = p;
    }
    if (cur.histogram == 0) {
      cur.t_found = st.moves;
      B[p].t_found = st.moves;
      if (st.method == SADMC_SAD) {
        int64_t ilo = e_index(st.too_lo), ihi = e_index(st.too_hi);
        if (i >= ilo && i <= ihi && st.moves > st.tmax) st.tmax = st.moves;
      }
    }
    cur.histogram += 1;
    cur.energy_total += E;
    cur.energy_squared_total += E * E;
    B[p].histogram = cur.histogram;
    B[p].energy_total = cur.energy_total;
    B[p].energy_squared_total = cur.energy_squared_total;
    if (SYS == SYS_WCA && A.collect_pressure) {
      if (st.moves % ((uint64_t)n * (uint64_t)n) == 0) {
        double pv = pressure();
        st.has_pressure = 1;
        X[p].count += 1;
        X[p].total += pv;
      }
    }
    update_weights(E, i);
    if (st.status) return;
    double li = cur.lnw;
    if (li > st.max_S) {
      st.max_S = li;
      st.max_S_index = i;
      st.epoch += 1;
      st.bulk_visited = 1;
    } else if (i == st.max_S_index) {
      if (e_index(e1) != i) {
        st.epoch += 1;
        st.bulk_visited = 0;
      }
    } else if (!(st.bulk_visited || cur.visit_epoch == st.epoch)) {
      cur.visit_epoch = st.epoch;
      cur.round_trips += 1;
      B[p].visit_epoch = cur.visit_epoch;
      B[p].round_trips = cur.round_trips;
    }
  }

  __device__ void run(uint64_t moves) {
    for (uint64_t k = 0; k < moves; k++) {
      if (st.status) break;
      move_once();
    }
  }

  // energy.rs:830-898 after the system exists (initial energy, greedy
  // descent below max_e, one-bin histogram, method state).
  __device__ void from_params(double e_lo, double width) {
    st.E = full_energy();
    st.error = 0.0;
    if (A.has_max_e) {
      for (uint64_t k = 0; k < 100000000ULL; k++) {
        double newe;
        if (plan_move(0.05, newe)) {
          if (newe < st.E) confirm();
          if (st.E < A.max_e) break;
        }
      }
    }
    double E = st.E;
    double emin = (round(E / width) - 0.5) * width;
    st.bins_width = width;
    st.bins_min = emin;
    st.len = 1;
    double b = floor((emin - e_lo) / width) + 2.0;
    if (!(b >= 0.0) || b >= (double)(A.cap - 2)) {
      st.status = SADMC_WS_CAPACITY;
      st.base = 0;
      return;
    }
    st.base = (int64_t)b;
    sadmc_bin r;
    r.lnw = 0.0;
    r.energy_total = E;
    r.energy_squared_total = E * E;
    r.histogram = 1;
    r.t_found = 0;
    r.round_trips = 1;
    r.visit_epoch = 0;
    r.wl_hist = 0;
    B[st.base] = r;
    st.max_S = 0.0;
    st.max_S_index = 0;
    st.epoch = 1;
    st.bulk_visited = 0;
    st.moves = 0;
    st.accepted_moves = 0;
    st.acceptance_rate = 0.5;
    st.translation_scale = A.move_plan == SADMC_TRANSLATION_SCALE ? A.move_param : 0.05;
    st.method = A.method;
    st.min_T = A.method == SADMC_SAD ? A.method_param : 0.0;
    st.t0 = A.method == SADMC_SAMC ? A.method_param : 0.0;
    st.too_lo = E;
    st.too_hi = E;
    st.tL = 0;
    st.tF = 0;
    st.num_states = 1;
    st.highest_hist = 1;
    st.latest_parameter = 0.0;
    st.tmax = 0;
    st.wl_gamma = 1.0;
    st.wl_lowest_hist = (A.has_min_e && A.has_max_e) ? 0 : 1;
    st.wl_highest_hist = 1;
    st.wl_total_hist = 0;
    st.wl_num_states = (A.has_min_e && A.has_max_e) ? (A.max_e - A.min_e) / width : 1.0;
    st.wl_len = 0;
    st.wl_min_energy = E;
    st.has_pressure = 0;
  }
};

}  // namespace sadmc
