/*
 * CPU oracle for the sad-monte-carlo walker hot path -- TEST INFRASTRUCTURE.
 *
 * A line-by-line restatement, in plain C, of the reference's EnergyMC loop
 * and of the systems it drives.  Each function cites the reference lines it
 * follows.  Built with -ffp-contract=off so that every floating-point
 * expression rounds exactly as written, like Rust's (which never contracts).
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg; the product never links it.  Pins are listed in sadmc_oracle.h.
 */
#include "sadmc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Seed tweak for the initial-configuration stream (LJ/WCA positions and
 * per-walker Ising spins), so that it is independent of the MC stream that
 * seed_from_u64(seed) drives.  No reference counterpart (its inits use the
 * fixed seeds 0 and 10137, lj.rs:128, wca.rs:402, ising.rs:45). */
#define OR_INIT_TAG 0xA5F0C3E1D2B49687ULL

enum { ST_OK = 0, ST_INDEX = 1, ST_VERIFY = 2, ST_INIT = 3 };

/* ===================================================================== */
/* Random numbers                                                         */
/* ===================================================================== */

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* SplitMix64 (Vigna), as rand_xoshiro 0.4's SplitMix64::next_u64. */
uint64_t or_splitmix64_next(uint64_t *x) {
  *x += 0x9e3779b97f4a7c15ULL;
  uint64_t z = *x;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* Xoroshiro128Plus::seed_from_u64 (rand_xoshiro 0.4, called at
 * energy.rs:835): SplitMix64 state = seed; the 16-byte seed is filled
 * little-endian from two SplitMix64 outputs, so s0, s1 = outputs 1, 2.
 * (Not src/rng.rs:73-93, which the crate does not use: rng.rs:27.) */
void or_rng_seed(uint64_t seed, uint64_t s[2]) {
  uint64_t x = seed;
  s[0] = or_splitmix64_next(&x);
  s[1] = or_splitmix64_next(&x);
}

/* xoroshiro128+ step, src/rng.rs:48-57 (a=24, b=16, c=37). */
uint64_t or_rng_next(uint64_t s[2]) {
  uint64_t s0 = s[0], s1 = s[1];
  uint64_t r = s0 + s1;
  s1 ^= s0;
  s[0] = rotl64(s0, 24) ^ s1 ^ (s1 << 16);
  s[1] = rotl64(s1, 37);
  return r;
}

/* rng.gen_range(low, high) for usize (ising.rs:105-106): rand 0.7
 * UniformInt::sample_single with the conservative zone
 * (range << leading_zeros(range)) - 1 and widening multiply. */
uint64_t or_gen_range(uint64_t s[2], uint64_t low, uint64_t high) {
  uint64_t range = high - low;
  uint64_t zone = (range << __builtin_clzll(range)) - 1;
  for (;;) {
    uint64_t v = or_rng_next(s);
    unsigned __int128 m = (unsigned __int128)v * range;
    if ((uint64_t)m <= zone) return low + (uint64_t)(m >> 64);
  }
}

/* rng.sample(Uniform::new(0, n)) for usize (lj.rs:368, wca.rs:345):
 * rand 0.7 UniformInt with exact zone MAX - (2^64 mod n). */
uint64_t or_uniform_index(uint64_t s[2], uint64_t n) {
  uint64_t reject = (UINT64_MAX - n + 1) % n;
  uint64_t zone = UINT64_MAX - reject;
  for (;;) {
    uint64_t v = or_rng_next(s);
    unsigned __int128 m = (unsigned __int128)v * n;
    if ((uint64_t)m <= zone) return (uint64_t)(m >> 64);
  }
}

/* rng.gen::<f64>() (energy.rs:465,489,498,508): rand 0.7 Standard,
 * (next_u64 >> 11) * 2^-53. */
double or_gen_f64(uint64_t s[2]) {
  uint64_t v = or_rng_next(s);
  return (1.0 / 9007199254740992.0) * (double)(v >> 11);
}

static inline double f64_from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t f64_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* rng.sample(Uniform::new(low, high)) for f64 (lj.rs:138 uses it): rand 0.7
 * UniformFloat -- scale = high-low shrunk until scale*max_rand+low < high,
 * sample = ((bits>>12 | 1.0) - 1) * scale + low. */
double or_uniform_f64(uint64_t s[2], double low, double high) {
  const double max_rand = f64_from_bits((UINT64_MAX >> 12) | 0x3FF0000000000000ULL) - 1.0;
  double scale = high - low;
  while (scale * max_rand + low >= high) scale = f64_from_bits(f64_bits(scale) - 1);
  uint64_t v = or_rng_next(s) >> 12;
  double v12 = f64_from_bits(v | 0x3FF0000000000000ULL);
  return (v12 - 1.0) * scale + low;
}

/* ===================================================================== */
/* Potentials                                                             */
/* ===================================================================== */

/* lj.rs:78-81: 4 eps ((s^2/r^2)^6 - (s^2/r^2)^3), sigma = eps = 1.  powi is
 * expanded by LLVM as binary powering: x^6 = x^2 * (x^2)^2, x^3 = x * x^2. */
double or_lj_pair(double r2) {
  double s = 1.0 / r2;
  double s2 = s * s;
  double s4 = s2 * s2;
  double s6 = s2 * s4;
  double s3 = s * s2;
  return 4.0 * (s6 - s3);
}

/* The same pair term in binary32 (precision variant; no reference
 * counterpart). */
float or_lj_pair_f32(float r2) {
  float s = 1.0f / r2;
  float s2 = s * s;
  float s4 = s2 * s2;
  float s6 = s2 * s4;
  float s3 = s * s2;
  return 4.0f * (s6 - s3);
}

/* wca.rs:66-76: shifted LJ, cut at r_c = 2^(1/6) sigma. */
double or_wca_pair(double r2, double rc2) {
  if (r2 < rc2) {
    double s = 1.0 / r2;
    double s2 = s * s;
    double s4 = s2 * s2;
    double s6 = s2 * s4;
    double s3 = s * s2;
    return 4.0 * (s6 - s3) + 1.0;
  }
  return 0.0;
}

/* wca.rs:79-92: -dV/dr r / 2 = 12 (2 s^6 - s^3). */
static double wca_pair_pressure(double r2, double rc2) {
  if (r2 < rc2) {
    double s = 1.0 / r2;
    double s2 = s * s;
    double s4 = s2 * s2;
    double s6 = s2 * s4;
    double s3 = s * s2;
    return 12.0 * (2.0 * s6 - s3);
  }
  return 0.0;
}

static inline double norm2(double x, double y, double z) { return x * x + y * y + z * z; }

/* Minimum-image separation component for positions inside [0, L). */
static inline double min_image(double d, double L, double halfL) {
  if (d > halfL) return d - L;
  if (d < -halfL) return d + L;
  return d;
}

/* Combine G per-lane partial sums the way an xor butterfly of GPU lanes
 * does (every lane ends with the same total). */
static double butterfly(double *acc, int G) {
  double tmp[64];
  for (int off = 1; off < G; off <<= 1) {
    for (int l = 0; l < G; l++) tmp[l] = acc[l] + acc[l ^ off];
    for (int l = 0; l < G; l++) acc[l] = tmp[l];
  }
  return acc[0];
}

/* ===================================================================== */
/* Walker state                                                           */
/* ===================================================================== */

struct or_mc {
  or_params p;
  int sys, n, G;
  /* system */
  double E, error;
  double *x, *y, *z;
  int8_t *S;
  double R2, L, halfL, rc2;
  int nsub;
  /* pending change (possible_change) */
  int ch_kind; /* 0 none, 1 move */
  int64_t ch_which;
  double ch_x, ch_y, ch_z, ch_e, ch_dabse, ch_lnh;
  /* scripted energies */
  double *script;
  int64_t script_len, script_pos;
  /* MC */
  uint64_t rng[2];
  uint64_t moves, accepted;
  double acc_rate, tscale;
  double width, min;
  int64_t len, cap;
  double *lnw, *etot, *esq, *ptotal;
  uint64_t *hist, *tfound, *rt, *pcount;
  uint8_t *vis;
  int has_pressure;
  double max_S;
  int64_t max_S_index;
  int method;
  double min_T, too_lo, too_hi, lambda;
  uint64_t tL, tF, Ns, highest;
  double t0, T;
  double wl_gamma, wl_Ns, wl_min_energy, min_gamma;
  uint64_t wl_lowest, wl_highest, wl_total;
  uint64_t *wlh;
  int64_t wl_len, wl_cap;
  int wl_inv_t, has_min_gamma;
  /* trace */
  double *tr_e;
  uint8_t *tr_a;
  uint64_t tr_cap, tr_n;
  /* the last move's decision inputs (test hook): proposed energy (NaN when
   * the move was None), log acceptance ratio and uniform of the
   * Metropolis-Hastings test (NaN when no uniform was drawn), |to|^2 - R^2
   * of a learned LJ move's wall test */
  double lm_e2, lm_lr, lm_u, lm_wall;
  int status;
};

/* ---- Ising (ising.rs) ------------------------------------------------ */

/* ising.rs:45-49: spins are the low bit of successive draws, in site order.
 * rule 0 uses the reference's fixed seed 10137; rule 1 a per-walker stream
 * (spins i.i.d. from each walker's own seed). */
void or_ising_init_spins(int32_t n, int32_t rule, uint64_t seed, int8_t *S) {
  uint64_t s[2];
  or_rng_seed(rule == 0 ? 10137ULL : (seed ^ OR_INIT_TAG), s);
  for (int64_t k = 0; k < (int64_t)n * n; k++) S[k] = (int8_t)(((int8_t)or_rng_next(s) & 1) * 2 - 1);
}

/* ising.rs:59-75. */
double or_ising_energy(int32_t N, const int8_t *S) {
  double e = 0.0;
  for (int i1 = 0; i1 < N; i1++)
    for (int j1 = 0; j1 < N; j1++) {
      int j2 = (j1 + 1) % N;
      int nt = S[i1 + j2 * N];
      int i2 = (i1 + 1) % N;
      nt += S[i2 + j1 * N];
      e += (double)nt * (double)S[i1 + j1 * N];
    }
  return e;
}

/* ---- LJ (lj.rs) ------------------------------------------------------- */

/* Initial positions (the workloads' rule, no reference counterpart;
 * lj.rs:126-205 is the reference's own rule): atoms added one at a time, each drawn
 * uniformly in the sphere by rejection from the cube [-R,R)^3 and redrawn
 * until every earlier atom is at least 0.8 sigma away. */
int or_lj_init(int32_t n, double radius, uint64_t seed, double *c) {
  uint64_t s[2];
  or_rng_seed(seed ^ OR_INIT_TAG, s);
  double *x = c, *y = c + n, *z = c + 2 * n;
  const double R2 = radius * radius, dmin2 = 0.8 * 0.8;
  for (int a = 0; a < n; a++) {
    uint64_t tries = 0;
    for (;;) {
      if (++tries > 100000000ULL) return -1;
      double px = or_uniform_f64(s, -radius, radius);
      double py = or_uniform_f64(s, -radius, radius);
      double pz = or_uniform_f64(s, -radius, radius);
      if (norm2(px, py, pz) >= R2) continue;
      int ok = 1;
      for (int b = 0; b < a; b++)
        if (norm2(px - x[b], py - y[b], pz - z[b]) < dmin2) { ok = 0; break; }
      if (!ok) continue;
      x[a] = px; y[a] = py; z[a] = pz;
      break;
    }
  }
  return 0;
}

/* lj.rs:236-244 (order 0) or the lane-group order. */
double or_lj_energy(int32_t n, const double *c, int32_t G) {
  const double *x = c, *y = c + n, *z = c + 2 * n;
  if (G <= 0) {
    double e = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < i; j++) e += or_lj_pair(norm2(x[i] - x[j], y[i] - y[j], z[i] - z[j]));
    return e;
  }
  double acc[64];
  for (int l = 0; l < G; l++) {
    acc[l] = 0.0;
    for (int i = l; i < n; i += G)
      for (int j = 0; j < i; j++) acc[l] += or_lj_pair(norm2(x[i] - x[j], y[i] - y[j], z[i] - z[j]));
  }
  return butterfly(acc, G);
}

double or_lj_energy_ld(int32_t n, const double *c) {
  const double *x = c, *y = c + n, *z = c + 2 * n;
  long double e = 0.0L;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) {
      long double dx = (long double)x[i] - x[j], dy = (long double)y[i] - y[j],
                  dz = (long double)z[i] - z[j];
      long double r2 = dx * dx + dy * dy + dz * dz;
      long double s = 1.0L / r2, s3 = s * s * s;
      e += 4.0L * (s3 * s3 - s3);
    }
  return (double)e;
}

/* ---- WCA (wca.rs, optcell.rs) ------------------------------------------ */

/* Initial positions (no reference counterpart; wca.rs:392-498 keeps the
 * lowest of N^2 random trials).  Of the simple-cubic, bcc and fcc lattices,
 * each with the fewest cubic cells per edge m that give at least n sites,
 * take the one with the largest nearest-neighbour distance (factor/m, ties
 * to fcc, then bcc); choose n of its sites by a partial Fisher-Yates
 * shuffle driven by the init stream. */
static const double WCA_BASIS[3][4][3] = {
    {{0.0, 0.0, 0.0}, {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}}, /* fcc */
    {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}},                                   /* bcc */
    {{0.0, 0.0, 0.0}}};                                                   /* sc */
static const int WCA_NBASIS[3] = {4, 2, 1};
static const double WCA_NN[3] = {0.7071067811865476, 0.8660254037844386, 1.0};

int or_wca_init(int32_t n, double box, uint64_t seed, double *c) {
  uint64_t s[2];
  or_rng_seed(seed ^ OR_INIT_TAG, s);
  int best = -1, bm = 0;
  double bestnn = -1.0;
  for (int t = 0; t < 3; t++) {
    int m = 1;
    while ((int64_t)WCA_NBASIS[t] * m * m * m < n) m++;
    double nn = WCA_NN[t] / (double)m;
    if (nn > bestnn) { bestnn = nn; best = t; bm = m; }
  }
  int m = bm, nb = WCA_NBASIS[best];
  int64_t ns = (int64_t)nb * m * m * m;
  double *sx = malloc(sizeof(double) * ns), *sy = malloc(sizeof(double) * ns),
         *sz = malloc(sizeof(double) * ns);
  double a = box / (double)m;
  int64_t k = 0;
  for (int ix = 0; ix < m; ix++)
    for (int iy = 0; iy < m; iy++)
      for (int iz = 0; iz < m; iz++)
        for (int b = 0; b < nb; b++) {
          sx[k] = ((double)ix + WCA_BASIS[best][b][0]) * a;
          sy[k] = ((double)iy + WCA_BASIS[best][b][1]) * a;
          sz[k] = ((double)iz + WCA_BASIS[best][b][2]) * a;
          k++;
        }
  for (int64_t i = 0; i < n; i++) {
    int64_t r = i + (int64_t)or_uniform_index(s, (uint64_t)(ns - i));
    double t;
    t = sx[i]; sx[i] = sx[r]; sx[r] = t;
    t = sy[i]; sy[i] = sy[r]; sy[r] = t;
    t = sz[i]; sz[i] = sz[r]; sz[r] = t;
  }
  for (int i = 0; i < n; i++) { c[i] = sx[i]; c[n + i] = sy[i]; c[2 * n + i] = sz[i]; }
  free(sx); free(sy); free(sz);
  return 0;
}

/* optcell.rs:61-73 (cell count), :112-124 (get_subcell). */
static int wca_nsub(double L, double rc2) { return (int)floor(L / sqrt(rc2)); }
static inline int64_t subcell1(double r, double L, int nsub) { return (int64_t)floor(r / L * (double)nsub); }

/* optcell.rs:373-405 stencil. */
static const int NB[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
    {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 0, 1},   {1, 0, -1}, {-1, 0, 1},
    {-1, 0, -1}, {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 1, 1},  {-1, 1, 1},
    {1, -1, 1},  {1, 1, -1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

static inline int64_t modulus(int64_t i, int64_t sz) { return (i + sz) % sz; }

/* Sum of f over the reference's maybe_interacting_atoms_excluding(r, which)
 * (optcell.rs:93-110): every (atom, image) entry that add_to_subcells
 * (optcell.rs:131-160) files under the subcell of r, with the image shifted
 * by -offset*L.  Entries are visited in atom order, then stencil order. */
typedef double (*pair_fn)(double r2, double rc2);
static double celllist_sum(double init, const double *x, const double *y, const double *z, int n,
                           int64_t which, double rx, double ry, double rz, double L, double rc2, int nsub,
                           pair_fn f, double sign_r_minus_img, double *abs_acc) {
  int64_t scx = subcell1(rx, L, nsub), scy = subcell1(ry, L, nsub), scz = subcell1(rz, L, nsub);
  int64_t qx = modulus(scx, nsub), qy = modulus(scy, nsub), qz = modulus(scz, nsub);
  double acc = init;
  for (int j = 0; j < n; j++) {
    if (j == which) continue;
    int64_t jx = subcell1(x[j], L, nsub), jy = subcell1(y[j], L, nsub), jz = subcell1(z[j], L, nsub);
    for (int t = 0; t < 27; t++) {
      int64_t px = jx + NB[t][0], py = jy + NB[t][1], pz = jz + NB[t][2];
      if (modulus(px, nsub) != qx || modulus(py, nsub) != qy || modulus(pz, nsub) != qz) continue;
      int ox = px < 0 ? -1 : (px == nsub ? 1 : 0);
      int oy = py < 0 ? -1 : (py == nsub ? 1 : 0);
      int oz = pz < 0 ? -1 : (pz == nsub ? 1 : 0);
      double ix = x[j] - (double)ox * L, iy = y[j] - (double)oy * L, iz = z[j] - (double)oz * L;
      double r2 = sign_r_minus_img > 0 ? norm2(rx - ix, ry - iy, rz - iz) : norm2(ix - rx, iy - ry, iz - rz);
      double v = f(r2, rc2);
      acc += v;
      if (abs_acc) *abs_acc += v;
    }
  }
  return acc;
}

/* wca.rs:222-230 through the cell lists. */
double or_wca_energy_celllist(int32_t n, const double *c, double L, double rc2) {
  const double *x = c, *y = c + n, *z = c + 2 * n;
  int nsub = wca_nsub(L, rc2);
  double e = 0.0;
  for (int i = 0; i < n; i++) {
    /* e += V(|r1 - r2|^2) for each neighbour entry r2 of r1 */
    int64_t scx = subcell1(x[i], L, nsub), scy = subcell1(y[i], L, nsub), scz = subcell1(z[i], L, nsub);
    int64_t qx = modulus(scx, nsub), qy = modulus(scy, nsub), qz = modulus(scz, nsub);
    for (int j = 0; j < n; j++) {
      if (j == i) continue;
      int64_t jx = subcell1(x[j], L, nsub), jy = subcell1(y[j], L, nsub), jz = subcell1(z[j], L, nsub);
      for (int t = 0; t < 27; t++) {
        int64_t px = jx + NB[t][0], py = jy + NB[t][1], pz = jz + NB[t][2];
        if (modulus(px, nsub) != qx || modulus(py, nsub) != qy || modulus(pz, nsub) != qz) continue;
        int ox = px < 0 ? -1 : (px == nsub ? 1 : 0);
        int oy = py < 0 ? -1 : (py == nsub ? 1 : 0);
        int oz = pz < 0 ? -1 : (pz == nsub ? 1 : 0);
        double ix = x[j] - (double)ox * L, iy = y[j] - (double)oy * L, iz = z[j] - (double)oz * L;
        e += or_wca_pair(norm2(x[i] - ix, y[i] - iy, z[i] - iz), rc2);
      }
    }
  }
  return e * 0.5;
}

/* Minimum-image energy, reference-like sequential or lane-group order. */
double or_wca_energy(int32_t n, const double *c, double L, double rc2, int32_t G) {
  const double *x = c, *y = c + n, *z = c + 2 * n;
  double h = 0.5 * L;
  if (G <= 0) G = 1;
  double acc[64];
  for (int l = 0; l < G; l++) {
    acc[l] = 0.0;
    for (int i = l; i < n; i += G)
      for (int j = 0; j < i; j++)
        acc[l] += or_wca_pair(norm2(min_image(x[i] - x[j], L, h), min_image(y[i] - y[j], L, h),
                                    min_image(z[i] - z[j], L, h)),
                              rc2);
  }
  return butterfly(acc, G);
}

double or_wca_energy_ld(int32_t n, const double *c, double L, double rc2) {
  const double *x = c, *y = c + n, *z = c + 2 * n;
  long double e = 0.0L, h = 0.5L * L;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) {
      long double d[3] = {(long double)x[i] - x[j], (long double)y[i] - y[j], (long double)z[i] - z[j]};
      for (int k = 0; k < 3; k++) {
        if (d[k] > h) d[k] -= L;
        else if (d[k] < -h) d[k] += L;
      }
      long double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 < (long double)rc2) {
        long double s = 1.0L / r2, s3 = s * s * s;
        e += 4.0L * (s3 * s3 - s3) + 1.0L;
      }
    }
  return (double)e;
}

/* optcell.rs:277-309. */
static inline double put_in_cell1(double r, double L) {
  if (r < 0.0) {
    do { r += L; } while (r < 0.0);
  } else {
    while (r >= L) r -= L;
  }
  return r;
}

/* ===================================================================== */
/* System dispatch (any.rs) for the walker                               */
/* ===================================================================== */

static double sys_compute_energy(const or_mc *m, int G) {
  if (m->sys == OR_ISING) return or_ising_energy(m->n, m->S);
  if (m->sys == OR_SCRIPTED) return m->E;
  int n = m->n;
  double *c = malloc(sizeof(double) * 3 * n);
  memcpy(c, m->x, sizeof(double) * n);
  memcpy(c + n, m->y, sizeof(double) * n);
  memcpy(c + 2 * n, m->z, sizeof(double) * n);
  double e;
  if (m->sys == OR_LJ) e = or_lj_energy(n, c, G);
  else if (G <= 0) e = or_wca_energy_celllist(n, c, m->L, m->rc2);
  else e = or_wca_energy(n, c, m->L, m->rc2, G);
  free(c);
  return e;
}

/* Pair-term difference for atom `w` moving from f to t against atom j. */
static inline double lj_term(const or_mc *m, int j, double tx, double ty, double tz, double fx, double fy,
                             double fz) {
  double xj = m->x[j], yj = m->y[j], zj = m->z[j];
  if (m->p.precision == 1) {
    float an = (float)(xj - tx), bn = (float)(yj - ty), cn = (float)(zj - tz);
    float ao = (float)(xj - fx), bo = (float)(yj - fy), co = (float)(zj - fz);
    float r2n = an * an + bn * bn + cn * cn;
    float r2o = ao * ao + bo * bo + co * co;
    return (double)or_lj_pair_f32(r2n) - (double)or_lj_pair_f32(r2o);
  }
  return or_lj_pair(norm2(xj - tx, yj - ty, zj - tz)) - or_lj_pair(norm2(xj - fx, yj - fy, zj - fz));
}

/* lj.rs:365-374 with the cube move, lj.rs:86-105 for the energy. */
static int lj_plan_move(or_mc *m, double d, double *e_out) {
  int64_t w = (int64_t)or_uniform_index(m->rng, (uint64_t)m->n);
  double ux = or_uniform_f64(m->rng, -d, d);
  double uy = or_uniform_f64(m->rng, -d, d);
  double uz = or_uniform_f64(m->rng, -d, d);
  double fx = m->x[w], fy = m->y[w], fz = m->z[w];
  double tx = fx + ux, ty = fy + uy, tz = fz + uz;
  double r2to = norm2(tx, ty, tz), r2from = norm2(fx, fy, fz);
  if (r2to > m->R2 && r2to > r2from) return 0;
  double e;
  if (m->G <= 0) {
    e = m->E;
    for (int j = 0; j < m->n; j++)
      if (j != w) e += lj_term(m, j, tx, ty, tz, fx, fy, fz);
  } else {
    double acc[64];
    for (int l = 0; l < m->G; l++) {
      acc[l] = 0.0;
      for (int j = l; j < m->n; j += m->G)
        if (j != w) acc[l] += lj_term(m, j, tx, ty, tz, fx, fy, fz);
    }
    e = m->E + butterfly(acc, m->G);
  }
  m->ch_kind = 1; m->ch_which = w; m->ch_x = tx; m->ch_y = ty; m->ch_z = tz; m->ch_e = e;
  *e_out = e;
  return 1;
}

/* wca.rs:341-353 with the cube move, wca.rs:119-140 for the energy. */
static int wca_plan_move(or_mc *m, double d, double *e_out) {
  int64_t w = (int64_t)or_uniform_index(m->rng, (uint64_t)m->n);
  double ux = or_uniform_f64(m->rng, -d, d);
  double uy = or_uniform_f64(m->rng, -d, d);
  double uz = or_uniform_f64(m->rng, -d, d);
  double fx = m->x[w], fy = m->y[w], fz = m->z[w];
  double tx = put_in_cell1(fx + ux, m->L), ty = put_in_cell1(fy + uy, m->L), tz = put_in_cell1(fz + uz, m->L);
  double e, dabse;
  if (m->G <= 0) {
    dabse = 0.0;
    e = celllist_sum(m->E, m->x, m->y, m->z, m->n, w, tx, ty, tz, m->L, m->rc2, m->nsub, or_wca_pair, -1.0,
                     &dabse);
    /* e -= de for each entry, one at a time (wca.rs:128-132) */
    {
      int64_t scx = subcell1(fx, m->L, m->nsub), scy = subcell1(fy, m->L, m->nsub), scz = subcell1(fz, m->L, m->nsub);
      int64_t qx = modulus(scx, m->nsub), qy = modulus(scy, m->nsub), qz = modulus(scz, m->nsub);
      for (int j = 0; j < m->n; j++) {
        if (j == w) continue;
        int64_t jx = subcell1(m->x[j], m->L, m->nsub), jy = subcell1(m->y[j], m->L, m->nsub),
                jz = subcell1(m->z[j], m->L, m->nsub);
        for (int t = 0; t < 27; t++) {
          int64_t px = jx + NB[t][0], py = jy + NB[t][1], pz = jz + NB[t][2];
          if (modulus(px, m->nsub) != qx || modulus(py, m->nsub) != qy || modulus(pz, m->nsub) != qz) continue;
          int ox = px < 0 ? -1 : (px == m->nsub ? 1 : 0);
          int oy = py < 0 ? -1 : (py == m->nsub ? 1 : 0);
          int oz = pz < 0 ? -1 : (pz == m->nsub ? 1 : 0);
          double ix = m->x[j] - (double)ox * m->L, iy = m->y[j] - (double)oy * m->L, iz = m->z[j] - (double)oz * m->L;
          double de = or_wca_pair(norm2(ix - fx, iy - fy, iz - fz), m->rc2);
          e -= de;
          dabse += de;
        }
      }
    }
  } else {
    double acc[64], accabs[64];
    double h = m->halfL, L = m->L;
    for (int l = 0; l < m->G; l++) {
      acc[l] = 0.0; accabs[l] = 0.0;
      for (int j = l; j < m->n; j += m->G) {
        if (j == w) continue;
        double vn = or_wca_pair(norm2(min_image(m->x[j] - tx, L, h), min_image(m->y[j] - ty, L, h),
                                      min_image(m->z[j] - tz, L, h)), m->rc2);
        double vo = or_wca_pair(norm2(min_image(m->x[j] - fx, L, h), min_image(m->y[j] - fy, L, h),
                                      min_image(m->z[j] - fz, L, h)), m->rc2);
        acc[l] += vn - vo;
        accabs[l] += vn + vo;
      }
    }
    e = m->E + butterfly(acc, m->G);
    dabse = butterfly(accabs, m->G);
  }
  m->ch_kind = 1; m->ch_which = w; m->ch_x = tx; m->ch_y = ty; m->ch_z = tz; m->ch_e = e; m->ch_dabse = dabse;
  *e_out = e;
  return 1;
}

/* ising.rs:104-119. */
static int ising_plan_move(or_mc *m, double *e_out) {
  int N = m->n;
  int64_t i = (int64_t)or_gen_range(m->rng, 0, (uint64_t)N);
  int64_t j = (int64_t)or_gen_range(m->rng, 0, (uint64_t)N);
  int64_t j2 = (j + 1) % N;
  int nt = m->S[i + j2 * N];
  j2 = (j + N - 1) % N;
  nt += m->S[i + j2 * N];
  int64_t i2 = (i + 1) % N;
  nt += m->S[i2 + j * N];
  i2 = (i + N - 1) % N;
  nt += m->S[i2 + j * N];
  double e = m->E - (double)nt * (double)m->S[i + j * N] * 2.0;
  m->ch_kind = 1; m->ch_which = i + j * N; m->ch_e = e;
  *e_out = e;
  return 1;
}

static int sys_plan_move(or_mc *m, double d, double *e) {
  switch (m->sys) {
    case OR_ISING: return ising_plan_move(m, e);
    case OR_LJ: return lj_plan_move(m, d, e);
    case OR_WCA: return wca_plan_move(m, d, e);
    default: {
      double v = m->script[m->script_pos % m->script_len];
      m->script_pos++;
      if (isnan(v)) return 0;
      m->ch_kind = 1; m->ch_e = v;
      *e = v;
      return 1;
    }
  }
}

/* lj.rs:106-123 (error-estimate refresh). */
static void lj_set_energy(or_mc *m, double new_e) {
  double n = (double)m->n;
  double new_error = fabs(new_e) > fabs(m->E) ? fabs(new_e) * 1e-15 * n : fabs(m->E) * 1e-15 * n;
  m->error = new_error + m->error;
  if (m->error > fabs(new_e) * 1e-14 * n * n) {
    m->error *= 0.0;
    m->E = sys_compute_energy(m, m->G);
  } else {
    m->E = new_e;
  }
}

/* wca.rs:164-180. */
static void wca_set_energy(or_mc *m, double new_e, double dabse) {
  double n = (double)m->n;
  double single = dabse > fabs(new_e) ? 1e-14 * dabse * n : 1e-14 * fabs(new_e) * n;
  m->error += single * n;
  if (m->error > fabs(new_e) * 1e-13 * n * n) {
    m->E = sys_compute_energy(m, m->G);
    m->error = 1e-15 * m->E * n;
  } else {
    m->E = new_e;
  }
}

/* ConfirmSystem::confirm: ising.rs:94-101, lj.rs:338-358, wca.rs:276-301. */
static void sys_confirm(or_mc *m) {
  if (!m->ch_kind) return;
  m->ch_kind = 0;
  switch (m->sys) {
    case OR_ISING:
      m->S[m->ch_which] = (int8_t)(m->S[m->ch_which] * -1);
      m->E = m->ch_e;
      break;
    case OR_LJ:
      m->x[m->ch_which] = m->ch_x; m->y[m->ch_which] = m->ch_y; m->z[m->ch_which] = m->ch_z;
      if (m->p.refresh > 0) m->E = m->ch_e;
      else lj_set_energy(m, m->ch_e);
      break;
    case OR_WCA:
      m->x[m->ch_which] = m->ch_x; m->y[m->ch_which] = m->ch_y; m->z[m->ch_which] = m->ch_z;
      if (m->p.refresh > 0) m->E = m->ch_e;
      else wca_set_energy(m, m->ch_e, m->ch_dabse);
      break;
    default:
      m->E = m->ch_e;
  }
}

/* lj.rs:249-261, wca.rs:237-250. */
static void sys_verify_energy(or_mc *m) {
  if (m->sys != OR_LJ && m->sys != OR_WCA) return;
  double egood = sys_compute_energy(m, m->G);
  double n = (double)m->n;
  double expected = m->sys == OR_LJ ? fabs(m->E) * 1e-14 * n * n : fabs(m->E) * 1e-13 * n * n;
  if (fabs(egood - m->E) > expected) m->status = ST_VERIFY;
}

/* wca.rs:202-218: pressure every N^2 moves. */
static int sys_pressure(const or_mc *m, double *val) {
  if (m->sys != OR_WCA || !m->p.collect_pressure) return 0;
  uint64_t nn = (uint64_t)m->n * (uint64_t)m->n;
  if (m->moves % nn != 0) return 0;
  double p;
  int n = m->n;
  if (m->G <= 0) {
    p = 0.0;
    for (int i = 0; i < n; i++)
      p = celllist_sum(p, m->x, m->y, m->z, n, i, m->x[i], m->y[i], m->z[i], m->L, m->rc2, m->nsub,
                       wca_pair_pressure, 1.0, NULL);
  } else {
    double acc[64], L = m->L, h = m->halfL;
    for (int l = 0; l < m->G; l++) {
      acc[l] = 0.0;
      for (int i = l; i < n; i += m->G)
        for (int j = 0; j < n; j++)
          if (j != i)
            acc[l] += wca_pair_pressure(norm2(min_image(m->x[i] - m->x[j], L, h), min_image(m->y[i] - m->y[j], L, h),
                                              min_image(m->z[i] - m->z[j], L, h)), m->rc2);
    }
    p = butterfly(acc, m->G);
  }
  *val = p / (3.0 * (m->L * m->L * m->L));
  return 1;
}

/* ===================================================================== */
/* EnergyMC (energy.rs)                                                   */
/* ===================================================================== */

static void bins_reserve(or_mc *m, int64_t need) {
  if (need <= m->cap) return;
  int64_t c = m->cap ? m->cap : 16;
  while (c < need) c *= 2;
#define GROW(ptr, T) ptr = (T *)realloc(ptr, sizeof(T) * c)
  GROW(m->lnw, double); GROW(m->etot, double); GROW(m->esq, double); GROW(m->ptotal, double);
  GROW(m->hist, uint64_t); GROW(m->tfound, uint64_t); GROW(m->rt, uint64_t); GROW(m->pcount, uint64_t);
  GROW(m->vis, uint8_t);
#undef GROW
  m->cap = c;
}

/* Bins::state_to_index (energy.rs:371-373): `as usize` saturates. */
static inline int64_t e_index(const or_mc *m, double e) {
  double v = (e - m->min) / m->width;
  if (!(v > 0.0)) return 0;
  if (v >= 9.2e18) return INT64_MAX;
  return (int64_t)v;
}
static inline double i_energy(const or_mc *m, int64_t i) { return m->min + ((double)i + 0.5) * m->width; }

/* energy.rs:400-434. */
static void prepare_for_state(or_mc *m, double e) {
  while (e < m->min) {
    bins_reserve(m, m->len + 1);
#define INS0(a) memmove(m->a + 1, m->a, sizeof(m->a[0]) * m->len)
    INS0(lnw); INS0(etot); INS0(esq); INS0(ptotal); INS0(hist); INS0(tfound); INS0(rt); INS0(pcount); INS0(vis);
#undef INS0
    m->lnw[0] = 0.0; m->etot[0] = 0.0; m->esq[0] = 0.0; m->ptotal[0] = 0.0;
    m->hist[0] = 0; m->tfound[0] = 0; m->pcount[0] = 0;
    m->vis[0] = 1;
    m->rt[0] = 1;
    m->len++;
    m->min -= m->width;
  }
  while (e >= m->min + m->width * (double)m->len) {
    bins_reserve(m, m->len + 1);
    int64_t k = m->len;
    m->lnw[k] = 0.0; m->hist[k] = 0; m->tfound[k] = 0; m->pcount[k] = 0; m->ptotal[k] = 0.0;
    m->etot[k] = 0.0; m->esq[k] = 0.0; m->vis[k] = 1; m->rt[k] = 1;
    m->len++;
  }
}

/* energy.rs:22-38. */
double or_sad_gamma(double lambda, double t, double tF, double ns) {
  if (lambda * tF * ns == 0.0) return 0.0;
  return (lambda + t / tF) / (lambda + t / ns * (t / tF));
}

/* energy.rs:799-824. */
static double mc_gamma(const or_mc *m) {
  switch (m->method) {
    case OR_CANONICAL: return 0.0;
    case OR_SAD: return or_sad_gamma(m->lambda, (double)m->moves, (double)m->tF, (double)m->Ns);
    case OR_SAMC: {
      double t = (double)m->moves;
      return t > m->t0 ? m->t0 / t : 1.0;
    }
    default: return m->wl_gamma;
  }
}

static int idx_ok(or_mc *m, int64_t i) {
  if (i < 0 || i >= m->len) { m->status = ST_INDEX; return 0; }
  return 1;
}

/* Metropolis-Hastings test on the log acceptance ratio lr = (ln w1 - ln w2)
 * + ln H of the learned move (the reference's tests when ln H = 0). */
static int reject_lr(or_mc *m, double lr) {
  m->lm_lr = lr;
  if (!(lr < 0.0)) return 0;
  m->lm_u = or_gen_f64(m->rng);
  return m->lm_u > exp(lr);
}

/* energy.rs:440-512; lnh = ln H of a learned move, 0 for the cube move. */
static int reject_move(or_mc *m, double e1, double e2, double lnh) {
  int64_t i1 = e_index(m, e1), i2 = e_index(m, e2);
  if (!idx_ok(m, i1) || !idx_ok(m, i2)) return 1;
  if (m->p.move_kind == 1) {
    int rejected;
    switch (m->method) {
      case OR_SAD: {
        int64_t ilo = e_index(m, m->too_lo), ihi = e_index(m, m->too_hi);
        if (!idx_ok(m, ilo) || !idx_ok(m, ihi)) return 1;
        double lnw1 = e1 < m->too_lo ? m->lnw[ilo] + (e1 - m->too_lo) / m->min_T
                      : e1 > m->too_hi ? m->lnw[ihi] : m->lnw[i1];
        double lnw2 = e2 < m->too_lo ? m->lnw[ilo] + (e2 - m->too_lo) / m->min_T
                      : e2 > m->too_hi ? m->lnw[ihi] : m->lnw[i2];
        rejected = reject_lr(m, lnw1 - lnw2 + lnh);
        if (!rejected && m->hist[i2] == 0 && e2 < m->too_hi && e2 > m->too_lo) {
          m->lambda = (m->too_hi - m->too_lo) / m->min_T;
          m->Ns += 1;
          m->tL = m->moves;
        }
        return rejected;
      }
      case OR_SAMC:
        return reject_lr(m, m->lnw[i1] - m->lnw[i2] + lnh);
      case OR_WL:
      case OR_INV_T_WL:
        rejected = reject_lr(m, m->lnw[i1] - m->lnw[i2] + lnh);
        if (!rejected && m->hist[i2] == 0 && m->wl_lowest > 0) m->wl_Ns += 1.0;
        return rejected;
      default:
        return reject_lr(m, (e1 - e2) / m->T + lnh);
    }
  }
  switch (m->method) {
    case OR_SAD: {
      int64_t ilo = e_index(m, m->too_lo), ihi = e_index(m, m->too_hi);
      if (!idx_ok(m, ilo) || !idx_ok(m, ihi)) return 1;
      double lnw1 = e1 < m->too_lo ? m->lnw[ilo] + (e1 - m->too_lo) / m->min_T
                    : e1 > m->too_hi ? m->lnw[ihi] : m->lnw[i1];
      double lnw2 = e2 < m->too_lo ? m->lnw[ilo] + (e2 - m->too_lo) / m->min_T
                    : e2 > m->too_hi ? m->lnw[ihi] : m->lnw[i2];
      int rejected = lnw2 > lnw1 && or_gen_f64(m->rng) > exp(lnw1 - lnw2);
      if (!rejected && m->hist[i2] == 0 && e2 < m->too_hi && e2 > m->too_lo) {
        m->lambda = (m->too_hi - m->too_lo) / m->min_T;
        m->Ns += 1;
        m->tL = m->moves;
      }
      return rejected;
    }
    case OR_SAMC: {
      double lnw1 = m->lnw[i1], lnw2 = m->lnw[i2];
      return lnw2 > lnw1 && or_gen_f64(m->rng) > exp(lnw1 - lnw2);
    }
    case OR_WL:
    case OR_INV_T_WL: {
      double lnw1 = m->lnw[i1], lnw2 = m->lnw[i2];
      int rejected = lnw2 > lnw1 && or_gen_f64(m->rng) > exp(lnw1 - lnw2);
      if (!rejected && m->hist[i2] == 0 && m->wl_lowest > 0) m->wl_Ns += 1.0;
      return rejected;
    }
    default: /* canonical */
      if (e1 >= e2) return 0;
      return or_gen_f64(m->rng) > exp((e1 - e2) / m->T);
  }
}

static void wl_reserve(or_mc *m, int64_t need) {
  if (need <= m->wl_cap) return;
  int64_t c = m->wl_cap ? m->wl_cap : 16;
  while (c < need) c *= 2;
  m->wlh = (uint64_t *)realloc(m->wlh, sizeof(uint64_t) * c);
  m->wl_cap = c;
}

/* energy.rs:514-761. */
static void update_weights(or_mc *m, double E) {
  int64_t i = e_index(m, E);
  if (!idx_ok(m, i)) return;
  double gamma = mc_gamma(m);
  double old_lnw = m->lnw[i];
  m->lnw[i] += gamma;
  int switch_to_samc = 0;
  double samc_t0 = 0.0;
  if (m->method == OR_SAD) {
    if (m->too_lo > m->too_hi || E < m->too_lo || E > m->too_hi) m->lnw[i] = old_lnw;
    if (m->hist[i] > m->highest) {
      m->highest = m->hist[i];
      if (E > m->too_hi) {
        int64_t ihi = e_index(m, m->too_hi);
        if (!idx_ok(m, ihi)) return;
        for (int64_t j = 0; j < m->len; j++) {
          double ej = i_energy(m, j);
          if (ej > m->too_hi && ej <= E) {
            if (m->hist[j] != 0) { m->lnw[j] = m->lnw[ihi]; m->Ns += 1; }
            else m->lnw[j] = 0.0;
          }
        }
        m->lambda = (E - m->too_lo) / m->min_T;
        m->tL = m->moves;
        m->too_hi = i_energy(m, e_index(m, E));
      } else if (E < m->too_lo) {
        int64_t ilo = e_index(m, m->too_lo);
        if (!idx_ok(m, ilo)) return;
        for (int64_t j = 0; j < m->len; j++) {
          double ej = i_energy(m, j);
          if (ej < m->too_lo && ej >= E) {
            if (m->hist[j] != 0) {
              m->lnw[j] = m->lnw[ilo] + (ej - m->too_lo) / m->min_T;
              if (m->lnw[j] < 0.0) m->lnw[j] = 0.0;
              m->Ns += 1;
            } else {
              m->lnw[j] = 0.0;
            }
          }
        }
        m->lambda = (m->too_hi - E) / m->min_T;
        m->tL = m->moves;
        m->too_lo = i_energy(m, e_index(m, E));
      }
    }
    if (m->tL == m->moves) {
      int64_t ilo = e_index(m, m->too_lo), ihi = e_index(m, m->too_hi);
      if (!idx_ok(m, ilo) || !idx_ok(m, ihi)) return;
      uint64_t mx = m->tfound[ilo];
      for (int64_t j = ilo; j <= ihi; j++)
        if (m->tfound[j] > mx) mx = m->tfound[j];
      uint64_t old_tF = m->tF;
      m->tF = mx;
      if (old_tF != m->tF && m->p.move_plan == 1) {
        double s = m->acc_rate / m->p.move_param;
        s = s < 0.8 ? 0.8 : (s > 1.2 ? 1.2 : s);
        m->tscale *= s;
      }
    }
  } else if (m->method == OR_WL || m->method == OR_INV_T_WL) {
    if (m->has_min_gamma && m->wl_gamma < m->min_gamma) {
      if (i >= m->wl_len) { m->status = ST_INDEX; return; }
      m->wlh[i] += 1;
      return;
    }
    if (m->wl_len != m->len) {
      if (m->wl_len == 0 || (m->wl_gamma != 1.0 && m->wl_lowest > 0)) {
        m->wl_gamma = 1.0;
        m->wl_lowest = 0; m->wl_highest = 0; m->wl_total = 0;
        wl_reserve(m, m->len);
        memset(m->wlh, 0, sizeof(uint64_t) * m->len);
        m->wl_len = m->len;
        m->wl_min_energy = m->min;
      } else {
        while (m->wl_min_energy > m->min) {
          m->wl_min_energy -= m->width;
          wl_reserve(m, m->wl_len + 1);
          memmove(m->wlh + 1, m->wlh, sizeof(uint64_t) * m->wl_len);
          m->wlh[0] = 0;
          m->wl_len++;
        }
        while (m->wl_len < m->len) {
          wl_reserve(m, m->wl_len + 1);
          m->wlh[m->wl_len++] = 0;
        }
        uint64_t mn = UINT64_MAX;
        for (int64_t j = 0; j < m->wl_len; j++)
          if (m->wlh[j] < mn) mn = m->wlh[j];
        m->wl_lowest = mn;
      }
    }
    if (i >= m->wl_len) { m->status = ST_INDEX; return; }
    m->wlh[i] += 1;
    if (m->wlh[i] > m->wl_highest) m->wl_highest = m->wlh[i];
    m->wl_total += 1;
    double max_energy = m->wl_min_energy + (double)m->wl_len * m->width;
    if (m->wlh[i] == m->wl_lowest + 1 && m->wl_len > 1 &&
        (!m->p.has_min_e || m->p.min_e >= m->wl_min_energy) &&
        (!m->p.has_max_e || m->p.max_e <= max_energy)) {
      int found = 0;
      uint64_t mn = UINT64_MAX;
      for (int64_t j = 0; j < m->wl_len; j++)
        if (m->hist[j] != 0) { found = 1; if (m->wlh[j] < mn) mn = m->wlh[j]; }
      if (found && mn == m->wl_lowest + 1) {
        m->wl_lowest = m->wlh[i];
        if ((m->wl_inv_t && m->wl_lowest > 0) ||
            (double)m->wl_lowest >= 0.8 * (double)m->wl_total / m->wl_Ns) {
          m->wl_gamma *= 0.5;
          memset(m->wlh, 0, sizeof(uint64_t) * m->wl_len);
          m->wl_total = 0; m->wl_lowest = 0; m->wl_highest = 0;
          if (m->has_min_gamma && m->wl_gamma < m->min_gamma) m->wl_gamma = 0.0;
        }
        if (m->wl_inv_t && m->wl_gamma < m->wl_Ns / (double)m->moves) {
          switch_to_samc = 1;
          samc_t0 = m->wl_Ns;
        }
      }
    }
  }
  if (switch_to_samc) {
    m->method = OR_SAMC;
    m->t0 = samc_t0;
  }
}

/* energy.rs:904-974. */
/* ---- learned proposal --------------------------------------------------
 * No reference counterpart.  The definitions are in sadmc_oracle.h. */
static const float kPropC0 = 0.9f, kPropDC = 1.6f / 7.0f, kPropA = 12.5f; /* 1/(2 w^2), w = 0.2 */
static const float kPropSMin = -4.605170186f, kPropSMax = 0.0f;          /* ln 0.01, ln 1 */

static void prop_features_xyz(int n, const double *x, const double *y, const double *z, int skip, double px,
                              double py, double pz, float *f) {
  for (int k = 0; k < OR_PROP_NF; k++) f[k] = 0.0f;
  for (int j = 0; j < n; j++) {
    if (j == skip) continue;
    float a = (float)(x[j] - px), b = (float)(y[j] - py), c = (float)(z[j] - pz);
    float r = sqrtf(a * a + b * b + c * c);
    for (int k = 0; k < OR_PROP_NF; k++) {
      float d = r - (kPropC0 + (float)k * kPropDC);
      f[k] += expf(-kPropA * d * d);
    }
  }
}

void or_prop_features(int32_t n, const double *c, int32_t skip, double px, double py, double pz, float *f) {
  prop_features_xyz(n, c, c + n, c + 2 * n, skip, px, py, pz, f);
}

/* Forward pass; h1, h2 (32 each) and the unclamped output may be returned. */
static float prop_forward(const float *w, const float *f, float *h1, float *h2, float *raw) {
  const float *W1 = w, *b1 = w + 256, *W2 = w + 288, *b2 = w + 1312, *w3 = w + 1344, *b3 = w + 1376;
  float a1[OR_PROP_H], a2[OR_PROP_H];
  for (int u = 0; u < OR_PROP_H; u++) {
    float z = b1[u];
    for (int k = 0; k < OR_PROP_NF; k++) z += W1[u * OR_PROP_NF + k] * f[k];
    a1[u] = z > 0.0f ? z : 0.0f;
  }
  for (int u = 0; u < OR_PROP_H; u++) {
    float z = b2[u];
    for (int k = 0; k < OR_PROP_H; k++) z += W2[u * OR_PROP_H + k] * a1[k];
    a2[u] = z > 0.0f ? z : 0.0f;
  }
  float s = b3[0];
  for (int u = 0; u < OR_PROP_H; u++) s += w3[u] * a2[u];
  if (h1) memcpy(h1, a1, sizeof a1);
  if (h2) memcpy(h2, a2, sizeof a2);
  if (raw) *raw = s;
  return s < kPropSMin ? kPropSMin : (s > kPropSMax ? kPropSMax : s);
}

float or_prop_mlp(const float *w, const float *f) { return prop_forward(w, f, NULL, NULL, NULL); }

void or_prop_gaussian(uint64_t rng[2], float *xi) {
  float u1 = (float)or_gen_f64(rng), u2 = (float)or_gen_f64(rng);
  float u3 = (float)or_gen_f64(rng), u4 = (float)or_gen_f64(rng);
  const float two_pi = 6.283185307f;
  float r1 = sqrtf(-2.0f * logf(1.0f - u1)), r2 = sqrtf(-2.0f * logf(1.0f - u3));
  xi[0] = r1 * cosf(two_pi * u2);
  xi[1] = r1 * sinf(two_pi * u2);
  xi[2] = r2 * cosf(two_pi * u4);
}

double or_prop_lnh(float s, float s_new, float xi2) {
  double ds = (double)s - (double)s_new;
  return 3.0 * ds + 0.5 * (double)xi2 * (1.0 - exp(2.0 * ds));
}

/* Sum over the other atoms of the pair-term difference for atom w moving
 * from f to t, in the walker's summation order. */
static double lj_delta(or_mc *m, int64_t w, double tx, double ty, double tz, double fx, double fy, double fz) {
  if (m->G <= 0) {
    double e = 0.0;
    for (int j = 0; j < m->n; j++)
      if (j != w) e += lj_term(m, j, tx, ty, tz, fx, fy, fz);
    return e;
  }
  double acc[64];
  for (int l = 0; l < m->G; l++) {
    acc[l] = 0.0;
    for (int j = l; j < m->n; j += m->G)
      if (j != w) acc[l] += lj_term(m, j, tx, ty, tz, fx, fy, fz);
  }
  return butterfly(acc, m->G);
}

static int lj_plan_learned(or_mc *m, double *e_out, double *lnh) {
  int64_t w = (int64_t)or_uniform_index(m->rng, (uint64_t)m->n);
  double fx = m->x[w], fy = m->y[w], fz = m->z[w];
  float f[OR_PROP_NF], xi[3];
  prop_features_xyz(m->n, m->x, m->y, m->z, (int)w, fx, fy, fz, f);
  float s = prop_forward(m->p.mlp, f, NULL, NULL, NULL);
  or_prop_gaussian(m->rng, xi);
  float sig = expf(s);
  double tx = fx + (double)(sig * xi[0]), ty = fy + (double)(sig * xi[1]), tz = fz + (double)(sig * xi[2]);
  double r2to = norm2(tx, ty, tz), r2from = norm2(fx, fy, fz);
  m->lm_wall = r2to - m->R2;
  if (r2to > m->R2 && r2to > r2from) return 0;
  prop_features_xyz(m->n, m->x, m->y, m->z, (int)w, tx, ty, tz, f);
  float s2 = prop_forward(m->p.mlp, f, NULL, NULL, NULL);
  double e = m->E + lj_delta(m, w, tx, ty, tz, fx, fy, fz);
  *lnh = m->p.drop_hastings ? 0.0 : or_prop_lnh(s, s2, xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
  m->ch_kind = 1; m->ch_which = w; m->ch_x = tx; m->ch_y = ty; m->ch_z = tz; m->ch_e = e; m->ch_lnh = *lnh;
  *e_out = e;
  return 1;
}

int or_mc_plan_learned(or_mc *m, double *e_new, double *lnh, double *to) {
  if (!lj_plan_learned(m, e_new, lnh)) return 0;
  to[0] = m->ch_x; to[1] = m->ch_y; to[2] = m->ch_z;
  return 1;
}

/* ln w(E) of a SAD walker with the clamped extension (energy.rs:445-463);
 * returns 0 if E has no bin. */
static int sad_lnw(const or_mc *m, double e, double *out) {
  int64_t ilo = e_index(m, m->too_lo), ihi = e_index(m, m->too_hi), i = e_index(m, e);
  if (ilo < 0 || ilo >= m->len || ihi < 0 || ihi >= m->len) return 0;
  if (e < m->too_lo) { *out = m->lnw[ilo] + (e - m->too_lo) / m->min_T; return 1; }
  if (e > m->too_hi) { *out = m->lnw[ihi]; return 1; }
  if (i < 0 || i >= m->len) return 0;
  *out = m->lnw[i];
  return 1;
}

void or_prop_probe(const or_mc *mc, uint64_t seed, int32_t count, float *out) {
  or_mc *m = (or_mc *)mc; /* lj_delta reads only */
  uint64_t rng[2];
  or_rng_seed(seed, rng);
  double lnw1;
  int have1 = sad_lnw(m, m->E, &lnw1);
  for (int32_t c = 0; c < count; c++) {
    float *rec = out + (int64_t)c * OR_PROP_REC;
    int64_t w = (int64_t)or_uniform_index(rng, (uint64_t)m->n);
    double fx = m->x[w], fy = m->y[w], fz = m->z[w];
    prop_features_xyz(m->n, m->x, m->y, m->z, (int)w, fx, fy, fz, rec);
    for (int k = 0; k < OR_PROP_NS; k++) {
      float sig = expf(kPropSMin + (float)k * ((kPropSMax - kPropSMin) / (float)(OR_PROP_NS - 1)));
      float jsum = 0.0f;
      for (int r = 0; r < OR_PROP_R; r++) {
        float xi[3];
        or_prop_gaussian(rng, xi);
        float dx = sig * xi[0], dy = sig * xi[1], dz = sig * xi[2];
        double tx = fx + (double)dx, ty = fy + (double)dy, tz = fz + (double)dz;
        double r2to = norm2(tx, ty, tz), r2from = norm2(fx, fy, fz);
        if (r2to > m->R2 && r2to > r2from) continue;
        double e2 = m->E + lj_delta(m, w, tx, ty, tz, fx, fy, fz);
        if (m->p.has_max_e && e2 > m->p.max_e && e2 > m->E) continue;
        double lnw2;
        if (!have1 || !sad_lnw(m, e2, &lnw2)) continue;
        double lr = lnw1 - lnw2;
        float a = lr >= 0.0 ? 1.0f : (float)exp(lr);
        jsum += a * (dx * dx + dy * dy + dz * dz);
      }
      rec[OR_PROP_NF + k] = jsum / (float)OR_PROP_R;
    }
  }
}

/* Jhat(s) by linear interpolation on the scale grid, its slope, and the
 * normaliser max_m J_m (0 if every J_m is 0). */
static void prop_target(const float *J, double s, double *val, double *slope, double *jmax) {
  const double dS = ((double)kPropSMax - (double)kPropSMin) / (double)(OR_PROP_NS - 1);
  double mx = 0.0;
  for (int k = 0; k < OR_PROP_NS; k++) mx = J[k] > mx ? J[k] : mx;
  *jmax = mx;
  double q = (s - (double)kPropSMin) / dS;
  int k = (int)floor(q);
  if (k < 0) k = 0;
  if (k > OR_PROP_NS - 2) k = OR_PROP_NS - 2;
  double t = q - (double)k;
  *val = (double)J[k] * (1.0 - t) + (double)J[k + 1] * t;
  *slope = ((double)J[k + 1] - (double)J[k]) / dS;
}

/* d(loss)/d(raw output) for one record before normalisation: -slope inside
 * the clamp range; at a clamp, -slope only when it points back inside (a
 * projected gradient, so a clamped output is not stuck). */
static double prop_dloss(double raw, double slope) {
  if (raw > kPropSMin && raw < kPropSMax) return -slope;
  if ((raw >= kPropSMax && slope < 0.0) || (raw <= kPropSMin && slope > 0.0)) return -slope;
  return 0.0;
}

double or_prop_loss_grad(const float *w, const float *records, int64_t count, double *grad) {
  const float *W2 = w + 288, *w3 = w + 1344;
  if (grad) memset(grad, 0, sizeof(double) * OR_PROP_NP);
  double loss = 0.0;
  for (int64_t n = 0; n < count; n++) {
    const float *rec = records + n * OR_PROP_REC;
    float h1[OR_PROP_H], h2[OR_PROP_H], raw;
    prop_forward(w, rec, h1, h2, &raw);
    double val, slope, mx;
    double s = raw < kPropSMin ? kPropSMin : (raw > kPropSMax ? kPropSMax : raw);
    prop_target(rec + OR_PROP_NF, s, &val, &slope, &mx);
    if (mx <= 0.0) continue;
    loss -= val / mx;
    if (!grad) continue;
    double d3 = prop_dloss(raw, slope) / mx / (double)count;
    if (d3 == 0.0) continue;
    double d2[OR_PROP_H];
    grad[1376] += d3;
    for (int u = 0; u < OR_PROP_H; u++) {
      grad[1344 + u] += d3 * h2[u];
      d2[u] = h2[u] > 0.0f ? d3 * w3[u] : 0.0;
    }
    for (int u = 0; u < OR_PROP_H; u++) {
      grad[1312 + u] += d2[u];
      for (int k = 0; k < OR_PROP_H; k++) grad[288 + u * OR_PROP_H + k] += d2[u] * h1[k];
    }
    for (int k = 0; k < OR_PROP_H; k++) {
      if (!(h1[k] > 0.0f)) continue;
      double d1 = 0.0;
      for (int u = 0; u < OR_PROP_H; u++) d1 += d2[u] * W2[u * OR_PROP_H + k];
      grad[256 + k] += d1;
      for (int j = 0; j < OR_PROP_NF; j++) grad[k * OR_PROP_NF + j] += d1 * rec[j];
    }
  }
  return loss / (double)count;
}

double or_prop_loss_grad_f64(const double *w, const float *records, int64_t count, double *grad) {
  const double *W1 = w, *b1 = w + 256, *W2 = w + 288, *b2 = w + 1312, *w3 = w + 1344, *b3 = w + 1376;
  if (grad) memset(grad, 0, sizeof(double) * OR_PROP_NP);
  double loss = 0.0;
  for (int64_t n = 0; n < count; n++) {
    const float *rec = records + n * OR_PROP_REC;
    double h1[OR_PROP_H], h2[OR_PROP_H];
    for (int u = 0; u < OR_PROP_H; u++) {
      double z = b1[u];
      for (int k = 0; k < OR_PROP_NF; k++) z += W1[u * OR_PROP_NF + k] * (double)rec[k];
      h1[u] = z > 0.0 ? z : 0.0;
    }
    for (int u = 0; u < OR_PROP_H; u++) {
      double z = b2[u];
      for (int k = 0; k < OR_PROP_H; k++) z += W2[u * OR_PROP_H + k] * h1[k];
      h2[u] = z > 0.0 ? z : 0.0;
    }
    double raw = b3[0];
    for (int u = 0; u < OR_PROP_H; u++) raw += w3[u] * h2[u];
    double s = raw < kPropSMin ? kPropSMin : (raw > kPropSMax ? kPropSMax : raw);
    double val, slope, mx;
    prop_target(rec + OR_PROP_NF, s, &val, &slope, &mx);
    if (mx <= 0.0) continue;
    loss -= val / mx;
    if (!grad) continue;
    double d3 = prop_dloss(raw, slope) / mx / (double)count;
    if (d3 == 0.0) continue;
    double d2[OR_PROP_H];
    grad[1376] += d3;
    for (int u = 0; u < OR_PROP_H; u++) {
      grad[1344 + u] += d3 * h2[u];
      d2[u] = h2[u] > 0.0 ? d3 * w3[u] : 0.0;
    }
    for (int u = 0; u < OR_PROP_H; u++) {
      grad[1312 + u] += d2[u];
      for (int k = 0; k < OR_PROP_H; k++) grad[288 + u * OR_PROP_H + k] += d2[u] * h1[k];
    }
    for (int k = 0; k < OR_PROP_H; k++) {
      if (!(h1[k] > 0.0)) continue;
      double d1 = 0.0;
      for (int u = 0; u < OR_PROP_H; u++) d1 += d2[u] * W2[u * OR_PROP_H + k];
      grad[256 + k] += d1;
      for (int j = 0; j < OR_PROP_NF; j++) grad[k * OR_PROP_NF + j] += d1 * (double)rec[j];
    }
  }
  return loss / (double)count;
}

static void move_once(or_mc *m) {
  m->moves += 1;
  if (m->sys == OR_LJ || m->sys == OR_WCA) {
    if (m->p.refresh > 0) {
      /* the FP64 refresh of the FP32-pair energy: its drift can carry the
         energy past the edge of the bins, which then grow to it as for a
         proposed energy (energy.rs:400-434) */
      if (m->moves % (uint64_t)m->p.refresh == 0) {
        m->E = sys_compute_energy(m, m->G);
        prepare_for_state(m, m->E);
      }
    } else {
      uint64_t L = (uint64_t)m->len;
      if (m->moves % (L * L * 1000ULL) == 0) sys_verify_energy(m);
    }
  }
  double e1 = m->E;
  double rs = sqrt(1.0 / (double)m->moves);
  m->acc_rate *= 1.0 - rs;
  double e2;
  int acc = 0;
  double lnh = 0.0;
  m->lm_e2 = m->lm_lr = m->lm_u = m->lm_wall = NAN;
  int some = m->p.move_kind == 1 && m->sys == OR_LJ ? lj_plan_learned(m, &e2, &lnh) : sys_plan_move(m, m->tscale, &e2);
  if (some) m->lm_e2 = e2;
  if (some) {
    int oob = 0;
    if (m->p.has_max_e) oob = e2 > m->p.max_e && e2 > e1;
    if (m->p.has_min_e) oob = oob || (e2 < m->p.min_e && e2 < e1);
    if (!oob) {
      prepare_for_state(m, e2);
      if (!reject_move(m, e1, e2, lnh)) {
        m->accepted += 1;
        m->acc_rate += rs;
        sys_confirm(m);
        acc = 1;
      }
    }
  }
  if (m->status) return;
  double E = m->E;
  int64_t i = e_index(m, E);
  if (!idx_ok(m, i)) return;
  if (m->hist[i] == 0) m->tfound[i] = m->moves;
  m->hist[i] += 1;
  m->etot[i] += E;
  m->esq[i] += E * E;
  double pv;
  if (sys_pressure(m, &pv)) {
    m->has_pressure = 1;
    m->pcount[i] += 1;
    m->ptotal[i] += pv;
  }
  update_weights(m, E);
  if (m->status) return;
  if (m->lnw[i] > m->max_S) {
    m->max_S = m->lnw[i];
    m->max_S_index = i;
    memset(m->vis, 1, (size_t)m->len);
  } else if (i == m->max_S_index) {
    if (e_index(m, e1) != i) memset(m->vis, 0, (size_t)m->len);
  } else if (!m->vis[i]) {
    m->vis[i] = 1;
    m->rt[i] += 1;
  }
  if (m->tr_n < m->tr_cap) {
    m->tr_e[m->tr_n] = m->E;
    m->tr_a[m->tr_n] = (uint8_t)acc;
    m->tr_n++;
  }
}

static or_mc *mc_alloc(const or_params *p) {
  or_mc *m = (or_mc *)calloc(1, sizeof(or_mc));
  m->p = *p;
  m->sys = p->system;
  m->n = p->n;
  m->G = p->order_g;
  return m;
}

/* energy.rs:830-898 (from_params) after the system exists. */
static void mc_finish_init(or_mc *m) {
  const or_params *p = &m->p;
  double w = p->energy_bin > 0 ? p->energy_bin : (m->sys == OR_ISING ? 4.0 : 1.0);
  or_rng_seed(p->seed, m->rng);
  if (p->has_max_e) {
    for (uint64_t k = 0; k < 100000000ULL; k++) {
      double newe;
      if (sys_plan_move(m, 0.05, &newe)) {
        if (newe < m->E) sys_confirm(m);
        if (m->E < p->max_e) break;
      }
    }
  }
  double E = m->E;
  double emin = (round(E / w) - 0.5) * w;
  m->width = w;
  m->min = emin;
  m->len = 1;
  bins_reserve(m, 16);
  m->hist[0] = 1; m->tfound[0] = 0; m->lnw[0] = 0.0; m->etot[0] = E; m->esq[0] = E * E;
  m->pcount[0] = 0; m->ptotal[0] = 0.0;
  m->vis[0] = 0; m->rt[0] = 1;
  m->max_S = 0.0; m->max_S_index = 0;
  m->moves = 0; m->accepted = 0; m->acc_rate = 0.5;
  m->tscale = p->move_plan == 0 ? p->move_param : 0.05;
  m->method = p->method;
  switch (p->method) {
    case OR_SAD:
      m->min_T = p->method_param; m->too_lo = E; m->too_hi = E; m->tL = 0; m->tF = 0;
      m->Ns = 1; m->highest = 1; m->lambda = 0.0;
      break;
    case OR_SAMC: m->t0 = p->method_param; break;
    case OR_WL:
    case OR_INV_T_WL:
      m->wl_gamma = 1.0;
      m->wl_lowest = (p->has_min_e && p->has_max_e) ? 0 : 1;
      m->wl_highest = 1; m->wl_total = 0;
      m->wl_Ns = (p->has_min_e && p->has_max_e) ? (p->max_e - p->min_e) / w : 1.0;
      m->wl_len = 0;
      m->wl_min_energy = E;
      m->wl_inv_t = p->method == OR_INV_T_WL;
      m->has_min_gamma = p->method == OR_WL && p->has_min_gamma;
      m->min_gamma = p->method_param;
      break;
    default: m->T = p->method_param;
  }
}

or_mc *or_mc_new(const or_params *p, const double *coords, const int8_t *spins) {
  or_mc *m = mc_alloc(p);
  int n = p->n;
  if (m->sys == OR_ISING) {
    m->S = (int8_t *)malloc((size_t)n * n);
    if (spins) memcpy(m->S, spins, (size_t)n * n);
    else or_ising_init_spins(n, p->ising_init, p->seed, m->S);
    m->E = or_ising_energy(n, m->S);
  } else {
    double *c = (double *)malloc(sizeof(double) * 3 * n);
    int rc = 0;
    if (coords) memcpy(c, coords, sizeof(double) * 3 * n);
    else if (m->sys == OR_LJ) rc = or_lj_init(n, p->radius, p->seed, c);
    else rc = or_wca_init(n, p->box, p->seed, c);
    m->x = c; m->y = c + n; m->z = c + 2 * n;
    m->R2 = p->radius * p->radius;
    m->L = p->box; m->halfL = 0.5 * p->box; m->rc2 = p->rc2;
    if (m->sys == OR_WCA) m->nsub = wca_nsub(m->L, m->rc2);
    if (rc) m->status = ST_INIT;
    m->E = sys_compute_energy(m, m->G);
  }
  mc_finish_init(m);
  return m;
}

or_mc *or_mc_new_scripted(const or_params *p, double e0, const double *script, int64_t len) {
  or_mc *m = mc_alloc(p);
  m->sys = OR_SCRIPTED;
  m->script = (double *)malloc(sizeof(double) * len);
  memcpy(m->script, script, sizeof(double) * len);
  m->script_len = len;
  m->E = e0;
  mc_finish_init(m);
  return m;
}

void or_mc_free(or_mc *m) {
  if (!m) return;
  free(m->x); free(m->S); free(m->script);
  free(m->lnw); free(m->etot); free(m->esq); free(m->ptotal);
  free(m->hist); free(m->tfound); free(m->rt); free(m->pcount); free(m->vis); free(m->wlh);
  free(m->tr_e); free(m->tr_a);
  free(m);
}

void or_mc_trace(or_mc *m, uint64_t cap) {
  free(m->tr_e); free(m->tr_a);
  m->tr_e = (double *)malloc(sizeof(double) * (cap ? cap : 1));
  m->tr_a = (uint8_t *)malloc(cap ? cap : 1);
  m->tr_cap = cap; m->tr_n = 0;
}

int or_mc_run(or_mc *m, uint64_t moves) {
  for (uint64_t k = 0; k < moves && !m->status; k++) move_once(m);
  return m->status;
}

void or_mc_scalars(const or_mc *m, or_scalars *o) {
  memset(o, 0, sizeof(*o));
  o->E = m->E; o->error = m->error; o->moves = m->moves; o->accepted = m->accepted;
  o->acceptance_rate = m->acc_rate; o->translation_scale = m->tscale;
  o->rng[0] = m->rng[0]; o->rng[1] = m->rng[1];
  o->bins_min = m->min; o->bins_width = m->width; o->len = m->len;
  o->max_S = m->max_S; o->max_S_index = m->max_S_index;
  o->method = m->method; o->status = m->status;
  o->too_lo = m->too_lo; o->too_hi = m->too_hi; o->lambda = m->lambda; o->min_T = m->min_T;
  o->tL = m->tL; o->tF = m->tF; o->num_states = m->Ns; o->highest_hist = m->highest;
  o->t0 = m->t0;
  o->wl_gamma = m->wl_gamma; o->wl_num_states = m->wl_Ns; o->wl_min_energy = m->wl_min_energy;
  o->wl_lowest = m->wl_lowest; o->wl_highest = m->wl_highest; o->wl_total = m->wl_total;
  o->wl_len = m->wl_len; o->has_pressure = m->has_pressure;
}

void or_mc_bins(const or_mc *m, double *lnw, uint64_t *hist, uint64_t *t_found, double *e_tot, double *e_sq,
                uint8_t *visited, uint64_t *round_trips, uint64_t *wl_hist, uint64_t *p_count, double *p_total) {
  size_t n = (size_t)m->len;
  if (lnw) memcpy(lnw, m->lnw, 8 * n);
  if (hist) memcpy(hist, m->hist, 8 * n);
  if (t_found) memcpy(t_found, m->tfound, 8 * n);
  if (e_tot) memcpy(e_tot, m->etot, 8 * n);
  if (e_sq) memcpy(e_sq, m->esq, 8 * n);
  if (visited) memcpy(visited, m->vis, n);
  if (round_trips) memcpy(round_trips, m->rt, 8 * n);
  if (wl_hist) {
    memset(wl_hist, 0, 8 * n);
    size_t k = (size_t)m->wl_len < n ? (size_t)m->wl_len : n;
    if (k) memcpy(wl_hist, m->wlh, 8 * k);
  }
  if (p_count) memcpy(p_count, m->pcount, 8 * n);
  if (p_total) memcpy(p_total, m->ptotal, 8 * n);
}

void or_mc_coords(const or_mc *m, double *c) {
  if (m->sys != OR_LJ && m->sys != OR_WCA) return;
  memcpy(c, m->x, sizeof(double) * m->n);
  memcpy(c + m->n, m->y, sizeof(double) * m->n);
  memcpy(c + 2 * m->n, m->z, sizeof(double) * m->n);
}

void or_mc_spins(const or_mc *m, int8_t *s) {
  if (m->sys == OR_ISING) memcpy(s, m->S, (size_t)m->n * m->n);
}

double or_mc_compute_energy(const or_mc *m) { return sys_compute_energy(m, m->G); }

/* MovableSystem::plan_move / ConfirmSystem::confirm on the walker's own
 * system and RNG (system/mod.rs:101-120). */
int or_mc_plan_move(or_mc *m, double scale, double *e) { return sys_plan_move(m, scale, e); }
void or_mc_confirm(or_mc *m) { sys_confirm(m); }
double or_mc_energy(const or_mc *m) { return m->E; }
double or_mc_reference_energy(const or_mc *m) { return sys_compute_energy(m, 0); }

/* Trace accessors live on the struct; expose copies. */
int or_mc_load_state(or_mc *m, const or_scalars *s, const double *lnw, const uint64_t *hist,
                     const uint64_t *t_found, const double *e_tot, const double *e_sq, const uint8_t *visited,
                     const uint64_t *round_trips, const double *coords) {
  if (m->sys != OR_LJ && m->sys != OR_WCA) return -1;
  if (s->method != OR_SAD || m->method != OR_SAD) return -1;
  m->E = s->E;
  m->error = s->error;
  m->moves = s->moves;
  m->accepted = s->accepted;
  m->acc_rate = s->acceptance_rate;
  m->tscale = s->translation_scale;
  m->rng[0] = s->rng[0];
  m->rng[1] = s->rng[1];
  m->min = s->bins_min;
  m->width = s->bins_width;
  bins_reserve(m, s->len);
  m->len = s->len;
  const size_t n = (size_t)s->len;
  memcpy(m->lnw, lnw, sizeof(double) * n);
  memcpy(m->hist, hist, sizeof(uint64_t) * n);
  memcpy(m->tfound, t_found, sizeof(uint64_t) * n);
  memcpy(m->etot, e_tot, sizeof(double) * n);
  memcpy(m->esq, e_sq, sizeof(double) * n);
  memcpy(m->vis, visited, n);
  memcpy(m->rt, round_trips, sizeof(uint64_t) * n);
  memset(m->pcount, 0, sizeof(uint64_t) * n);
  memset(m->ptotal, 0, sizeof(double) * n);
  m->max_S = s->max_S;
  m->max_S_index = s->max_S_index;
  m->too_lo = s->too_lo;
  m->too_hi = s->too_hi;
  m->lambda = s->lambda;
  m->min_T = s->min_T;
  m->tL = s->tL;
  m->tF = s->tF;
  m->Ns = s->num_states;
  m->highest = s->highest_hist;
  m->status = s->status;
  memcpy(m->x, coords, sizeof(double) * 3 * (size_t)m->n);
  m->ch_kind = 0;
  return 0;
}

void or_mc_last_move(const or_mc *m, double *out) {
  out[0] = m->lm_e2;
  out[1] = m->lm_lr;
  out[2] = m->lm_u;
  out[3] = m->lm_wall;
}

uint64_t or_mc_trace_copy(const or_mc *m, double *energies, uint8_t *accepted) {
  if (energies) memcpy(energies, m->tr_e, sizeof(double) * m->tr_n);
  if (accepted) memcpy(accepted, m->tr_a, m->tr_n);
  return m->tr_n;
}

/* ===================================================================== */
/* CPU baseline: one walker per thread                                   */
/* ===================================================================== */

typedef struct {
  or_params p;
  uint64_t moves;
  pthread_barrier_t *bar;
  or_mc *mc;
} bench_arg;

static void *bench_thread(void *v) {
  bench_arg *a = (bench_arg *)v;
  a->mc = or_mc_new(&a->p, NULL, NULL);
  pthread_barrier_wait(a->bar);
  or_mc_run(a->mc, a->moves);
  pthread_barrier_wait(a->bar);
  return NULL;
}

/* ---- parallel tempering (tempering.rs) ------------------------------- */

/* The xoroshiro128+ jump of 2^64 steps (the published jump polynomial). */
void or_rng_jump(uint64_t s[2]) {
  static const uint64_t J[2] = {0xDF900294D8F554A5ULL, 0x170865DF4B3201FCULL};
  uint64_t a = 0, b = 0;
  for (int w = 0; w < 2; w++)
    for (int k = 0; k < 64; k++) {
      if ((J[w] >> k) & 1ULL) {
        a ^= s[0];
        b ^= s[1];
      }
      or_rng_next(s);
    }
  s[0] = a;
  s[1] = b;
}

struct or_pt {
  int R, steps;
  or_mc **rep;
  or_replica *st;
  uint64_t rng[2];
  double scale;
};

static uint64_t sys_moves_to_randomize(const or_mc *m) {
  return m->sys == OR_ISING ? (uint64_t)m->n * (uint64_t)m->n : (uint64_t)m->n;
}

or_pt *or_pt_new(const or_params *p, int32_t R, const double *T, uint64_t seed, const double *coords,
                 const int8_t *spins, int32_t canonical_steps) {
  or_pt *pt = (or_pt *)calloc(1, sizeof(or_pt));
  pt->R = R;
  pt->rep = (or_mc **)calloc((size_t)R, sizeof(or_mc *));
  pt->st = (or_replica *)calloc((size_t)R, sizeof(or_replica));
  or_params q = *p;
  q.seed = seed;
  q.has_max_e = 0;
  q.has_min_e = 0;
  q.method = OR_CANONICAL;
  q.method_param = 1.0;
  for (int r = 0; r < R; r++) {
    pt->rep[r] = or_mc_new(&q, coords, spins);  /* the generator is seed_from_u64(seed) */
    pt->st[r].T = T[r];
  }
  pt->steps = (int)(sys_moves_to_randomize(pt->rep[0]) * (uint64_t)canonical_steps);
  pt->scale = p->move_param;
  or_rng_seed(seed, pt->rng);
  or_rng_jump(pt->rng);
  return pt;
}

static void pt_collect(or_replica *s, double e) {
  s->total_energy += e;
  s->total_energy_squared += e * e;
  if (e >= 0.0) s->ignored += 1;
}

static void pt_swap_systems(or_mc *a, or_mc *b) {
  double *t;
  int8_t *u;
  double E = a->E, err = a->error;
  a->E = b->E; a->error = b->error; b->E = E; b->error = err;
  t = a->x; a->x = b->x; b->x = t;
  t = a->y; a->y = b->y; b->y = t;
  t = a->z; a->z = b->z; b->z = t;
  u = a->S; a->S = b->S; b->S = u;
}

void or_pt_run(or_pt *pt, uint64_t rounds) {
  for (uint64_t k = 0; k < rounds; k++) {
    for (int r = 0; r < pt->R; r++) {  /* tempering.rs:96-112 */
      or_mc *m = pt->rep[r];
      or_replica *s = &pt->st[r];
      for (int i = 0; i < pt->steps; i++) {
        double e;
        if (!sys_plan_move(m, pt->scale, &e)) continue;
        double bde = (e - m->E) / s->T;
        if (bde < 0.0 || or_gen_f64(m->rng) < exp(-bde)) {
          sys_confirm(m);
          s->accepted += 1;
        } else {
          s->rejected += 1;
        }
        pt_collect(s, m->E);
      }
    }
    /* tempering.rs:288-318: rand 0.7's bool is the top bit of next_u32, and
     * rand_xoshiro's next_u32 the high half of next_u64 */
    int first = (or_rng_next(pt->rng) >> 63) ? 0 : 1;
    for (int i = first; i + 1 < pt->R; i += 2) {
      or_mc *m0 = pt->rep[i], *m1 = pt->rep[i + 1];
      or_replica *s0 = &pt->st[i], *s1 = &pt->st[i + 1];
      double de_db = (m0->E - m1->E) * (1.0 / s0->T - 1.0 / s1->T);
      if (de_db >= 0.0 || or_gen_f64(m1->rng) < exp(de_db)) {
        s0->accepted_swaps += 1;
        s1->accepted_swaps += 1;
        pt_swap_systems(m0, m1);
      } else {
        s0->rejected_swaps += 1;
        s1->rejected_swaps += 1;
      }
      pt_collect(s0, m0->E);
      pt_collect(s1, m1->E);
    }
  }
}

void or_pt_replica(const or_pt *pt, int32_t r, or_replica *out) {
  *out = pt->st[r];
  out->E = pt->rep[r]->E;
  out->error = pt->rep[r]->error;
  out->rng[0] = pt->rep[r]->rng[0];
  out->rng[1] = pt->rep[r]->rng[1];
}

void or_pt_coords(const or_pt *pt, int32_t r, double *c) {
  const or_mc *m = pt->rep[r];
  if (m->sys == OR_ISING) return;
  memcpy(c, m->x, sizeof(double) * m->n);
  memcpy(c + m->n, m->y, sizeof(double) * m->n);
  memcpy(c + 2 * m->n, m->z, sizeof(double) * m->n);
}

void or_pt_spins(const or_pt *pt, int32_t r, int8_t *s) { or_mc_spins(pt->rep[r], s); }

void or_pt_rng(const or_pt *pt, uint64_t out[2]) {
  out[0] = pt->rng[0];
  out[1] = pt->rng[1];
}

void or_pt_free(or_pt *pt) {
  if (!pt) return;
  for (int r = 0; r < pt->R; r++) or_mc_free(pt->rep[r]);
  free(pt->rep);
  free(pt->st);
  free(pt);
}

double or_bench(const or_params *p, int32_t threads, uint64_t moves, uint64_t seed0) {
  pthread_t *t = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  bench_arg *a = (bench_arg *)malloc(sizeof(bench_arg) * threads);
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, (unsigned)threads + 1);
  for (int i = 0; i < threads; i++) {
    a[i].p = *p;
    a[i].p.seed = seed0 + (uint64_t)i;
    a[i].moves = moves;
    a[i].bar = &bar;
    pthread_create(&t[i], NULL, bench_thread, &a[i]);
  }
  pthread_barrier_wait(&bar);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  pthread_barrier_wait(&bar);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  for (int i = 0; i < threads; i++) {
    pthread_join(t[i], NULL);
    or_mc_free(a[i].mc);
  }
  pthread_barrier_destroy(&bar);
  free(t); free(a);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ===================================================================== */
/* EnergyMC over a histogram Binning (energy_binning.rs, binning/histogram.rs) */
/* ===================================================================== */

/* BinCounts (binning/histogram.rs:11-33): one array of totals and counts with
 * the running statistics the reference keeps beside them. */
typedef struct {
  double *total;
  uint64_t *count;
  double *fc;  /* linear.rs:24-31: fractional counts (linear binning) */
  double min_total, max_total, e_max_total, e_max_count;
  uint64_t min_count, max_count, total_count;
  double fmin_count, fmax_count;
  int exists;
} bcounts;

struct or_bmc {
  or_mc *m;  /* the system, its generator and energy */
  or_params bp;
  /* histogram::Bins (binning/histogram.rs:98-111) */
  double min, width, min_e, max_e;
  int64_t len, cap;
  bcounts lnw, en, tf, hist, pr;
  /* EnergyMC (energy_binning.rs:77-100) */
  uint64_t moves, accepted;
  double acc_rate, tscale;
  int method;
  uint64_t num_states, tL;
  double min_T, too_lo, too_hi, tF, latest;
  double t0;
  double gamma;
  int inv_t, has_min_gamma;
  double min_gamma;
  int status;
  int linear;  /* binning/linear.rs instead of binning/histogram.rs */
  /* EnergyMC::high_resolution (energy_binning.rs:125): a histogram::Bins
   * of its own width whose ln w BinCounts is incremented with gamma 0 */
  int has_hr;
  double hr_min, hr_width, hr_min_e, hr_max_e;
  int64_t hr_len, hr_cap;
  double *hr_total;
  uint64_t *hr_count;
  double hr_min_total, hr_max_total, hr_e_max_total, hr_e_max_count;
  uint64_t hr_min_count, hr_max_count, hr_total_count;
};

/* f64::max / f64::min as folded from NaN (binning/histogram.rs:366-372):
 * a NaN accumulator takes the other value. */
static inline double fold_max(double r, double x) { return (isnan(r) || x > r) ? x : r; }
static inline double fold_min(double r, double x) { return (isnan(r) || x < r) ? x : r; }

static void bc_new(bcounts *b, int64_t sz, int64_t cap) {
  b->total = (double *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(double));
  b->count = (uint64_t *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(uint64_t));
  b->fc = (double *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(double));
  (void)sz;
  b->fmin_count = 0.0;
  b->fmax_count = 0.0;
  b->min_total = 0.0;
  b->max_total = 0.0;
  b->e_max_total = -INFINITY;
  b->e_max_count = -INFINITY;
  b->min_count = 0;
  b->max_count = 0;
  b->total_count = 0;
  b->exists = 1;
}

static void bc_free(bcounts *b) {
  free(b->total);
  free(b->count);
  free(b->fc);
}

static double bc_max_of(const bcounts *b, int64_t len) {
  double r = NAN;
  for (int64_t i = 0; i < len; i++) r = fold_max(r, b->total[i]);
  return r;
}
static double bc_min_of(const bcounts *b, int64_t len) {
  double r = NAN;
  for (int64_t i = 0; i < len; i++) r = fold_min(r, b->total[i]);
  return r;
}
static uint64_t bc_min_count(const bcounts *b, int64_t len) {
  uint64_t r = UINT64_MAX;
  for (int64_t i = 0; i < len; i++) r = b->count[i] < r ? b->count[i] : r;
  return r;
}

static void bmc_grow(or_bmc *b, int64_t need) {
  if (need <= b->cap) return;
  int64_t nc = b->cap ? b->cap : 16;
  while (nc < need) nc *= 2;
  bcounts *all[5] = {&b->lnw, &b->en, &b->tf, &b->hist, &b->pr};
  for (int k = 0; k < 5; k++) {
    if (!all[k]->exists) continue;
    all[k]->total = (double *)realloc(all[k]->total, sizeof(double) * nc);
    all[k]->count = (uint64_t *)realloc(all[k]->count, sizeof(uint64_t) * nc);
    all[k]->fc = (double *)realloc(all[k]->fc, sizeof(double) * nc);
  }
  b->cap = nc;
}

/* histogram.rs:47-58: insert_zero / push_zero on lnw and every extra. */
static void bmc_insert_zero(or_bmc *b) {
  bmc_grow(b, b->len + 1);
  bcounts *all[5] = {&b->lnw, &b->en, &b->tf, &b->hist, &b->pr};
  for (int k = 0; k < 5; k++) {
    bcounts *c = all[k];
    if (!c->exists) continue;
    memmove(c->total + 1, c->total, sizeof(double) * (size_t)b->len);
    memmove(c->count + 1, c->count, sizeof(uint64_t) * (size_t)b->len);
    memmove(c->fc + 1, c->fc, sizeof(double) * (size_t)b->len);
    c->total[0] = 0.0;
    c->count[0] = 0;
    c->fc[0] = 0.0;
    c->min_total = 0.0;
    c->min_count = 0;
    c->fmin_count = 0.0;
  }
  b->len += 1;
}
static void bmc_push_zero(or_bmc *b) {
  bmc_grow(b, b->len + 1);
  bcounts *all[5] = {&b->lnw, &b->en, &b->tf, &b->hist, &b->pr};
  for (int k = 0; k < 5; k++) {
    bcounts *c = all[k];
    if (!c->exists) continue;
    c->total[b->len] = 0.0;
    c->count[b->len] = 0;
    c->fc[b->len] = 0.0;
    c->min_total = 0.0;
    c->min_count = 0;
    c->fmin_count = 0.0;
  }
  b->len += 1;
}

/* histogram.rs:129-140; -1 stands for the usize::MAX of the out-of-range
 * cases (e below min, or the top edge of an empty histogram). */
static int64_t bmc_index(const or_bmc *b, double e) {
  if (e < b->min) return -1;
  double i = (e - b->min) / b->width;
  if (i == (double)b->len) return b->len == 0 ? -1 : (int64_t)i - 1;
  if (!(i < 9.2e18)) return INT64_MAX;
  return (int64_t)i;
}
static inline double bmc_energy(const or_bmc *b, int64_t i) { return b->min + ((double)i + 0.5) * b->width; }
static inline double bc_total_at(const bcounts *c, int64_t len, int64_t idx) {
  return (idx >= 0 && idx < len) ? c->total[idx] : 0.0;
}
static inline uint64_t bc_count_at(const bcounts *c, int64_t len, int64_t idx) {
  return (idx >= 0 && idx < len) ? c->count[idx] : 0;
}
static inline double bmc_get_lnw(const or_bmc *b, double e) { return bc_total_at(&b->lnw, b->len, bmc_index(b, e)); }
/* get_count as an integer count: count/width compares with 0 and with
 * other counts/width exactly as the counts do. */
static inline uint64_t bmc_get_count(const or_bmc *b, double e) {
  return bc_count_at(&b->lnw, b->len, bmc_index(b, e));
}

/* The fractional index of linear.rs:201-203 and its interpretation
 * (linear.rs:84-98): the bins and weights in the order the reference's
 * iterator yields them. */
typedef struct {
  int n;
  int64_t i[2];
  double f[2];
} lin_idx;

static lin_idx lin_interpret(int64_t len, double fidx) {
  lin_idx r;
  r.n = 0;
  if (fidx < -1.0) return r;
  if (fidx < 0.0) {
    r.n = 1;
    r.i[0] = 0;
    r.f[0] = 1.0 - (-fidx);
  } else if (fidx < (double)len - 1.0) {
    int64_t i = (int64_t)fidx;
    double o = fidx - (double)i;
    r.n = 2;
    r.i[0] = i + 1;
    r.f[0] = o;
    r.i[1] = i;
    r.f[1] = 1.0 - o;
  } else if (fidx < (double)len) {
    r.n = 1;
    r.i[0] = len - 1;
    r.f[0] = 1.0 - (fidx - ((double)len - 1.0));
  }
  return r;
}

/* BinCounts::get_total / get_count (linear.rs:139-152). */
static double lin_get(const double *a, int64_t len, double fidx) {
  lin_idx r = lin_interpret(len, fidx);
  double t = 0.0;
  for (int k = 0; k < r.n; k++) t += a[r.i[k]] * r.f[k];
  return t;
}

static inline double lin_fidx(const or_bmc *b, double e) { return (e - b->min) / b->width; }

/* get_lnw and get_count (as PerEnergy, count / width) of either binning. */
static double bm_get_lnw(const or_bmc *b, double e) {
  if (b->linear) return b->len ? lin_get(b->lnw.total, b->len, lin_fidx(b, e)) : 0.0;
  return bmc_get_lnw(b, e);
}
static double bm_get_count(const or_bmc *b, double e) {
  if (b->linear) return (b->len ? lin_get(b->lnw.fc, b->len, lin_fidx(b, e)) : 0.0) / b->width;
  return (double)bmc_get_count(b, e) / b->width;
}

/* histogram.rs:141-158 / linear.rs:204-222 (the linear binning keeps one
 * more bin above the energy). */
static void bmc_prep(or_bmc *b, double e) {
  if (b->len == 0) b->min = floor(e / b->width) * b->width;
  while (e < b->min) {
    bmc_insert_zero(b);
    b->min -= b->width;
  }
  if (b->linear) {
    while (e >= b->min + b->width * ((double)b->len - 1.0)) bmc_push_zero(b);
  } else {
    while (e >= b->min + b->width * (double)b->len) bmc_push_zero(b);
  }
}

static double fc_min_of(const bcounts *c, int64_t len) {
  double r = NAN;
  for (int64_t i = 0; i < len; i++) r = fold_min(r, c->fc[i]);
  return r;
}

/* BinCounts::increment_count of linear.rs:99-137 on ln w. */
static void lin_increment(or_bmc *b, double elo, double ehi, double fidx, double value) {
  bcounts *c = &b->lnw;
  int64_t len = b->len;
  int64_t idx = (int64_t)fidx;
  double offset = fidx - (double)idx;
  c->total_count += 1;
  double oc = c->fc[idx], opc = c->fc[idx + 1];
  c->fc[idx] += 1.0 - offset;
  c->fc[idx + 1] += offset;
  double ot = c->total[idx], opt = c->total[idx + 1];
  c->total[idx] += value * (1.0 - offset);
  c->total[idx + 1] += value * offset;
  if (c->total[idx] > c->max_total && bc_max_of(c, len) == c->total[idx]) {
    c->max_total = c->total[idx];
    c->e_max_total = elo;
  }
  if (c->total[idx + 1] > c->max_total && bc_max_of(c, len) == c->total[idx + 1]) {
    c->max_total = c->total[idx + 1];
    c->e_max_total = ehi;
  }
  if (ot == c->min_total || opt == c->min_total) c->min_total = bc_min_of(c, len);
  if (oc == c->fmin_count || opc == c->fmin_count) c->fmin_count = fc_min_of(c, len);
  if (c->fc[idx] > c->fmax_count) {
    c->fmax_count = c->fc[idx];
    c->e_max_count = elo;
  }
  if (c->fc[idx + 1] > c->fmax_count) {
    c->fmax_count = c->fc[idx + 1];
    c->e_max_count = ehi;
  }
}

/* BinCounts::increment_count (histogram.rs:59-81). */
static void bc_increment(bcounts *c, int64_t len, double e, int64_t idx, double value) {
  c->total_count += 1;
  uint64_t old_count = c->count[idx];
  c->count[idx] += 1;
  double old_total = c->total[idx];
  c->total[idx] += value;
  if (c->total[idx] > c->max_total && bc_max_of(c, len) == c->total[idx]) {
    c->max_total = c->total[idx];
    c->e_max_total = e;
  }
  if (old_total == c->min_total) c->min_total = bc_min_of(c, len);
  if (old_count == c->min_count) c->min_count = bc_min_count(c, len);
  if (c->count[idx] > c->max_count) {
    c->max_count = c->count[idx];
    c->e_max_count = e;
  }
}

/* Bins::accumulate_extra (histogram.rs:219-245); a missing extra is created
 * with the current length (BinCounts::new). */
static int lin_accumulate(or_bmc *b, bcounts *c, double e, double value);

static int bmc_accumulate(or_bmc *b, bcounts *c, double e, double value) {
  if (b->linear) return lin_accumulate(b, c, e, value);
  bmc_prep(b, e);
  int64_t idx = bmc_index(b, e);
  if (idx < 0 || idx >= b->len) {  /* the reference's assert (histogram.rs:222) */
    b->status = ST_INDEX;
    return 0;
  }
  if (!c->exists) {
    bc_new(c, b->len, b->cap);
    memset(c->total, 0, sizeof(double) * (size_t)b->cap);
    memset(c->count, 0, sizeof(uint64_t) * (size_t)b->cap);
  }
  c->total_count += 1;
  c->count[idx] += 1;
  if (c->count[idx] == c->min_count + 1) c->min_count = bc_min_count(c, b->len);
  if (c->count[idx] > c->max_count) c->max_count = c->count[idx];
  double old_total = c->total[idx];
  c->total[idx] = old_total + value;
  if (c->total[idx] > c->max_total) c->max_total = c->total[idx];
  if (old_total == c->min_total) c->min_total = bc_min_of(c, b->len);
  return 1;
}

/* Bins::accumulate_extra of linear.rs:267-310. */
static int lin_accumulate(or_bmc *b, bcounts *c, double e, double value) {
  bmc_prep(b, e);
  double fidx = lin_fidx(b, e);
  int64_t i = (int64_t)fidx;
  if (!(fidx >= 0.0) || i + 1 >= b->len) {  /* an assert / index panic in the reference */
    b->status = ST_INDEX;
    return 0;
  }
  if (!c->exists) {
    bc_new(c, b->len, b->cap);
  }
  double off = fidx - (double)i;
  c->total_count += 1;
  double oc = c->fc[i], opc = c->fc[i + 1];
  c->fc[i] += 1.0 - off;
  c->fc[i + 1] += off;
  if (oc == c->fmin_count || opc == c->fmin_count) c->fmin_count = fc_min_of(c, b->len);
  if (c->fc[i] > c->fmax_count) c->fmax_count = c->fc[i];
  if (c->fc[i + 1] > c->fmax_count) c->fmax_count = c->fc[i + 1];
  double ot = c->total[i], opt = c->total[i + 1];
  c->total[i] = ot + value * (1.0 - off);
  c->total[i + 1] = opt + value * off;
  if (c->total[i] > c->max_total) c->max_total = c->total[i];
  if (c->total[i + 1] > c->max_total) c->max_total = c->total[i + 1];
  if (ot == c->min_total || opt == c->min_total) c->min_total = bc_min_of(c, b->len);
  return 1;
}

/* Bins::zero_out_extra (histogram.rs:246-259). */
static void bc_zero_out(bcounts *c, int64_t len) {
  if (!c->exists) return;
  memset(c->count, 0, sizeof(uint64_t) * (size_t)len);
  for (int64_t i = 0; i < len; i++) c->total[i] = 0.0;
  for (int64_t i = 0; i < len; i++) c->fc[i] = 0.0;
  c->min_total = 0.0;
  c->max_total = -INFINITY;
  c->min_count = 0;
  c->max_count = 0;
  c->fmin_count = 0.0;
  c->fmax_count = 0.0;
  c->total_count = 0;
}

/* Bins::increment_count (histogram.rs:171-181). */
static int bmc_increment(or_bmc *b, double e, double gamma) {
  if (e > b->max_e) b->max_e = e;
  if (e < b->min_e) b->min_e = e;
  bmc_prep(b, e);
  if (b->linear) {  /* linear.rs:237-253 */
    double fidx = lin_fidx(b, e);
    int64_t ii = (int64_t)fidx;
    if (!(fidx >= 0.0) || ii + 1 >= b->len) {
      b->status = ST_INDEX;
      return 0;
    }
    double rg = gamma * 1.0 / b->width;
    lin_increment(b, bmc_energy(b, ii), bmc_energy(b, ii + 1), fidx, rg);
    return 1;
  }
  int64_t idx = bmc_index(b, e);
  if (idx < 0 || idx >= b->len) {  /* an index panic in the reference */
    b->status = ST_INDEX;
    return 0;
  }
  bc_increment(&b->lnw, b->len, e, idx, gamma);
  return 1;
}

/* count_states (histogram.rs:191-200) with the SAD predicate
 * e >= too_lo && e <= too_hi (count >= 0 always holds). */
static uint64_t bmc_count_states(const or_bmc *b) {
  uint64_t t = 0;
  for (int64_t i = 0; i < b->len; i++) {
    double e = bmc_energy(b, i);
    if (e >= b->too_lo && e <= b->too_hi) t++;
  }
  return t;
}

/* energy_binning.rs:503-530. */
static double bmc_gamma(const or_bmc *b) {
  if (b->method == OR_SAD) {
    double ns = (double)b->num_states;
    if (b->latest * b->tF * ns == 0.0) return 0.0;
    double t = (double)b->moves;
    return (b->latest + t / b->tF) / (b->latest + t / ns * (t / b->tF));
  }
  if (b->method == OR_SAMC) {
    double t = (double)b->moves;
    return t > b->t0 ? b->t0 / t : 1.0;
  }
  return b->gamma;
}

/* energy_binning.rs:276-319. */
static int bmc_reject(or_bmc *b, double e1, double e2) {
  double lnw1 = bm_get_lnw(b, e1), lnw2 = bm_get_lnw(b, e2);
  if (b->method == OR_SAD) {
    if (e1 < b->too_lo) lnw1 = bm_get_lnw(b, b->too_lo) + (e1 - b->too_lo) / b->min_T;
    else if (e1 > b->too_hi) lnw1 = bm_get_lnw(b, b->too_hi);
    if (e2 < b->too_lo) lnw2 = bm_get_lnw(b, b->too_lo) + (e2 - b->too_lo) / b->min_T;
    else if (e2 > b->too_hi) lnw2 = bm_get_lnw(b, b->too_hi);
    int rejected = lnw2 > lnw1 && or_gen_f64(b->m->rng) > exp(lnw1 - lnw2);
    if (!rejected && bm_get_count(b, e2) == 0.0 && e2 < b->too_hi && e2 > b->too_lo) b->tL = b->moves;
    return rejected;
  }
  return lnw2 > lnw1 && or_gen_f64(b->m->rng) > exp(lnw1 - lnw2);
}

/* energy_binning.rs:320-500. */
/* The side histogram's Bins::increment_count(e, 0.0): histogram.rs:182-192
 * (min_e / max_e, prep_for_e at :147-167 with insert_zero / push_zero at
 * :47-58, energy_to_index at :129-140) and BinCounts::increment_count at
 * :59-81 with value 0, its rescans included. */
static int hr_increment(or_bmc *b, double e) {
  if (e > b->hr_max_e) b->hr_max_e = e;
  if (e < b->hr_min_e) b->hr_min_e = e;
  if (b->hr_len == 0) b->hr_min = floor(e / b->hr_width) * b->hr_width;
  while (e < b->hr_min || e >= b->hr_min + b->hr_width * (double)b->hr_len) {
    if (b->hr_len + 1 > b->hr_cap) {
      b->hr_cap = b->hr_cap ? 2 * b->hr_cap : 16;
      b->hr_total = (double *)realloc(b->hr_total, sizeof(double) * b->hr_cap);
      b->hr_count = (uint64_t *)realloc(b->hr_count, sizeof(uint64_t) * b->hr_cap);
    }
    if (e < b->hr_min) { /* insert_zero */
      memmove(b->hr_total + 1, b->hr_total, sizeof(double) * (size_t)b->hr_len);
      memmove(b->hr_count + 1, b->hr_count, sizeof(uint64_t) * (size_t)b->hr_len);
      b->hr_total[0] = 0.0;
      b->hr_count[0] = 0;
      b->hr_min -= b->hr_width;
    } else { /* push_zero */
      b->hr_total[b->hr_len] = 0.0;
      b->hr_count[b->hr_len] = 0;
    }
    b->hr_len += 1;
    b->hr_min_total = 0.0;
    b->hr_min_count = 0;
  }
  double fi = (e - b->hr_min) / b->hr_width;
  int64_t idx = fi == (double)b->hr_len ? (int64_t)fi - 1 : (int64_t)fi;
  if (!(e >= b->hr_min) || idx < 0 || idx >= b->hr_len) { /* an index panic in the reference */
    b->status = ST_INDEX;
    return 0;
  }
  b->hr_total_count += 1;
  uint64_t old_count = b->hr_count[idx];
  b->hr_count[idx] += 1;
  double old_total = b->hr_total[idx];
  b->hr_total[idx] += 0.0;
  if (b->hr_total[idx] > b->hr_max_total) {
    double mx = NAN;
    for (int64_t i = 0; i < b->hr_len; i++) mx = fold_max(mx, b->hr_total[i]);
    if (mx == b->hr_total[idx]) {
      b->hr_max_total = b->hr_total[idx];
      b->hr_e_max_total = e;
    }
  }
  if (old_total == b->hr_min_total) {
    double mn = NAN;
    for (int64_t i = 0; i < b->hr_len; i++) mn = fold_min(mn, b->hr_total[i]);
    b->hr_min_total = mn;
  }
  if (old_count == b->hr_min_count) {
    uint64_t mn = UINT64_MAX;
    for (int64_t i = 0; i < b->hr_len; i++) mn = b->hr_count[i] < mn ? b->hr_count[i] : mn;
    b->hr_min_count = mn;
  }
  if (b->hr_count[idx] > b->hr_max_count) {
    b->hr_max_count = b->hr_count[idx];
    b->hr_e_max_count = e;
  }
  return 1;
}

static void bmc_update(or_bmc *b, double energy) {
  double gamma = bmc_gamma(b);
  double old_highest = (b->linear ? b->lnw.fmax_count : (double)b->lnw.max_count) / b->width;
  double old_here = bm_get_count(b, energy);
  if (!bmc_increment(b, energy, gamma)) return;
  if (b->has_hr && !hr_increment(b, energy)) return;
  if (b->method == OR_SAD) {
    double here = bm_get_count(b, energy);
    if (old_here == 0.0 && !bmc_accumulate(b, &b->tf, energy, (double)b->moves)) return;
    if (here > old_highest) {
      if (energy > b->too_hi) {
        double v = bm_get_lnw(b, b->too_hi);
        for (int64_t i = 0; i < b->len; i++) {  /* set_lnw, histogram.rs:182-190, linear.rs:254-262 */
          double e = bmc_energy(b, i);
          if (e > b->too_hi && bm_get_count(b, e) > 0.0) {
            b->lnw.total[i] = v;
            b->lnw.count[i] = 0;
            b->lnw.fc[i] = 0.0;
          }
        }
        b->latest = (energy - b->too_lo) / b->min_T;
        b->tL = b->moves;
        b->too_hi = energy;
        b->num_states = bmc_count_states(b);
      } else if (energy < b->too_lo) {
        double v = bm_get_lnw(b, b->too_lo);
        double lo = b->too_lo;
        for (int64_t i = 0; i < b->len; i++) {
          double e = bmc_energy(b, i);
          if (e < lo && bm_get_count(b, e) > 0.0) {
            b->lnw.total[i] = v + (e - lo) / b->min_T;
            b->lnw.count[i] = 0;
            b->lnw.fc[i] = 0.0;
          }
        }
        b->latest = (b->too_hi - energy) / b->min_T;
        b->tL = b->moves;
        b->too_lo = energy;
        b->num_states = bmc_count_states(b);
      }
    }
    if (b->tL == b->moves) {
      double old_tF = b->tF;
      b->tF = b->tf.exists ? b->tf.max_total : 0.0;
      if (old_tF != b->tF && b->bp.move_plan == 1) {
        double s = b->acc_rate / b->bp.move_param;
        s = s < 0.8 ? 0.8 : (s > 1.2 ? 1.2 : s);
        b->tscale *= s;
      }
    }
  } else if (b->method == OR_WL) {
    double w = b->width;
    double old_lowest = b->hist.exists ? (b->linear ? b->hist.fmin_count : (double)b->hist.min_count) / w : 0.0;
    if (!bmc_accumulate(b, &b->hist, energy, 0.0)) return;
    if (b->has_min_gamma && b->gamma < b->min_gamma) return;
    double lowest = (b->linear ? b->hist.fmin_count : (double)b->hist.min_count) / w;
    int ok_lo = !b->bp.has_min_e || bm_get_count(b, b->bp.min_e) > 0.0;
    int ok_hi = !b->bp.has_max_e || bm_get_count(b, b->bp.max_e) > 0.0;
    if (lowest > old_lowest && ok_lo && ok_hi) {
      double mean = (double)b->hist.total_count / (w * (double)b->len);
      if ((b->inv_t && lowest > 0.0) || lowest >= 0.8 * mean) {
        b->gamma *= 0.5;
        bc_zero_out(&b->hist, b->len);
        if (b->has_min_gamma && b->gamma < b->min_gamma) b->gamma = 0.0;
      }
      if (b->inv_t && b->gamma < (double)b->len / (double)b->moves) {
        b->method = OR_SAMC;
        b->t0 = (double)b->len;
      }
    }
  }
}

/* energy_binning.rs:566-611. */
static void bmc_move_once(or_bmc *b) {
  or_mc *m = b->m;
  b->moves += 1;
  m->moves = b->moves;
  if (b->moves % 100000000ULL == 0) {
    sys_verify_energy(m);
    if (m->status) {
      b->status = m->status;
      return;
    }
  }
  double e1 = m->E;
  if (!bmc_accumulate(b, &b->en, e1, e1)) return;
  double pv;
  if (sys_pressure(m, &pv) && !bmc_accumulate(b, &b->pr, e1, pv)) return;
  double rs = sqrt(1.0 / (double)b->moves);
  b->acc_rate *= 1.0 - rs;
  double e2;
  if (sys_plan_move(m, b->tscale, &e2)) {
    int oob = 0;
    if (b->bp.has_max_e) oob = e2 > b->bp.max_e && e2 > e1;
    if (b->bp.has_min_e) oob = oob || (e2 < b->bp.min_e && e2 < e1);
    if (!oob && !bmc_reject(b, e1, e2)) {
      b->accepted += 1;
      b->acc_rate += rs;
      sys_confirm(m);
    }
  }
  bmc_update(b, m->E);
}

/* energy_binning.rs:533-564 (from_params) on a built system. */
static or_bmc *bmc_from_system(const or_params *bp, or_mc *m) {
  or_bmc *b = (or_bmc *)calloc(1, sizeof(or_bmc));
  b->bp = *bp;
  b->m = m;
  if (bp->has_max_e) {
    for (uint64_t k = 0; k < 100000000ULL; k++) {
      double newe;
      if (sys_plan_move(m, 0.05, &newe)) {
        if (newe < m->E) sys_confirm(m);
        if (m->E < bp->max_e) break;
      }
    }
  }
  double E = m->E;
  b->method = bp->method == OR_INV_T_WL ? OR_WL : bp->method;
  b->min_T = bp->method_param;
  b->too_lo = E;
  b->too_hi = E;
  b->t0 = bp->method_param;
  b->gamma = 1.0;
  b->inv_t = bp->method == OR_INV_T_WL;
  b->has_min_gamma = bp->method == OR_WL && bp->has_min_gamma;
  b->min_gamma = bp->method_param;
  b->width = bp->energy_bin > 0 ? bp->energy_bin : 1.0;
  b->linear = bp->binning == 1;
  b->min = (round(E / b->width) - 0.5) * b->width;  /* Bins::new, histogram.rs:162-170, linear.rs:226-236 */
  b->min_e = E;
  b->max_e = E;
  b->len = 0;
  if (bp->high_resolution_de > 0) { /* Bins::new(system.energy(), de), energy_binning.rs:568-570 */
    b->has_hr = 1;
    b->hr_width = bp->high_resolution_de;
    b->hr_min = (round(E / b->hr_width) - 0.5) * b->hr_width;
    b->hr_min_e = E;
    b->hr_max_e = E;
    b->hr_e_max_total = -INFINITY;
    b->hr_e_max_count = -INFINITY;
  }
  bmc_grow(b, 16);
  bc_new(&b->lnw, 0, b->cap);
  b->acc_rate = 0.5;
  b->tscale = bp->move_plan == 0 ? bp->move_param : 0.05;
  return b;
}

static or_params bmc_system_params(const or_params *p) {
  or_params q = *p;
  q.has_max_e = 0;
  q.has_min_e = 0;
  q.method = OR_CANONICAL;
  q.method_param = 1.0;
  return q;
}

or_bmc *or_bmc_new(const or_params *p, const double *coords, const int8_t *spins) {
  or_params q = bmc_system_params(p);
  return bmc_from_system(p, or_mc_new(&q, coords, spins));  /* generator: seed_from_u64(seed) */
}

/* The same walker over the scripted system (plan_move returns the script's
 * energies in turn, NaN meaning no move). */
or_bmc *or_bmc_new_scripted(const or_params *p, double e0, const double *script, int64_t len) {
  or_params q = bmc_system_params(p);
  return bmc_from_system(p, or_mc_new_scripted(&q, e0, script, len));
}

int or_bmc_run(or_bmc *b, uint64_t moves) {
  for (uint64_t k = 0; k < moves && !b->status; k++) bmc_move_once(b);
  return b->status;
}

static void bc_export(const bcounts *c, or_bcounts *o, int linear) {
  o->min_total = c->min_total;
  o->max_total = c->max_total;
  o->e_max_total = c->e_max_total;
  o->e_max_count = c->e_max_count;
  o->min_count = linear ? f64_bits(c->fmin_count) : c->min_count;
  o->max_count = linear ? f64_bits(c->fmax_count) : c->max_count;
  o->total_count = c->total_count;
  o->exists = c->exists;
}

void or_bmc_state(const or_bmc *b, or_bstate *o) {
  memset(o, 0, sizeof *o);
  o->rng[0] = b->m->rng[0];
  o->rng[1] = b->m->rng[1];
  o->E = b->m->E;
  o->error = b->m->error;
  o->moves = b->moves;
  o->accepted = b->accepted;
  o->acceptance_rate = b->acc_rate;
  o->translation_scale = b->tscale;
  o->min = b->min;
  o->width = b->width;
  o->min_e = b->min_e;
  o->max_e = b->max_e;
  o->len = b->len;
  o->method = b->method;
  o->status = b->status;
  o->num_states = b->num_states;
  o->tL = b->tL;
  o->min_T = b->min_T;
  o->too_lo = b->too_lo;
  o->too_hi = b->too_hi;
  o->tF = b->tF;
  o->latest_parameter = b->latest;
  o->t0 = b->t0;
  o->gamma = b->gamma;
  bc_export(&b->lnw, &o->lnw, b->linear);
  bc_export(&b->en, &o->energy, b->linear);
  bc_export(&b->tf, &o->t_found, b->linear);
  bc_export(&b->hist, &o->hist, b->linear);
  bc_export(&b->pr, &o->pressure, b->linear);
}

/* Logical bins [0, len): totals and counts of lnw and the extras (zeros for
 * an extra that does not exist). */
void or_bmc_bins(const or_bmc *b, double *lnw_total, uint64_t *lnw_count, double *energy_total,
                 uint64_t *energy_count, double *tf_total, uint64_t *tf_count, uint64_t *hist_count,
                 double *p_total, uint64_t *p_count) {
  int64_t n = b->len;
  const bcounts *cs[5] = {&b->lnw, &b->en, &b->tf, &b->hist, &b->pr};
  double *tt[5] = {lnw_total, energy_total, tf_total, NULL, p_total};
  uint64_t *cc[5] = {lnw_count, energy_count, tf_count, hist_count, p_count};
  for (int k = 0; k < 5; k++) {
    for (int64_t i = 0; i < n; i++) {
      if (tt[k]) tt[k][i] = cs[k]->exists ? cs[k]->total[i] : 0.0;
      if (cc[k]) cc[k][i] = cs[k]->exists ? (b->linear ? f64_bits(cs[k]->fc[i]) : cs[k]->count[i]) : 0;
    }
  }
}

/* get_lnw and get_count (PerEnergy) at any energy, as the walker sees them. */
double or_bmc_lnw_at(const or_bmc *b, double e) { return bm_get_lnw(b, e); }
double or_bmc_count_at(const or_bmc *b, double e) { return bm_get_count(b, e); }

void or_bmc_coords(const or_bmc *b, double *c) { or_mc_coords(b->m, c); }
void or_bmc_spins(const or_bmc *b, int8_t *s) { or_mc_spins(b->m, s); }

void or_bmc_free(or_bmc *b) {
  if (!b) return;
  bcounts *all[5] = {&b->lnw, &b->en, &b->tf, &b->hist, &b->pr};
  for (int k = 0; k < 5; k++)
    if (all[k]->exists) bc_free(all[k]);
  free(b->hr_total);
  free(b->hr_count);
  or_mc_free(b->m);
  free(b);
}

int64_t or_bmc_hires(const or_bmc *b, double *stats, uint64_t *counts) {
  if (!b->has_hr) return -1;
  if (stats) {
    stats[0] = b->hr_min;
    stats[1] = b->hr_width;
    stats[2] = b->hr_min_e;
    stats[3] = b->hr_max_e;
    stats[4] = (double)b->hr_len;
    stats[5] = (double)b->hr_total_count;
    stats[6] = (double)b->hr_min_count;
    stats[7] = (double)b->hr_max_count;
    stats[8] = b->hr_e_max_count;
    stats[9] = b->hr_min_total;
  }
  if (counts && b->hr_len > 0) memcpy(counts, b->hr_count, sizeof(uint64_t) * (size_t)b->hr_len);
  return b->hr_len;
}

/* ===================================================================== */
/* Energy-ceiling replicas ("ZMC", energy_replicas.rs)                   */
/* ===================================================================== */

/* A deep copy of a system (positions or spins, energy, error estimate). */
static or_mc *mc_clone(const or_mc *m) {
  or_mc *c = (or_mc *)malloc(sizeof(or_mc));
  *c = *m;
  c->lnw = NULL; c->etot = NULL; c->esq = NULL; c->ptotal = NULL; c->hist = NULL; c->tfound = NULL;
  c->rt = NULL; c->pcount = NULL; c->vis = NULL; c->wlh = NULL; c->tr_e = NULL; c->tr_a = NULL;
  c->script = NULL;
  if (m->x) {
    c->x = (double *)malloc(sizeof(double) * 3 * m->n);
    memcpy(c->x, m->x, sizeof(double) * m->n);
    memcpy(c->x + m->n, m->y, sizeof(double) * m->n);
    memcpy(c->x + 2 * m->n, m->z, sizeof(double) * m->n);
    c->y = c->x + m->n;
    c->z = c->x + 2 * m->n;
  }
  if (m->S) {
    c->S = (int8_t *)malloc((size_t)m->n * m->n);
    memcpy(c->S, m->S, (size_t)m->n * m->n);
  }
  return c;
}

/* The WCA energy of the first m atoms of SoA arrays, in lane-group order. */
static double or_wca_energy_first(const double *x, const double *y, const double *z, int m, double L, double rc2,
                                  int G) {
  double h = 0.5 * L, acc[64];
  for (int l = 0; l < G; l++) {
    acc[l] = 0.0;
    for (int i = l; i < m; i += G)
      for (int j = 0; j < i; j++)
        acc[l] += or_wca_pair(norm2(min_image(x[i] - x[j], L, h), min_image(y[i] - y[j], L, h),
                                    min_image(z[i] - z[j], L, h)),
                              rc2);
  }
  return butterfly(acc, G);
}

/* System::randomize: ising.rs:79-85 (the low bit of each draw, in site
 * order), lj.rs:262-279 (each atom uniform in the sphere by rejection from
 * Uniform(-1, 1)^3, scaled by the radius), wca.rs:251-267 (below).  The
 * energy is recomputed; Ising and LJ leave the error estimate as it was. */
static double sys_randomize(or_mc *m, uint64_t rng[2]) {
  if (m->sys == OR_ISING) {
    for (int64_t k = 0; k < (int64_t)m->n * m->n; k++) m->S[k] = (int8_t)(((int8_t)or_rng_next(rng) & 1) * 2 - 1);
    m->E = or_ising_energy(m->n, m->S);
  } else if (m->sys == OR_WCA) {
    /* wca.rs:251-267: every atom removed, then added back one at a time at
     * put_in_cell(Uniform(0, L)^3); each addition is a Change::Add whose
     * confirm goes through set_energy (wca.rs:105-116, 164-180) with the
     * atom count after the addition, so the error estimate (and, when it
     * trips, an energy over the atoms added so far) evolves from the old
     * system's E; the energy is recomputed at the end.  dabse, the new atom's
     * pair terms with the atoms already there, is summed in lane-group order
     * (the reference sums it in cell-list order). */
    const double L = m->L, h = m->halfL;
    const int n = m->n, G = m->G > 0 ? m->G : 1;
    for (int i = 0; i < n; i++) {
      double x = put_in_cell1(or_uniform_f64(rng, 0.0, L), L);
      double y = put_in_cell1(or_uniform_f64(rng, 0.0, L), L);
      double z = put_in_cell1(or_uniform_f64(rng, 0.0, L), L);
      double acc[64];
      for (int l = 0; l < G; l++) acc[l] = 0.0;
      for (int j = 0; j < i; j++)
        acc[j % G] += or_wca_pair(norm2(min_image(m->x[j] - x, L, h), min_image(m->y[j] - y, L, h),
                                        min_image(m->z[j] - z, L, h)),
                                  m->rc2);
      double dabse = butterfly(acc, G);
      m->x[i] = x;
      m->y[i] = y;
      m->z[i] = z;
      double new_e = m->E + dabse;
      double cnt = (double)(i + 1);
      double single = dabse > fabs(new_e) ? 1e-14 * dabse * cnt : 1e-14 * fabs(new_e) * cnt;
      m->error += single * cnt;
      if (m->error > fabs(new_e) * 1e-13 * cnt * cnt) {
        m->E = or_wca_energy_first(m->x, m->y, m->z, i + 1, L, m->rc2, G);
        m->error = 1e-15 * m->E * cnt;
      } else {
        m->E = new_e;
      }
    }
    m->E = sys_compute_energy(m, m->G);
  } else {
    double R = m->p.radius;
    for (int i = 0; i < m->n; i++) {
      double x, y, z;
      for (;;) {
        x = or_uniform_f64(rng, -1.0, 1.0);
        y = or_uniform_f64(rng, -1.0, 1.0);
        z = or_uniform_f64(rng, -1.0, 1.0);
        if (norm2(x, y, z) < 1.0) break;
      }
      m->x[i] = x * R;
      m->y[i] = y * R;
      m->z[i] = z * R;
    }
    m->E = sys_compute_energy(m, m->G);
  }
  return m->E;
}

struct or_zmc {
  or_params p;
  int R, cap;
  or_mc **sys;  /* replica r's system (swapped with the systems) */
  or_zrep *r;
  uint64_t rng[2];
  uint64_t moves, isbnb;
  double min_T;
  double *med;
  int med_len;
};

/* max_size: ising.rs:120-122, lj.rs:375-377, wca.rs:354-356. */
static double zmc_max_size(const or_mc *m) {
  if (m->sys == OR_ISING) return 0.5;
  if (m->sys == OR_WCA) return sqrt(norm2(m->L, m->L, m->L));
  return m->p.radius;
}
static uint64_t zmc_dim(const or_mc *m) { return m->sys == OR_ISING ? sys_moves_to_randomize(m) : 3 * sys_moves_to_randomize(m); }

/* Replica::new (energy_replicas.rs:145-168). */
static void zrep_new(or_zrep *r, double max_e, double cutoff, const or_mc *sys, const uint64_t rng[2]) {
  memset(r, 0, sizeof *r);
  r->max_energy = max_e;
  r->cutoff_energy = cutoff;
  r->translation_scale = zmc_max_size(sys);
  r->lowest_max_energy = max_e;
  r->unique_visitors = 1;
  r->collecting_data = 1;
  r->rng[0] = rng[0];
  r->rng[1] = rng[1];
}

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* MC::from_params (energy_replicas.rs:344-395). */
or_zmc *or_zmc_new(const or_params *p, uint64_t seed, double min_T, uint64_t isbnb, const double *coords,
                   const int8_t *spins) {
  or_params q = *p;
  q.has_max_e = 0;
  q.has_min_e = 0;
  q.method = OR_CANONICAL;
  q.method_param = 1.0;
  q.seed = seed;
  or_zmc *z = (or_zmc *)calloc(1, sizeof(or_zmc));
  z->p = q;
  z->min_T = min_T;
  z->isbnb = isbnb ? isbnb : 64;
  or_mc *system = or_mc_new(&q, coords, spins);
  or_rng_seed(seed, z->rng);
  const int MAX_INIT = 1 << 15;
  double *energies = (double *)malloc(sizeof(double) * MAX_INIT);
  for (int k = 0; k < MAX_INIT; k++) energies[k] = sys_randomize(system, z->rng);
  or_mc *high = mc_clone(system);
  sys_randomize(high, z->rng);
  while (system->E > energies[MAX_INIT / 2]) sys_randomize(system, z->rng);
  qsort(energies, MAX_INIT, sizeof(double), cmp_double);
  z->cap = 16;
  z->sys = (or_mc **)calloc(z->cap, sizeof(or_mc *));
  z->r = (or_zrep *)calloc(z->cap, sizeof(or_zrep));
  z->sys[0] = high;
  z->sys[1] = system;
  zrep_new(&z->r[0], INFINITY, energies[MAX_INIT / 2], high, z->rng);
  zrep_new(&z->r[1], energies[MAX_INIT / 2], energies[MAX_INIT / 4], system, z->rng);
  z->R = 2;
  or_rng_jump(z->rng);
  z->med = (double *)malloc(sizeof(double) * 4096);
  z->med[0] = energies[MAX_INIT / 4];
  z->med_len = 1;
  free(energies);
  return z;
}

/* Replica::run_once (energy_replicas.rs:200-245); LJ and Ising collect no
 * extra data. */
/* wca.rs:202-218: data_to_collect(iter) is the pressure every N^2 moves
 * (all ordered pairs, lane-group order). */
static int zmc_pressure(const or_mc *m, uint64_t iter, double *val) {
  if (m->sys != OR_WCA) return 0;
  if (iter % ((uint64_t)m->n * (uint64_t)m->n) != 0) return 0;
  double acc[64], L = m->L, h = m->halfL;
  int n = m->n, G = m->G > 0 ? m->G : 1;
  for (int l = 0; l < G; l++) {
    acc[l] = 0.0;
    for (int i = l; i < n; i += G)
      for (int j = 0; j < n; j++)
        if (j != i)
          acc[l] += wca_pair_pressure(norm2(min_image(m->x[i] - m->x[j], L, h), min_image(m->y[i] - m->y[j], L, h),
                                            min_image(m->z[i] - m->z[j], L, h)), m->rc2);
  }
  *val = butterfly(acc, G) / (3.0 * (L * L * L));
  return 1;
}

static void zrep_run_once(or_zrep *r, or_mc *m, double very_lowest_max, uint64_t iter) {
  m->rng[0] = r->rng[0];
  m->rng[1] = r->rng[1];
  if (isfinite(r->max_energy)) {
    double e;
    if (sys_plan_move(m, r->translation_scale, &e)) {
      if (e < r->max_energy) {
        sys_confirm(m);
        r->accepted_count += 1;
      } else {
        r->rejected_count += 1;
      }
    } else {
      r->rejected_count += 1;
    }
  } else {
    sys_randomize(m, m->rng);
    r->lowest_max_energy = r->max_energy;
  }
  r->rng[0] = m->rng[0];
  r->rng[1] = m->rng[1];
  double e = m->E;
  if (r->collecting_data) {
    if (e > r->cutoff_energy) {
      r->above_count += 1;
      r->above_total += e;
      r->above_total_squared += e * e;
      double pv;
      if (zmc_pressure(m, iter, &pv)) {  /* above_extra["pressure"] */
        r->pressure_total += pv;
        r->pressure_count += 1;
      }
    } else {
      r->below_count += 1;
      r->below_total += e;
      r->below_total_squared += e * e;
    }
    if (r->lowest_max_energy == very_lowest_max) r->upwelling_count += 1;
  }
}

/* MedianEstimator (energy_replicas.rs:47-96). */
static void zmed_add(or_zmc *z, double e) {
  if (z->med_len < 4096) {
    z->med[z->med_len++] = e;
  } else if (or_gen_f64(z->rng) < 1.0 / ((double)z->med_len + 1.0)) {
    uint64_t i = or_gen_range(z->rng, 0, (uint64_t)z->med_len);
    z->med[i] = e;
  }
}
static double zmed_median(or_zmc *z) {
  int n = z->med_len;
  qsort(z->med, n, sizeof(double), cmp_double);
  int middle = n / 2;
  double em = z->med[middle];
  for (int i = middle + 1; i < n; i++)
    if (z->med[i] != em) return 0.5 * (em + z->med[i]);
  for (int i = middle - 1; i >= 0; i--)
    if (z->med[i] != em) return 0.5 * (em + z->med[i]);
  return em;
}

/* Replica::decimate (energy_replicas.rs:169-192). */
static void zrep_decimate(or_zrep *r) {
  r->upwelling_count = 0;
  if (r->above_count > 1) {
    r->above_total /= (double)r->above_count;
    r->above_total_squared /= (double)r->above_count;
    r->above_count = 1;
  }
  if (r->below_count > 1) {
    r->below_total /= (double)r->below_count;
    r->below_total_squared /= (double)r->below_count;
    r->below_count = 1;
  }
  if (r->pressure_count > 1) {
    r->pressure_total /= (double)r->pressure_count;
    r->pressure_count = 1;
  }
  r->lowest_max_energy = r->max_energy;
  r->accepted_count = 1;
  r->rejected_count = 1;
  r->unique_visitors = 1;
}

/* Replica::occasional_update (energy_replicas.rs:246-290). */
static void zrep_occasional(or_zrep *r, const or_mc *m) {
  if (r->rejected_count > 128 && r->accepted_count > 128 && isfinite(r->max_energy)) {
    double ratio = (double)r->accepted_count / (double)r->rejected_count;
    double mx = (double)sys_moves_to_randomize(m);
    if (ratio < 0.5 || ratio > 2.0 * mx) {
      double adj = ratio < 0.5 ? ratio / sqrt(mx) : ratio * sqrt(mx);
      if (adj > 2.0) adj = 2.0;
      else if (adj < 0.5) adj = 0.5;
      r->translation_scale *= adj;
      r->accepted_count = 0;
      r->rejected_count = 0;
    }
  }
}

/* MC::run_once (energy_replicas.rs:504-641), without the plugins. */
void or_zmc_run(or_zmc *z, uint64_t rounds) {
  for (uint64_t round = 0; round < rounds; round++) {
    uint64_t steps = sys_moves_to_randomize(z->sys[0]);
    double lowest = z->r[z->R - 1].max_energy;
    uint64_t these = 0;
    for (int k = 0; k < z->R; k++) {
      these += steps;
      if (isfinite(z->r[k].max_energy)) {
        for (uint64_t i = 0; i < steps; i++) zrep_run_once(&z->r[k], z->sys[k], lowest, z->moves + i);
      } else {
        zrep_run_once(&z->r[k], z->sys[k], lowest, z->moves);
      }
    }
    int first = (or_rng_next(z->rng) >> 63) ? 0 : 1;
    for (int i = first; i + 1 < z->R; i += 2) {
      or_zrep *r0 = &z->r[i], *r1 = &z->r[i + 1];
      if (z->sys[i]->E < r1->max_energy) {
        or_mc *t = z->sys[i];
        z->sys[i] = z->sys[i + 1];
        z->sys[i + 1] = t;
        double l = r0->lowest_max_energy;
        r0->lowest_max_energy = r1->lowest_max_energy;
        r1->lowest_max_energy = l;
        r0->collecting_data = 1;
        r1->collecting_data = 1;
        if (r1->lowest_max_energy > r1->max_energy) {
          r1->unique_visitors += 1;
          r1->lowest_max_energy = r1->max_energy;
        }
      }
    }
    {
      or_zrep *last = &z->r[z->R - 1];
      double le = z->sys[z->R - 1]->E;
      if (le < last->cutoff_energy) zmed_add(z, le);
    }
    or_zrep *r = &z->r[z->R - 1];
    if (r->unique_visitors >= z->isbnb && r->lowest_max_energy == r->max_energy) {
      double mean_below = r->below_total / (double)r->below_count;
      if (mean_below + z->min_T < r->cutoff_energy && z->sys[z->R - 1]->E < r->cutoff_energy) {
        double median_below = zmed_median(z);
        z->med[0] = median_below;
        z->med_len = 1;
        /* the new replica is built from a copy taken before the arrays may
           move (r points into z->r) */
        or_zrep nr = *r;
        if (z->R == z->cap) {
          z->cap *= 2;
          z->sys = (or_mc **)realloc(z->sys, sizeof(or_mc *) * z->cap);
          z->r = (or_zrep *)realloc(z->r, sizeof(or_zrep) * z->cap);
        }
        nr.max_energy = nr.cutoff_energy;
        nr.cutoff_energy = median_below;
        zrep_decimate(&nr);
        nr.lowest_max_energy = -INFINITY;
        nr.translation_scale = nr.translation_scale * pow(0.5, 1.0 / (double)zmc_dim(z->sys[z->R - 1]));
        or_rng_jump(nr.rng);
        z->sys[z->R] = mc_clone(z->sys[z->R - 1]);
        z->r[z->R] = nr;
        z->R += 1;
      }
    }
    for (int k = 0; k < z->R; k++) zrep_occasional(&z->r[k], z->sys[k]);
    z->moves += these;
  }
}

int32_t or_zmc_replicas(const or_zmc *z) { return z->R; }

void or_zmc_replica(const or_zmc *z, int32_t k, or_zrep *out) {
  *out = z->r[k];
  out->E = z->sys[k]->E;
  out->error = z->sys[k]->error;
}

void or_zmc_coords(const or_zmc *z, int32_t k, double *c) { or_mc_coords(z->sys[k], c); }
void or_zmc_spins(const or_zmc *z, int32_t k, int8_t *s) { or_mc_spins(z->sys[k], s); }

void or_zmc_rng(const or_zmc *z, uint64_t out[2]) {
  out[0] = z->rng[0];
  out[1] = z->rng[1];
}

int32_t or_zmc_median(const or_zmc *z, double *out) {
  if (out) memcpy(out, z->med, sizeof(double) * z->med_len);
  return z->med_len;
}

uint64_t or_zmc_moves(const or_zmc *z) { return z->moves; }

void or_zmc_free(or_zmc *z) {
  if (!z) return;
  for (int k = 0; k < z->R; k++) or_mc_free(z->sys[k]);
  free(z->sys);
  free(z->r);
  free(z->med);
  free(z);
}
