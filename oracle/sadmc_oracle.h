/*
 * CPU oracle for the sad-monte-carlo walker hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's per-walker loop (src/mc/energy.rs:904-974) and the three
 * systems it drives (src/system/ising.rs, lj.rs, wca.rs), together with the
 * random-number stream (rand_xoshiro 0.4 Xoroshiro128Plus seeded by
 * SplitMix64, restated step at src/rng.rs:48-57) and the rand 0.7 adapters
 * the reference calls.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg load it, and only as the checker.  The product library
 * never links it.
 *
 * Summation order.  `order_g == 0` follows the reference's sequential sums
 * (lj.rs:93-102, lj.rs:236-244, wca.rs:119-140, wca.rs:222-230).
 * `order_g == G > 0` sums in the order a group of G GPU lanes does: lane l
 * accumulates atoms l, l+G, l+2G, ... and the G partial sums are combined by
 * an xor butterfly.  Both orders are the same mathematics; they differ only
 * in rounding, and tests check that they agree to ~1e-12.
 *
 * Pinning (see tests/test_oracle_*.py): SplitMix64 against the published
 * seed-0 vector; the xoroshiro128 linear engine against its full period and
 * the published jump polynomial; Ising against brute force and the exact
 * 8x8 density of states; LJ against closed-form potentials and the known
 * cluster minima; SAD gamma against its closed form; the SAD/WL/SAMC
 * bookkeeping against hand-built scripted-energy trajectories; the
 * precision variant's pair term against numpy's binary32 arithmetic and its
 * rounding bound, its refreshes against long-double recomputes, and the bin
 * growth a refresh past the bins' edge causes.  The rand 0.7
 * integer/float adapters (gen_range, Uniform) are restated from the rand
 * 0.7.3 sources as recalled -- parity unpinned against the Rust crate,
 * which cannot be built here.
 */
#ifndef SADMC_ORACLE_H
#define SADMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers -------------------------------------------------- */
uint64_t or_splitmix64_next(uint64_t *x);
void or_rng_seed(uint64_t seed, uint64_t s[2]);
uint64_t or_rng_next(uint64_t s[2]);
uint64_t or_gen_range(uint64_t s[2], uint64_t low, uint64_t high);
uint64_t or_uniform_index(uint64_t s[2], uint64_t n);
double or_gen_f64(uint64_t s[2]);
double or_uniform_f64(uint64_t s[2], double low, double high);

/* ---- potentials ------------------------------------------------------- */
double or_lj_pair(double r2);
float or_lj_pair_f32(float r2);
double or_wca_pair(double r2, double rc2);

/* ---- walker ----------------------------------------------------------- */
enum { OR_ISING = 0, OR_LJ = 1, OR_WCA = 2, OR_SCRIPTED = 3 };
enum { OR_SAD = 0, OR_SAMC = 1, OR_WL = 2, OR_INV_T_WL = 3, OR_CANONICAL = 4 };

typedef struct {
  int32_t system;
  int32_t n;
  int32_t method;
  int32_t move_plan;        /* 0 TranslationScale, 1 AcceptanceRate */
  double method_param;      /* SAD min_T, SAMC t0, canonical T, WL min_gamma */
  int32_t has_min_gamma;
  int32_t precision;        /* 0 fp64 pairs, 1 fp32 pairs + fp64 sums */
  double move_param;        /* translation scale, or target acceptance rate */
  double energy_bin;        /* <= 0: system default */
  int32_t has_min_e;
  int32_t has_max_e;
  double min_e;
  double max_e;
  double radius;            /* LJ sphere radius */
  double box;               /* WCA cube edge */
  double rc2;               /* WCA cutoff squared */
  int32_t refresh;          /* 0: reference error model; >0 fixed interval */
  int32_t order_g;          /* 0 reference order; G>0 lane-group order */
  int32_t ising_init;       /* 0 reference seed 10137; 1 per-walker seed */
  int32_t collect_pressure; /* WCA pressure every N^2 moves */
  uint64_t seed;
  int32_t move_kind;        /* 0 cube move, 1 learned Gaussian move (LJ) */
  int32_t drop_hastings;    /* test hook: omit ln H (a planted mistake) */
  const float *mlp;         /* learned move: OR_PROP_NP weights (BF16 values as float) */
  int32_t binning;          /* binning walker: 0 histogram (histogram.rs), 1 linear (linear.rs) */
  double high_resolution_de; /* binning walker: > 0 keeps the high_resolution side histogram */
} or_params;

typedef struct {
  double E;
  double error;
  uint64_t moves;
  uint64_t accepted;
  double acceptance_rate;
  double translation_scale;
  uint64_t rng[2];
  double bins_min;
  double bins_width;
  int64_t len;
  double max_S;
  int64_t max_S_index;
  int32_t method;           /* may switch 1/t-WL -> SAMC */
  int32_t status;
  /* SAD */
  double too_lo, too_hi, lambda, min_T;
  uint64_t tL, tF, num_states, highest_hist;
  /* SAMC */
  double t0;
  /* WL */
  double wl_gamma, wl_num_states, wl_min_energy;
  uint64_t wl_lowest, wl_highest, wl_total;
  int64_t wl_len;
  int32_t has_pressure;
} or_scalars;

typedef struct or_mc or_mc;

/* Build a walker.  coords: SoA x[n],y[n],z[n] (NULL: generate with the
 * documented init rule); spins: n*n int8 (NULL: generate). */
or_mc *or_mc_new(const or_params *p, const double *coords, const int8_t *spins);
void or_mc_free(or_mc *mc);
int or_mc_run(or_mc *mc, uint64_t moves);
void or_mc_scalars(const or_mc *mc, or_scalars *out);
/* Copy bins (logical order, len entries).  Any pointer may be NULL. */
void or_mc_bins(const or_mc *mc, double *lnw, uint64_t *hist, uint64_t *t_found,
                double *e_tot, double *e_sq, uint8_t *visited, uint64_t *round_trips,
                uint64_t *wl_hist, uint64_t *p_count, double *p_total);
void or_mc_coords(const or_mc *mc, double *coords);
void or_mc_spins(const or_mc *mc, int8_t *spins);
double or_mc_compute_energy(const or_mc *mc);     /* order-aware, fp64 */
double or_mc_reference_energy(const or_mc *mc);   /* reference order */
int or_mc_plan_move(or_mc *mc, double scale, double *e_new);
void or_mc_confirm(or_mc *mc);
double or_mc_energy(const or_mc *mc);

/* Per-move trace: energy after each move and whether it was accepted.
 * or_mc_trace arms a buffer for the next cap moves; or_mc_trace_copy
 * copies what was recorded and returns the count. */
void or_mc_trace(or_mc *mc, uint64_t cap);
uint64_t or_mc_trace_copy(const or_mc *mc, double *energies, uint8_t *accepted);
/* The last move's decision inputs: out[0] the proposed energy (NaN: None),
 * out[1] the log acceptance ratio of a learned move's Metropolis-Hastings
 * test and out[2] its uniform (NaN: none drawn), out[3] |to|^2 - R^2 of a
 * learned move's wall test.  Lets a float-tolerance comparison tell a
 * decision taken inside rounding of its threshold from a real mismatch. */
void or_mc_last_move(const or_mc *mc, double *out);
/* Overwrite an LJ/WCA SAD walker's whole state -- scalars, the len logical
 * bins, positions (SoA) -- with another engine's (test hook: a comparison
 * at float tolerance resynchronises the oracle on the device walker after
 * the two part at a decision within rounding of its threshold).  Returns -1
 * for other systems or methods. */
int or_mc_load_state(or_mc *mc, const or_scalars *s, const double *lnw, const uint64_t *hist,
                     const uint64_t *t_found, const double *e_tot, const double *e_sq, const uint8_t *visited,
                     const uint64_t *round_trips, const double *coords);

/* Scripted system: plan_move returns script[k % len] (NaN = None). */
or_mc *or_mc_new_scripted(const or_params *p, double e0, const double *script, int64_t len);

/* System-level helpers for invariant tests. */
void or_ising_init_spins(int32_t n, int32_t rule, uint64_t seed, int8_t *spins);
double or_ising_energy(int32_t n, const int8_t *spins);
int or_lj_init(int32_t n, double radius, uint64_t seed, double *coords);
int or_wca_init(int32_t n, double box, uint64_t seed, double *coords);
double or_lj_energy(int32_t n, const double *coords, int32_t order_g);
double or_wca_energy_celllist(int32_t n, const double *coords, double box, double rc2);
double or_wca_energy(int32_t n, const double *coords, double box, double rc2, int32_t order_g);
double or_wca_energy_ld(int32_t n, const double *coords, double box, double rc2);
double or_lj_energy_ld(int32_t n, const double *coords);
double or_sad_gamma(double lambda, double t, double tF, double num_states);

/* ---- learned proposal (no reference counterpart) ----------------------
 * Features of atom i at position p: f_k = sum_{j != i} exp(-(r_ij - c_k)^2 /
 * (2 w^2)), c_k = 0.9 + k (1.6/7), w = 0.2, k = 0..7, in float with r_ij =
 * sqrtf of the float r^2 of the double differences.  MLP 8 -> 32 -> 32 -> 1,
 * ReLU, weights in the order W1[32][8], b1[32], W2[32][32], b2[32], w3[32],
 * b3; the output s = ln(sigma) is clamped to [ln 0.01, 0].  Move: atom i
 * uniform, xi ~ N(0, I3) by Box-Muller from four gen_f64 draws, p' = p +
 * sigma(p) xi; ln H = 3 (s - s') + |xi|^2 (1 - exp(2 (s - s'))) / 2 (the
 * two-sided Gaussian proposal ratio q(p'->p)/q(p->p')), added to the
 * acceptance exponent.  Probe samples for training: features of a random
 * atom and, for each of the 16 scales ln sigma_m = ln 0.01 + m (ln 100)/15,
 * the mean over 2 Gaussian draws of alpha |d|^2 (alpha = the walker's
 * acceptance probability of that displacement, Hastings term excluded). */
#define OR_PROP_NF 8
#define OR_PROP_H 32
#define OR_PROP_NP 1377
#define OR_PROP_NS 16
#define OR_PROP_R 2
#define OR_PROP_REC 24 /* floats per probe record: f[8], J[16] */
void or_prop_features(int32_t n, const double *coords, int32_t skip, double px, double py, double pz, float *f);
float or_prop_mlp(const float *w, const float *f);
void or_prop_gaussian(uint64_t rng[2], float *xi);
double or_prop_lnh(float s, float s_new, float xi2);
/* Training loss -mean_n Jhat_n(s_n)/max_m J_n,m and its gradient (float
 * forward, double accumulation).  At a clamp of the output the gradient is
 * projected: kept only when it points back into the range.  grad may be
 * NULL. */
double or_prop_loss_grad(const float *w, const float *records, int64_t count, double *grad);
/* The same in double throughout (for finite-difference checks). */
double or_prop_loss_grad_f64(const double *w, const float *records, int64_t count, double *grad);
/* `count` probe records of walker `mc` with the probe stream seeded by seed. */
void or_prop_probe(const or_mc *mc, uint64_t seed, int32_t count, float *out);
/* Learned move of walker `mc` (plan only): returns 0 if the wall rejects;
 * e_new, ln H and the moved atom's new position are written. */
int or_mc_plan_learned(or_mc *mc, double *e_new, double *lnh, double *to);

/* ---- parallel tempering (src/mc/tempering.rs) --------------------------
 * One ladder of R replicas at temperatures T[0..R), all starting from the
 * same configuration with clones of one generator seeded by `seed`
 * (tempering.rs:147-158); the ladder's own generator is that one after a
 * 2^64-step jump.  A round (run_once, tempering.rs:272-341) makes
 * canonical_steps * min_moves_to_randomize moves per replica (accept when
 * beta dE < 0 or u < exp(-beta dE), tempering.rs:96-112), then, with pairs
 * (0,1),(2,3),... or (1,2),(3,4),... chosen by the ladder generator's
 * bool, swaps the systems of a pair when d = (E0 - E1)(1/T0 - 1/T1) >= 0
 * or u < exp(d) with u from the second replica's generator. */
typedef struct {
  double E, error, T, total_energy, total_energy_squared;
  uint64_t accepted, rejected, ignored, accepted_swaps, rejected_swaps;
  uint64_t rng[2];
} or_replica;
typedef struct or_pt or_pt;
void or_rng_jump(uint64_t s[2]);
or_pt *or_pt_new(const or_params *p, int32_t replicas, const double *T, uint64_t seed, const double *coords,
                 const int8_t *spins, int32_t canonical_steps);
void or_pt_run(or_pt *pt, uint64_t rounds);
void or_pt_replica(const or_pt *pt, int32_t r, or_replica *out);
void or_pt_coords(const or_pt *pt, int32_t r, double *coords);
void or_pt_spins(const or_pt *pt, int32_t r, int8_t *spins);
void or_pt_rng(const or_pt *pt, uint64_t out[2]);
void or_pt_free(or_pt *pt);

/* ---- EnergyMC over a histogram Binning (src/mc/energy_binning.rs with
 * src/mc/binning/histogram.rs) --------------------------------------------
 * The `binning` binary's walker: SAD (un-rounded too_lo/too_hi, num_states
 * from count_states, set_lnw zeroing counts), SAMC, WL and 1/t-WL over a
 * histogram of width energy_bin (<= 0: 1.0) that grows at both ends, with
 * the "energy" extra every move, the "t_found" extra (SAD), the "hist" extra
 * (WL) and the WCA "pressure" extra every N^2 moves.  p->binning = 1
 * selects the linear interpolation of binning/linear.rs (fractional counts,
 * exported as the bits of doubles in the count fields).  The method comes from
 * p->method (OR_CANONICAL is not a method of this walker); p->max_e /
 * p->min_e are max/min_allowed_energy; the generator is seed_from_u64(seed)
 * and a max_allowed_energy triggers the greedy descent of from_params. */
typedef struct {
  double min_total, max_total, e_max_total, e_max_count;
  uint64_t min_count, max_count, total_count;
  int32_t exists, reserved;
} or_bcounts;
typedef struct {
  uint64_t rng[2];
  double E, error;
  uint64_t moves, accepted;
  double acceptance_rate, translation_scale;
  double min, width, min_e, max_e;
  int64_t len;
  int32_t method, status;
  uint64_t num_states, tL;
  double min_T, too_lo, too_hi, tF, latest_parameter, t0, gamma;
  or_bcounts lnw, energy, t_found, hist, pressure;
} or_bstate;
typedef struct or_bmc or_bmc;
or_bmc *or_bmc_new(const or_params *p, const double *coords, const int8_t *spins);
or_bmc *or_bmc_new_scripted(const or_params *p, double e0, const double *script, int64_t len);
int or_bmc_run(or_bmc *b, uint64_t moves);
void or_bmc_state(const or_bmc *b, or_bstate *out);
void or_bmc_bins(const or_bmc *b, double *lnw_total, uint64_t *lnw_count, double *energy_total,
                 uint64_t *energy_count, double *tf_total, uint64_t *tf_count, uint64_t *hist_count,
                 double *p_total, uint64_t *p_count);
double or_bmc_lnw_at(const or_bmc *b, double e);
double or_bmc_count_at(const or_bmc *b, double e);
void or_bmc_coords(const or_bmc *b, double *coords);
void or_bmc_spins(const or_bmc *b, int8_t *spins);
void or_bmc_free(or_bmc *b);
/* The high_resolution side histogram (energy_binning.rs:125, 328-330):
 * stats[10] = min, width, min_e, max_e, len, total_count, min_count,
 * max_count, e_max_count, min_total (as doubles); counts[len] when non-null.
 * Returns len, or -1 when the walker keeps none. */
int64_t or_bmc_hires(const or_bmc *b, double *stats, uint64_t *counts);

/* ---- energy-ceiling replicas (src/mc/energy_replicas.rs, "ZMC") --------
 * Ising, LJ or WCA systems.  from_params (energy_replicas.rs:344-395): 2^15
 * randomizations of the system with the generator seed_from_u64(seed), a
 * second randomized system for the unbounded replica, the starting system
 * randomized until its energy is below the 2^14-th sampled energy (taken
 * before the sort, as the reference does), replicas [inf, E(1/2)] and
 * [E(1/2), E(1/4)] of the sorted energies with clones of the generator,
 * which is then jumped; the median estimator starts at E(1/4).  A round is
 * MC::run_once (energy_replicas.rs:504-641) without plugins: every bounded
 * replica makes min_moves_to_randomize moves (accept iff E' < max_energy,
 * moves of half-width translation_scale, which starts at max_size), the
 * unbounded one randomizes once; neighbour swaps when the lower system fits
 * under the upper replica's ceiling; the median estimator; a new replica
 * after `isbnb` unique visitors; occasional translation-scale updates. */
typedef struct {
  double max_energy, cutoff_energy;
  uint64_t rejected_count, accepted_count, above_count, below_count, upwelling_count;
  double above_total, below_total, above_total_squared, below_total_squared;
  double lowest_max_energy;
  uint64_t unique_visitors;
  int32_t collecting_data, reserved;
  uint64_t rng[2];
  double translation_scale;
  double E, error;  /* the replica's system */
  double pressure_total;   /* above_extra["pressure"] (WCA): total and count */
  uint64_t pressure_count;
} or_zrep;
typedef struct or_zmc or_zmc;
or_zmc *or_zmc_new(const or_params *p, uint64_t seed, double min_T, uint64_t isbnb, const double *coords,
                   const int8_t *spins);
void or_zmc_run(or_zmc *z, uint64_t rounds);
int32_t or_zmc_replicas(const or_zmc *z);
void or_zmc_replica(const or_zmc *z, int32_t k, or_zrep *out);
void or_zmc_coords(const or_zmc *z, int32_t k, double *coords);
void or_zmc_spins(const or_zmc *z, int32_t k, int8_t *spins);
void or_zmc_rng(const or_zmc *z, uint64_t out[2]);
int32_t or_zmc_median(const or_zmc *z, double *out);
uint64_t or_zmc_moves(const or_zmc *z);
void or_zmc_free(or_zmc *z);

/* CPU baseline: run `threads` independent walkers (one per thread) for
 * `moves` moves each; returns wall seconds (init excluded). */
double or_bench(const or_params *p, int32_t threads, uint64_t moves, uint64_t seed0);

#ifdef __cplusplus
}
#endif
#endif
