/*
 * sadmc -- MI355X (gfx950) walker engine for the sad-monte-carlo hot path.
 *
 * C ABI.  Plain pointers, sizes and status codes; no C++ or torch types.
 * The reference's FFI-less Rust types this boundary stands in for are cited
 * on each entry point; a Rust caller declares these functions in an
 * extern "C" block and links libsadmc_amd.so.
 *
 * Model.  A context owns one HIP stream and any number of walker *sets*.  A
 * set is a batch of independent walkers of one system kind and size (one
 * reference `EnergyMC<S>` each, src/mc/energy.rs:167-210).  All walker state
 * -- positions or spins, the RNG stream, the SAD/WL/SAMC method state and the
 * per-walker energy bins -- lives in device memory for the life of the
 * context.  `sadmc_run` advances every walker of every set by K moves, i.e.
 * K calls of the reference's `EnergyMC::move_once` (energy.rs:904-974), in
 * one launch.
 *
 * Threading: one host thread per context; calls are not re-entrant.
 * Errors: every entry point returns SADMC_OK (0) or a negative code;
 * sadmc_last_error() gives the message.  Walker-level failures that the
 * reference turns into a panic (bin index out of range, energy drift
 * beyond verify_energy's bound) or that are specific to the fixed device bin
 * window (capacity) set that walker's `status`, stop that walker, and make
 * sadmc_sync() return the matching code.
 */
#ifndef SADMC_H
#define SADMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SADMC_ABI_VERSION 1

enum sadmc_status {
  SADMC_OK = 0,
  SADMC_E_ARG = -1,      /* bad argument */
  SADMC_E_HIP = -2,      /* HIP runtime error */
  SADMC_E_CAPACITY = -3, /* a walker's energy left its bin window */
  SADMC_E_VERIFY = -4,   /* verify_energy bound exceeded (lj.rs:249-261) */
  SADMC_E_INIT = -5,     /* initial configuration could not be built */
  SADMC_E_STATE = -6,    /* call out of order */
  SADMC_E_NODEVICE = -7, /* no usable gfx950 device */
  SADMC_E_INDEX = -8     /* energy index outside the bins (reference panics) */
};

/* Walker `status` values (sadmc_walker_state.status). */
enum { SADMC_WS_OK = 0, SADMC_WS_INDEX = 1, SADMC_WS_VERIFY = 2, SADMC_WS_INIT = 3, SADMC_WS_CAPACITY = 4 };

/* src/system/any.rs:61-162 -- the systems the `histogram` binary drives. */
enum sadmc_system { SADMC_ISING = 0, SADMC_LJ = 1, SADMC_WCA = 2 };
/* src/mc/energy.rs:44-69 MethodParams. */
enum sadmc_method { SADMC_SAD = 0, SADMC_SAMC = 1, SADMC_WL = 2, SADMC_INV_T_WL = 3, SADMC_CANONICAL = 4 };
/* src/mc/energy.rs:73-78 MoveParams. */
enum sadmc_move_plan { SADMC_TRANSLATION_SCALE = 0, SADMC_ACCEPTANCE_RATE = 1 };
/* Pair-term precision of the LJ energy difference. */
enum sadmc_precision { SADMC_FP64 = 0, SADMC_FP32_PAIRS = 1 };
/* Displacement proposal for LJ/WCA. */
enum sadmc_move_kind { SADMC_MOVE_CUBE = 0, SADMC_MOVE_LEARNED = 1 };
/* The Binning of binning-path walkers (src/mc/binning.rs:48-61). */
enum sadmc_binning { SADMC_BIN_HISTOGRAM = 0, SADMC_BIN_LINEAR = 1 };

/* Parameters shared by every walker of a set.  Mirrors the reference's
 * EnergyMCParams (energy.rs:81-97) plus the system parameters
 * (ising.rs:12-15, lj.rs:14-19, wca.rs:370-380). */
typedef struct {
  int32_t system;        /* sadmc_system */
  int32_t n;             /* Ising lattice width N (N*N spins, N <= 8), or atom count */
  int32_t method;        /* sadmc_method */
  int32_t move_plan;     /* sadmc_move_plan */
  double method_param;   /* SAD min_T | SAMC t0 | WL min_gamma | canonical T */
  int32_t has_min_gamma; /* WL only: min_gamma given */
  int32_t precision;     /* sadmc_precision (LJ only) */
  double move_param;     /* translation scale (cube half-width) or target acceptance rate */
  double energy_bin;     /* <= 0: system default (Ising 4, else 1) -- energy.rs:831-833 */
  int32_t has_min_e;     /* --min-allowed-energy given */
  int32_t has_max_e;     /* --max-allowed-energy given */
  double min_e;
  double max_e;
  double rc2;            /* WCA cutoff squared (2^(1/3) sigma^2) */
  int32_t refresh;       /* 0: reference error-estimate refresh; >0: FP64 refresh every `refresh` moves
                            (a refreshed energy past the bins' edge grows them, energy.rs:400-434) */
  int32_t ising_init;    /* 0: reference spins (seed 10137, ising.rs:45); 1: per-walker seed */
  int32_t collect_pressure; /* WCA: virial pressure every N^2 moves (wca.rs:202-218) */
  int32_t lanes;         /* lanes per walker for LJ/WCA (power of two 2..64); 0 = auto */
  double e_lo_bound;     /* bin window: lowest energy a walker may reach */
  double e_hi_bound;     /* bin window: highest energy a walker may reach */
  int32_t move_kind;     /* sadmc_move_kind */
  int32_t binning;       /* binning-path walkers: sadmc_binning (0 histogram) */
  double high_resolution_de; /* binning-path walkers: > 0 keeps the high_resolution side
                                histogram of this width (energy_binning.rs:63, 125) */
  double reserved[3];
} sadmc_set_params;

/* Device-resident per-walker state; `sadmc_download_raw` returns an array of
 * these.  Field meanings follow EnergyMC (energy.rs:167-244). */
typedef struct {
  uint64_t rng[2];            /* xoroshiro128+ state */
  double E;                   /* System::energy() */
  double error;               /* LJ/WCA accumulated error estimate */
  uint64_t moves, accepted_moves;
  double acceptance_rate, translation_scale;
  double bins_min, bins_width;
  int64_t base;               /* physical slot of logical bin 0 */
  int64_t len;                /* logical number of bins */
  double max_S;
  int64_t max_S_index;
  uint64_t epoch;             /* round-trip visit epoch */
  uint64_t bulk_visited;      /* value of have_visited for bins not marked this epoch */
  int32_t method;             /* current method (1/t-WL switches to SAMC) */
  int32_t status;             /* SADMC_WS_* */
  double too_lo, too_hi, latest_parameter, min_T;
  uint64_t tL, tF, num_states, highest_hist;
  uint64_t tmax;              /* max t_found over [too_lo, too_hi] (tracked incrementally) */
  double t0;                  /* SAMC */
  double wl_gamma, wl_num_states, wl_min_energy;
  uint64_t wl_lowest_hist, wl_highest_hist, wl_total_hist;
  int64_t wl_len;
  int32_t has_pressure;
  int32_t reserved0;
  double geom;                /* LJ sphere radius or WCA box edge */
  int64_t cap;                /* bin capacity */
  uint64_t reserved[3];
} sadmc_walker_state;

/* One energy bin (64 bytes).  The per-bin vectors of energy.rs:146-163 and
 * 202-205, interleaved so that one walker's touch of one bin is one half
 * cache line. */
typedef struct {
  double lnw, energy_total, energy_squared_total;
  uint64_t histogram, t_found, round_trips, visit_epoch, wl_hist;
} sadmc_bin;

/* One replica of a parallel-tempering ladder (tempering.rs:44-71). */
typedef struct {
  uint64_t rng[2];              /* the replica's generator */
  double E, error, T;           /* its system's energy and error estimate; its temperature */
  double total_energy, total_energy_squared;
  uint64_t accepted, rejected, ignored, accepted_swaps, rejected_swaps;
  double geom;                  /* sphere radius or box edge of the ladder */
  int32_t status, reserved0;
} sadmc_replica;

typedef struct sadmc_ctx sadmc_ctx;

/* ---- context --------------------------------------------------------- */
int sadmc_abi_version(void);
/* Create a context on HIP device `device`.  Fails with SADMC_E_NODEVICE when
 * no gfx950 device is present. */
int sadmc_create(sadmc_ctx **out, int device);
void sadmc_destroy(sadmc_ctx *ctx);
const char *sadmc_last_error(const sadmc_ctx *ctx);

/* ---- building walkers: MonteCarlo::from_params (energy.rs:830-898) ----
 * seeds[w]: the --seed of walker w.  geom[w]: LJ sphere radius or WCA box
 * edge (ignored for Ising; may be NULL then).  coords: optional initial
 * positions, n_walkers x [x[n], y[n], z[n]]; NULL generates them with the
 * documented init rule.  spins: optional n_walkers x n*n int8 (+1/-1). */
int sadmc_add_set(sadmc_ctx *ctx, const sadmc_set_params *p, int32_t n_walkers, const uint64_t *seeds,
                  const double *geom, const double *coords, const int8_t *spins, int32_t *set_id);
/* Allocate device state, upload, and run from_params on the device (initial
 * energy, optional greedy descent below max_e, one-bin histogram). */
int sadmc_init(sadmc_ctx *ctx);

/* ---- the hot path: K x move_once for every walker (energy.rs:904-974) - */
/* Enqueue one launch per set advancing every walker of every set by
 * `moves`.  With several sets the launches run side by side on up to four
 * streams forked from and joined back to the context's stream, so work
 * enqueued before or after the call stays ordered with all of them. */
int sadmc_run(sadmc_ctx *ctx, uint64_t moves);
/* Enqueue one launch advancing only the walkers of set `set` by `moves`
 * (move_once x moves for each of its walkers, energy.rs:904-974): the drain
 * a MonteCarlo shim shares between the walkers of one set. */
int sadmc_run_set(sadmc_ctx *ctx, int32_t set, uint64_t moves);
/* Wait for the stream; returns the first walker failure as an error code. */
int sadmc_sync(sadmc_ctx *ctx);
/* Duration in ms of the last launch of sadmc_run, sadmc_run_set,
 * sadmc_pt_run, sadmc_bin_run or sadmc_zmc_run (HIP events on the context
 * stream around the kernels; for sadmc_run with several sets, around the
 * fork and the join of its streams). */
int sadmc_last_run_ms(const sadmc_ctx *ctx, double *ms);

/* ---- mixed batches (walker sets of different systems and sizes) ------
 * One persistent launch runs every set (SAD, LJ/WCA, FP64 pair terms) for
 * `moves` moves: workgroups map to (set, walker range) tiles, heaviest
 * first.  Results are bit-identical to sadmc_run.  No reference counterpart
 * (the reference runs one walker per process, src/bin/histogram.rs:6-11). */
int sadmc_run_batched(sadmc_ctx *ctx, uint64_t moves);
/* The same launch captured once per move count in a HIP graph and replayed. */
int sadmc_graph_run(sadmc_ctx *ctx, uint64_t moves);
/* Lanes per walker: times `trial_moves` moves (after `warm_moves`) of a
 * scratch set for each power-of-two lane count that fits (<= 16 atoms per
 * lane) and returns the fastest.  rates[k] = moves/s at 2^k lanes (k = 0..6,
 * 0 where not tried); may be NULL. */
int sadmc_autotune_lanes(int device, const sadmc_set_params *p, int32_t n_walkers, const uint64_t *seeds,
                         const double *geom, uint64_t warm_moves, uint64_t trial_moves, int32_t *best_lanes,
                         double *rates);

/* Wavefronts per workgroup (1, 2 or 4) of set `set`'s run launch, or with
 * set = -1 of the batched launch (whose workgroup then runs that many
 * consecutive tiles); results do not depend on it.  SADMC_E_ARG beyond 64 KiB
 * of LDS per workgroup.  No reference counterpart (launch shape). */
int sadmc_set_block_waves(sadmc_ctx *ctx, int32_t set, int32_t waves);
/* Joint tuning of a mixed batch (no reference counterpart): the lanes per
 * walker of each of the nsets sets and the wavefronts per workgroup of the
 * batched launch, chosen by timing the whole batch (graph replay of
 * `trial_moves` moves after `warm_moves`): a coordinate descent over each
 * set's power-of-two lane counts (<= 16 atoms per lane), starting from
 * ps[k].lanes (0: the default), then 1, 2 or 4 wavefronts per workgroup.
 * seeds / geom: the sets' walkers concatenated in set order.  rate_out:
 * the chosen configuration's moves/s. */
int sadmc_autotune_batch(int device, int32_t nsets, const sadmc_set_params *ps, const int32_t *n_walkers,
                         const uint64_t *seeds, const double *geom, uint64_t warm_moves, uint64_t trial_moves,
                         int32_t *lanes_out, int32_t *waves_out, double *rate_out);

/* Diagnostic, no reference counterpart: the reciprocal the FP32 pair terms
 * use (the hardware reciprocal and one FMA Newton step) against the IEEE
 * quotient 1.0f / x, for every binary32 bit pattern x in [lo, hi)
 * (hi <= 2^32), on `device`.  what = 0: *mismatches = the number of x whose
 * two reciprocals differ bit for bit (NaN results count as equal); 0 over
 * [2^-126, 2^126).  what = 1: the number of x whose fast pair term is
 * neither NaN nor bit-identical to the division's; 0 over every
 * non-negative x, which is what lets the pair loops take the fast sum
 * unless it is NaN and recompute with the division then. */
int sadmc_check_recip_f32(int device, int32_t what, uint64_t lo, uint64_t hi, uint64_t *mismatches);

/* ---- single-walker pair for parity tests: MovableSystem::plan_move and
 * ConfirmSystem::confirm (system/mod.rs:101-120).  plan_move consumes the
 * walker's RNG exactly as inside move_once; *some = 0 means None. */
int sadmc_plan_move(sadmc_ctx *ctx, int32_t set, int32_t walker, double scale, double *e_new, int32_t *some);
int sadmc_confirm(sadmc_ctx *ctx, int32_t set, int32_t walker);

/* ---- learned proposal (no reference counterpart) ----------------------
 * LJ walker sets with move_kind = SADMC_MOVE_LEARNED propose Gaussian steps
 * whose scale comes from a small network of the moved atom's radial
 * features, with the two-sided Hastings term in the acceptance; the fixed
 * cube move (lj.rs:365-374 with a uniform step) stays the default.  The
 * network (1377 parameters, BF16 on the device, FP32 master copy) belongs to
 * the context and is shared by its learned sets. */
/* Single-walker learned plan: e_new, ln H and the moved atom's position. */
int sadmc_plan_learned(sadmc_ctx *ctx, int32_t set, int32_t walker, double *e_new, double *lnh, double *to,
                       int32_t *some);
/* Default initial weights from a seed; resets the optimiser. */
int sadmc_prop_init(sadmc_ctx *ctx, uint64_t seed);
int sadmc_prop_set_weights(sadmc_ctx *ctx, const float *weights, int32_t count);
int sadmc_prop_get_weights(sadmc_ctx *ctx, float *master, uint16_t *bf16);
/* Append per_walker probe records (24 floats: 8 features, 16 acceptance-
 * weighted squared jumps on the ln sigma grid) from every walker of an LJ
 * SAD set; walker w draws from the stream seeded by seed + w. */
int sadmc_prop_collect(sadmc_ctx *ctx, int32_t set, uint64_t seed, int32_t per_walker);
int sadmc_prop_records(sadmc_ctx *ctx, float *out, int64_t max_records, int64_t *count);
int sadmc_prop_load_records(sadmc_ctx *ctx, const float *records, int64_t count);
/* Mean loss and gradient over records [first, first+count) (cyclic). */
int sadmc_prop_gradient(sadmc_ctx *ctx, int64_t first, int32_t count, float *grad, float *loss);
/* `steps` Adam steps on consecutive minibatches; losses[k] = step k's loss. */
int sadmc_prop_train(sadmc_ctx *ctx, int32_t steps, int32_t batch, float lr, float *losses);

/* ---- mean entropy on the device (plotting/parse-binning.py:150-171) ---
 * S(E) of every SAD walker of a set (ln w inside [too_lo, too_hi], extended
 * below with slope 1/min_T and outside with ln(H/<H>)), each walker shifted
 * to mean 0 over the energy window [e_a, e_b] (either one NaN: the
 * intersection of the walkers' [too_lo, too_hi]; e_a >= e_b: SADMC_E_ARG),
 * averaged over the walkers covering each bin.  Writes bin k's centre as e_first + k * width, S[k] and count[k];
 * returns SADMC_E_CAPACITY (with *nbins set) if max_bins is too small. */
int sadmc_mean_entropy(sadmc_ctx *ctx, int32_t set, double e_a, double e_b, int64_t max_bins, double *e_first,
                       int64_t *nbins, double *S, int64_t *count);
/* The same reduction allocating nothing: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its scratch is a
 * fixed 32 MiB device array per GPU, held until the process ends, instead of
 * the context's growable buffer; a set whose reduction needs more (about
 * 96 bytes per walker plus 16 bytes per grid bin and chunk) is refused with
 * SADMC_E_CAPACITY.  Same arguments and results as sadmc_mean_entropy.  All
 * calls on one GPU share the array: calls that overlap on one GPU race. */
int sadmc_mean_entropy_noalloc(sadmc_ctx *ctx, int32_t set, double e_a, double e_b, int64_t max_bins,
                               double *e_first, int64_t *nbins, double *S, int64_t *count);

/* ---- EnergyMC over a histogram Binning (src/mc/energy_binning.rs with
 * src/mc/binning/histogram.rs: the `binning` binary) -----------------------
 * A separate kind of walker set: SAD (un-rounded too_lo/too_hi, num_states
 * by count_states, set_lnw zeroing counts), SAMC, WL and 1/t-WL over a
 * histogram (p->binning = SADMC_BIN_HISTOGRAM, binning/histogram.rs) or a
 * linear interpolation (SADMC_BIN_LINEAR, binning/linear.rs) of width
 * p->energy_bin (<= 0: 1.0, BinningParams' default) that grows at both ends
 * inside a fixed window of bins per walker.  Every
 * move accumulates the "energy" extra, SAD the "t_found" extra, WL the
 * "hist" extra, WCA with collect_pressure the "pressure" extra every N^2
 * moves, each with BinCounts' running statistics (histogram.rs:11-33).
 * Method from p->method (SADMC_CANONICAL rejected), FP64, refresh 0, cube
 * moves; p->min_e/max_e are min/max_allowed_energy (a max triggers the
 * greedy descent of from_params, energy_binning.rs:537-550). */
typedef struct {
  double min_total, max_total, e_max_total, e_max_count;
  uint64_t min_count, max_count, total_count;
  int32_t exists, reserved;
} sadmc_bcounts;

typedef struct {
  uint64_t rng[2];
  double E, error;
  uint64_t moves, accepted_moves;
  double acceptance_rate, translation_scale;
  double min, width, min_e, max_e; /* histogram::Bins (histogram.rs:98-111) */
  int64_t len;
  int32_t method, status;          /* SADMC_SAD/SAMC/WL (1/t-WL may switch to SAMC); SADMC_WS_* */
  uint64_t num_states, tL;
  double min_T, too_lo, too_hi, tF, latest_parameter, t0, gamma;
  sadmc_bcounts lnw, energy, t_found, hist, pressure;
  /* placement of the logical bins in the walker's window */
  int64_t base, cap;
  double geom;
  int32_t inv_t, has_min_gamma;
  double min_gamma;
  uint64_t reserved[3];
} sadmc_bwalker_state;

/* One logical bin: lnw and the extras' totals and counts.  With
 * SADMC_BIN_LINEAR the counts are fractional (linear.rs:24-31): every
 * *_count field here, and min_count / max_count of sadmc_bcounts, then hold
 * the bits of an IEEE double. */
typedef struct {
  double lnw_total;
  uint64_t lnw_count;
  double energy_total;
  uint64_t energy_count;
  double t_found_total;
  uint64_t t_found_count;
  uint64_t hist_count;
  double pressure_total;
  uint64_t pressure_count;
  uint64_t reserved;
} sadmc_bbin;

/* EnergyMC::from_params for W walkers (runs on the device at once). */
int sadmc_bin_add_set(sadmc_ctx *ctx, const sadmc_set_params *p, int32_t W, const uint64_t *seeds,
                      const double *geom, const double *coords, const int8_t *spins, int32_t *set_id);
/* move_once x moves for every walker of the set (synchronous; fails with
 * SADMC_E_CAPACITY / SADMC_E_VERIFY when a walker stopped). */
int sadmc_bin_run(sadmc_ctx *ctx, int32_t set, uint64_t moves);
int sadmc_bin_state(sadmc_ctx *ctx, int32_t set, int32_t w, sadmc_bwalker_state *out);
/* The logical bins [0, len) of walker w; SADMC_E_CAPACITY if max_bins < len. */
int sadmc_bin_bins(sadmc_ctx *ctx, int32_t set, int32_t w, int64_t max_bins, sadmc_bbin *out);
int sadmc_bin_coords(sadmc_ctx *ctx, int32_t set, int32_t w, double *coords);
int sadmc_bin_spins(sadmc_ctx *ctx, int32_t set, int32_t w, int8_t *spins);

/* EnergyMC::high_resolution (energy_binning.rs:125, 328-330, 568-570): a
 * histogram::Bins of width high_resolution_de made by Bins::new at the
 * walker's starting energy, and incremented with gamma 0 beside every ln w
 * update.  Its ln w totals stay 0 (so min_total = max_total = 0 and
 * e_max_total = -inf throughout); its counts and count statistics are kept
 * here.  Bins span the same energy window as the walker's own. */
typedef struct {
  double min, width;    /* histogram::Bins min and width */
  double min_e, max_e;  /* lowest and highest energy seen */
  int64_t len;          /* logical number of bins */
  int64_t base;         /* physical slot of logical bin 0 */
  uint64_t total_count, min_count, max_count;
  double e_max_count;
} sadmc_hires;

/* Walker w's side histogram: its statistics, and (counts non-null) the
 * counts of logical bins [0, len).  SADMC_E_ARG when the set was added
 * without high_resolution_de; SADMC_E_CAPACITY if max_bins < len. */
int sadmc_bin_hires(sadmc_ctx *ctx, int32_t set, int32_t w, sadmc_hires *out, int64_t max_bins, uint64_t *counts);

/* ---- energy-ceiling replicas (src/mc/energy_replicas.rs, "ZMC") ---------
 * `sims` independent runs of Ising, LJ or WCA systems (the replicas binary;
 * geom: the LJ sphere radius or the WCA box length of each run),
 * each on one wave with up to max_replicas replicas (the wave's lane
 * groups take the replicas in turn; a run's systems must fit 64 KiB of LDS).  from_params (energy_replicas.rs:344-395) runs on the host when
 * the set is added: 2^15 randomizations with seed_from_u64(seeds[s]), the
 * replicas [inf, E(1/2)] and [E(1/2), E(1/4)], the generator jumped, the
 * median estimator at E(1/4).  A round is MC::run_once
 * (energy_replicas.rs:504-641) without plugins: N (LJ, WCA) or N^2 (Ising)
 * moves per bounded replica, accepted iff E' < max_energy, cube moves of
 * half-width translation_scale (max_size at first: the radius, the box
 * diagonal, or 0.5), a randomization for the unbounded replica (WCA's adds
 * the atoms back one by one through its error estimate, wca.rs:251-267);
 * WCA replicas collect the pressure every N^2 moves while above their cutoff
 * (the "pressure" above_extra, wca.rs:202-218); neighbour swaps when the lower
 * system fits under the upper ceiling; the median estimator; a new replica
 * after `isbnb` (<= 0: 64) unique visitors when the mean below the cutoff is
 * min_T under it; occasional translation-scale updates.  The lowest replica
 * cannot split further once max_replicas exist: sadmc_zmc_run then fails
 * with SADMC_E_CAPACITY. */
typedef struct {
  double max_energy, cutoff_energy;
  uint64_t rejected_count, accepted_count, above_count, below_count, upwelling_count;
  double above_total, below_total, above_total_squared, below_total_squared;
  double lowest_max_energy; /* of the replica's system */
  uint64_t unique_visitors;
  int32_t collecting_data, reserved;
  uint64_t rng[2];
  double translation_scale;
  double E, error; /* the replica's system */
  double pressure_total;   /* above_extra["pressure"]: total and count (WCA) */
  uint64_t pressure_count;
} sadmc_zreplica;

int sadmc_zmc_add(sadmc_ctx *ctx, const sadmc_set_params *p, int32_t sims, const uint64_t *seeds,
                  const double *geom, const double *coords, const int8_t *spins, double min_T, int64_t isbnb,
                  int32_t max_replicas, int32_t *zmc_id);
int sadmc_zmc_run(sadmc_ctx *ctx, int32_t zmc, uint64_t rounds);
/* Replica count of run s, and up to max_out replica records, lowest index
 * (unbounded) first. */
int sadmc_zmc_replicas(sadmc_ctx *ctx, int32_t zmc, int32_t s, int32_t max_out, sadmc_zreplica *out,
                       int32_t *count);
int sadmc_zmc_coords(sadmc_ctx *ctx, int32_t zmc, int32_t s, int32_t replica, double *coords);
int sadmc_zmc_spins(sadmc_ctx *ctx, int32_t zmc, int32_t s, int32_t replica, int8_t *spins);
/* The run's own generator (2 words), its move count and its median
 * estimator's energies (up to 4096; *len set). */
int sadmc_zmc_state(sadmc_ctx *ctx, int32_t zmc, int32_t s, uint64_t *rng, uint64_t *moves, double *median,
                    int32_t *len);

/* ---- parallel tempering (src/mc/tempering.rs) ---------------------------
 * `ladders` independent ladders of `replicas` canonical replicas each, at
 * temperatures temps[0..replicas).  Ladder l starts every replica from one
 * configuration (coords/spins of ladder l, or the init rule with seeds[l])
 * and every replica's generator from seed_from_u64(seeds[l]); the ladder's
 * own generator is that one jumped by 2^64 (tempering.rs:147-158).  A round
 * (MC::run_once, tempering.rs:272-341) makes canonical_steps *
 * min_moves_to_randomize moves per replica (Replica::run_once, :96-112),
 * cube moves of half-width p->move_param (the reference's replicas keep a
 * translation scale of 1.0, tempering.rs:90; pass 1.0 for that), then tries
 * the odd or even neighbour swaps.  One workgroup per ladder: replicas * lanes <= 512;
 * the pair terms of the moves and initial energies are summed in lane-group
 * order.  Rejected: FP32 pair terms, refresh > 0, the learned move. */
int sadmc_pt_add(sadmc_ctx *ctx, const sadmc_set_params *p, int32_t ladders, int32_t replicas, const double *temps,
                 const uint64_t *seeds, const double *geom, const double *coords, const int8_t *spins,
                 int32_t canonical_steps, int32_t *pt_id);
int sadmc_pt_run(sadmc_ctx *ctx, int32_t pt, uint64_t rounds);
/* ladders * replicas records, ladder-major. */
int sadmc_pt_replicas(sadmc_ctx *ctx, int32_t pt, sadmc_replica *out);
int sadmc_pt_coords(sadmc_ctx *ctx, int32_t pt, int32_t ladder, int32_t replica, double *coords);
int sadmc_pt_spins(sadmc_ctx *ctx, int32_t pt, int32_t ladder, int32_t replica, int8_t *spins);
/* The ladders' own generators: 2 words per ladder. */
int sadmc_pt_ladder_rng(sadmc_ctx *ctx, int32_t pt, uint64_t *out);

/* ---- state: the serde view of EnergyMC (mc/mod.rs:111-120) ----------- */
int sadmc_walker_count(const sadmc_ctx *ctx, int32_t set, int32_t *n_walkers, int64_t *bin_capacity);
int sadmc_get_walker_state(sadmc_ctx *ctx, int32_t set, int32_t walker, sadmc_walker_state *out);
/* Logical bins [0, len): any output pointer may be NULL; `cap` bounds them. */
int sadmc_walker_bins(sadmc_ctx *ctx, int32_t set, int32_t walker, int64_t cap, double *lnw,
                      uint64_t *histogram, uint64_t *t_found, double *energy_total,
                      double *energy_squared_total, uint8_t *have_visited, uint64_t *round_trips,
                      uint64_t *wl_hist, uint64_t *pressure_count, double *pressure_total);
int sadmc_walker_coords(sadmc_ctx *ctx, int32_t set, int32_t walker, double *coords);
int sadmc_walker_spins(sadmc_ctx *ctx, int32_t set, int32_t walker, int8_t *spins);
/* Whole-set dumps: n_walkers states, n_walkers*cap bins. */
int sadmc_download_raw(sadmc_ctx *ctx, int32_t set, sadmc_walker_state *states, sadmc_bin *bins);
/* Resume (mc/mod.rs:70-84): overwrite a whole set's device state. */
int sadmc_upload_raw(sadmc_ctx *ctx, int32_t set, const sadmc_walker_state *states, const sadmc_bin *bins,
                     const double *coords, const int8_t *spins);

/* ---- System trait queries (src/system/mod.rs:54-98) ------------------- */
/* System::compute_energy (ising.rs:59-75, lj.rs:236-244, wca.rs:222-230):
 * the walker's energy recomputed from its configuration on the device, FP64,
 * pair terms in lane-group order. */
int sadmc_compute_energy(sadmc_ctx *ctx, int32_t set, int32_t walker, double *e);
/* System::verify_energy (lj.rs:249-261, wca.rs:237-250): recompute and
 * compare with the walker's tracked E against expected_accuracy (|E| 1e-14
 * N^2 for LJ, |E| 1e-13 N^2 for WCA; Ising has the trait's default, which
 * checks nothing).  *egood gets the recomputed energy, *ok 1 or 0; a failed
 * check returns SADMC_E_VERIFY (the reference panics). */
int sadmc_verify_energy(sadmc_ctx *ctx, int32_t set, int32_t walker, double *egood, int32_t *ok);
/* The constant trait queries of a system of this parameter set and geometry
 * (sphere radius or box edge): lowest_possible_energy (lj.rs:245-248,
 * wca.rs:234-236; none for Ising), delta_energy (ising.rs:76-78; none for
 * the continuous systems), min_moves_to_randomize and dimensionality
 * (ising.rs:86-91, lj.rs:280-285, wca.rs:268-273), max_size (ising.rs:120-122,
 * lj.rs:375-377, wca.rs:354-356).  No device needed. */
typedef struct {
  int32_t has_lowest_possible_energy, has_delta_energy;
  double lowest_possible_energy, delta_energy;
  uint64_t min_moves_to_randomize, dimensionality;
  double max_size;
} sadmc_system_info_t;
int sadmc_system_info(const sadmc_set_params *p, double geom, sadmc_system_info_t *out);
/* Resume of the WCA "pressure" extra (energy.rs:146-163 Bins::extra;
 * wca.rs:202-218): per physical bin, n_walkers x cap counts and totals in the
 * layout of sadmc_upload_raw's bins.  SADMC_E_ARG for a set without
 * collect_pressure. */
int sadmc_upload_pressure(sadmc_ctx *ctx, int32_t set, const uint64_t *count, const double *total);

/* ---- host-side helpers (no device needed) ----------------------------- */
/* The initial configurations sadmc_add_set generates when coords/spins are
 * NULL.  Returns SADMC_E_INIT if LJ placement fails. */
int sadmc_init_coords(int32_t system, int32_t n, double geom, uint64_t seed, double *coords);
int sadmc_init_spins(int32_t n, int32_t rule, uint64_t seed, int8_t *spins);
/* Bin window (capacity, e_lo, e_hi) chosen for a parameter set when
 * e_lo_bound == e_hi_bound. */
int sadmc_bin_window(const sadmc_set_params *p, double geom, double *e_lo, double *e_hi, int64_t *cap);

#ifdef __cplusplus
}
#endif
#endif
