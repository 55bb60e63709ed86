// Host side of the C ABI (include/sadmc.h): walker sets, device buffers,
// initial configurations, kernel dispatch and state transfer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sadmc_kernels.h"
#include "sadmc_internal.h"

using namespace sadmc;

static int method_family(int method) {
  switch (method) {
    case SADMC_SAD: return M_SAD;
    case SADMC_SAMC: return M_SAMC;
    case SADMC_WL:
    case SADMC_INV_T_WL: return M_WL;
    default: return M_CAN;
  }
}

static hipError_t launch(int sys, int prec, int method, int kmax, KernelKind kk, const SetArgs &a,
                         const LaunchArgs &la, hipStream_t s) {
  const int m = method_family(method);
  if (sys == SYS_ISING) return launch_ising(m, kk, a, la, s);
  if (sys == SYS_LJ && a.move_kind == SADMC_MOVE_LEARNED && (kk == KK_RUN || kk == KK_PLAN || kk == KK_PROBE)) {
    if (prec == SADMC_FP32_PAIRS) return launch_learned32(m, kmax, kk, a, la, s);
    return launch_learned64(m, kmax, kk, a, la, s);
  }
  if (sys == SYS_LJ && kk == KK_PROBE) {
    if (prec == SADMC_FP32_PAIRS) return launch_learned32(m, kmax, kk, a, la, s);
    return launch_learned64(m, kmax, kk, a, la, s);
  }
  if (sys == SYS_LJ) {
    if (prec == SADMC_FP32_PAIRS) return launch_lj32(m, kmax, kk, a, la, s);
    return launch_lj64(m, kmax, kk, a, la, s);
  }
  return launch_wca(m, kmax, kk, a, la, s);
}

// ============================================================ host side

namespace {

struct SetHost {
  sadmc_set_params p;
  int32_t W = 0;
  int G = 1, K = 1, KMAX = 1, npad = 0;
  int64_t cap = 0;
  double e_lo = 0, e_hi = 0, width = 1;
  std::vector<uint64_t> seeds;
  std::vector<double> geom;
  std::vector<double> coords;  // W * 3 * n, unpadded
  std::vector<int8_t> spins;   // W * n * n
  sadmc_walker_state *d_ws = nullptr;
  double *d_pos = nullptr;
  uint64_t *d_spins = nullptr;
  HotBin *d_hot = nullptr;
  ColdBin *d_cold = nullptr;
  uint64_t *d_wlh = nullptr;
  PExtra *d_extra = nullptr;
  Pending *d_pending = nullptr;
  int32_t *d_flag = nullptr;
  double *d_eout = nullptr;  // a recomputed energy (sadmc_compute_energy)
  int waves = 1;             // wavefronts per workgroup of the run kernel
  SetArgs args;
};

// A parallel-tempering ladder set on the device.
struct PtHost {
  int L = 0, R = 0, G = 1, K = 1, KMAX = 1, npad = 0, threads = 64, sys = 0;
  SetArgs args;
  PtArgs pa;
  sadmc_replica *d_rep = nullptr;
  double *d_slot_e = nullptr, *d_pos = nullptr, *d_geom = nullptr;
  int32_t *d_slot_of = nullptr;
  uint64_t *d_spins = nullptr, *d_lrng = nullptr;
};

// Energy-ceiling replica runs on the device.
struct ZmcHost {
  int S = 0, RM = 0, G = 1, K = 1, KMAX = 1, npad = 0, threads = 64, sys = 0, n = 0;
  SetArgs args;
  ZArgs za;
  ZSim *d_sim = nullptr;
  sadmc_zreplica *d_rep = nullptr;
  double *d_slot = nullptr, *d_med = nullptr, *d_pos = nullptr, *d_geom = nullptr;
  int32_t *d_slot_of = nullptr;
  uint64_t *d_spins = nullptr;
};

// A walker set of the binning path on the device.
struct BinSetHost {
  sadmc_set_params p;
  int32_t W = 0;
  int G = 1, K = 1, KMAX = 1, npad = 0, sys = 0;
  int64_t cap = 0;
  SetArgs args;
  BinArgs ba;
  sadmc_bwalker_state *d_ws = nullptr;
  sadmc_bbin *d_bins = nullptr;
  double *d_pos = nullptr;
  uint64_t *d_spins = nullptr;
  int32_t *d_flag = nullptr;
  int64_t hr_cap = 0;            // side histogram bins per walker (0: none)
  sadmc_hires *d_hrs = nullptr;
  uint64_t *d_hr = nullptr;
};

}  // namespace

constexpr int kSideStreams = 4;

struct sadmc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<SetHost> sets;
  std::string err;
  bool inited = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  // sadmc_run with several sets: their launches fork from `stream` onto up
  // to kSideStreams streams (default priority) and join back, so sets that
  // each fill only part of the GPU run side by side
  hipStream_t side[kSideStreams] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kSideStreams] = {};
  // batched launch: per-set arguments and the block -> tile table on the device
  SetArgs *d_sets = nullptr;
  BatchTile *d_tiles = nullptr;
  int ntiles = 0;
  size_t batch_lds = 0;  // LDS per wavefront of the batched launch
  int batch_waves = 1;   // wavefronts per workgroup of the batched launch
  std::map<uint64_t, hipGraphExec_t> graphs;  // captured batched launch per move count
  std::vector<hipGraph_t> graph_defs;
  // learned proposal: FP32 master weights, Adam moments, BF16 copy the
  // kernels read; probe records; training scratch
  float *d_pw = nullptr, *d_pm = nullptr, *d_pv = nullptr;
  uint16_t *d_pbf = nullptr;
  int prop_step = 0;
  int64_t prop_cursor = 0;
  float *d_rec = nullptr;
  int64_t nrec = 0, rec_cap = 0;
  float *d_partial = nullptr;
  int partial_blocks = 0;
  float *d_grad = nullptr, *d_loss = nullptr;
  int loss_cap = 0;
  std::vector<PtHost> pts;
  std::vector<BinSetHost> bsets;
  std::vector<ZmcHost> zmcs;
  // entropy reduction scratch, grown on demand and kept for later calls
  void *d_ent = nullptr;
  size_t ent_bytes = 0;
};

static int fail(sadmc_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHECK(ctx, call)                                                                          \
  do {                                                                                               \
    hipError_t _e = (call);                                                                          \
    if (_e != hipSuccess) return fail(ctx, SADMC_E_HIP, std::string(#call ": ") + hipGetErrorString(_e)); \
  } while (0)

// Frees the device buffers of a set under construction when an error
// returns early, unless release() is called once the set is complete.
struct AllocGuard {
  std::vector<void *> ptrs;
  bool armed = true;
  template <typename T>
  hipError_t alloc(T **p, size_t bytes) {
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e == hipSuccess) ptrs.push_back((void *)*p);
    return e;
  }
  void release() { armed = false; }
  ~AllocGuard() {
    if (armed)
      for (void *p : ptrs) (void)hipFree(p);
  }
};

// ---------------------------------------------------------------- init rules

// LJ initial positions (uniform in the sphere, all pairs >= 0.8 sigma):
// atoms are added one at a time, each drawn by rejection from [-R, R)^3 and
// redrawn until it is inside the sphere and 0.8 sigma from every earlier atom.
static int lj_init_coords(int n, double radius, uint64_t seed, double *c) {
  uint64_t s0, s1;
  rng_seed(seed ^ kInitTag, s0, s1);
  double *x = c, *y = c + n, *z = c + 2 * n;
  const double R2 = radius * radius, dmin2 = 0.8 * 0.8;
  const double sc = uniform_scale(-radius, radius);
  for (int a = 0; a < n; a++) {
    uint64_t tries = 0;
    for (;;) {
      if (++tries > 100000000ULL) return SADMC_E_INIT;
      double px = uniform_sample(s0, s1, -radius, sc);
      double py = uniform_sample(s0, s1, -radius, sc);
      double pz = uniform_sample(s0, s1, -radius, sc);
      if (norm2(px, py, pz) >= R2) continue;
      bool ok = true;
      for (int b = 0; b < a && ok; b++) ok = norm2(px - x[b], py - y[b], pz - z[b]) >= dmin2;
      if (!ok) continue;
      x[a] = px;
      y[a] = py;
      z[a] = pz;
      break;
    }
  }
  return SADMC_OK;
}

// WCA initial positions: of the simple-cubic, bcc and fcc lattices with the
// fewest cells per edge m giving >= n sites, the one with the largest
// nearest-neighbour distance (factor / m; ties to fcc, then bcc); n of its
// sites chosen by a partial Fisher-Yates shuffle.
static int wca_init_coords(int n, double box, uint64_t seed, double *c) {
  static const double basis[3][4][3] = {
      {{0.0, 0.0, 0.0}, {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}},
      {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}},
      {{0.0, 0.0, 0.0}}};
  static const int nbasis[3] = {4, 2, 1};
  static const double nnf[3] = {0.7071067811865476, 0.8660254037844386, 1.0};
  uint64_t s0, s1;
  rng_seed(seed ^ kInitTag, s0, s1);
  int best = 0, bm = 1;
  double bestnn = -1.0;
  for (int t = 0; t < 3; t++) {
    int m = 1;
    while ((int64_t)nbasis[t] * m * m * m < n) m++;
    double nn = nnf[t] / (double)m;
    if (nn > bestnn) {
      bestnn = nn;
      best = t;
      bm = m;
    }
  }
  const int m = bm, nb = nbasis[best];
  const int64_t ns = (int64_t)nb * m * m * m;
  std::vector<double> sx(ns), sy(ns), sz(ns);
  const double a = box / (double)m;
  int64_t k = 0;
  for (int ix = 0; ix < m; ix++)
    for (int iy = 0; iy < m; iy++)
      for (int iz = 0; iz < m; iz++)
        for (int b = 0; b < nb; b++) {
          sx[k] = ((double)ix + basis[best][b][0]) * a;
          sy[k] = ((double)iy + basis[best][b][1]) * a;
          sz[k] = ((double)iz + basis[best][b][2]) * a;
          k++;
        }
  for (int64_t i = 0; i < n; i++) {
    int64_t r = i + (int64_t)uniform_index(s0, s1, (uint64_t)(ns - i));
    std::swap(sx[i], sx[r]);
    std::swap(sy[i], sy[r]);
    std::swap(sz[i], sz[r]);
  }
  for (int i = 0; i < n; i++) {
    c[i] = sx[i];
    c[n + i] = sy[i];
    c[2 * n + i] = sz[i];
  }
  return SADMC_OK;
}

// Known global minima of the free LJ clusters (Wales & Doye 1997); the
// confined cluster cannot go below them.
static double lj_lowest(int n) {
  switch (n) {
    case 13: return -44.326801;
    case 19: return -72.659782;
    case 31: return -133.586422;
    case 38: return -173.928427;
    default: return -0.5 * n * (n - 1.0);  // lj.rs:245-248
  }
}

// Assemble the public 64-byte records of walker w's physical slots
// [first, first + count) from the hot, cold and WL arrays.
static int gather_bins(sadmc_ctx *c, const SetHost &s, int w, int64_t first, int64_t count, sadmc_bin *out) {
  std::vector<HotBin> h(count);
  std::vector<ColdBin> cb(count);
  std::vector<uint64_t> wl(count, 0);
  const size_t off = (size_t)w * s.cap + first;
  HIPCHECK(c, hipMemcpy(h.data(), s.d_hot + off, sizeof(HotBin) * count, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(cb.data(), s.d_cold + off, sizeof(ColdBin) * count, hipMemcpyDeviceToHost));
  if (s.d_wlh) HIPCHECK(c, hipMemcpy(wl.data(), s.d_wlh + off, sizeof(uint64_t) * count, hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < count; k++) {
    out[k].lnw = h[k].lnw;
    out[k].histogram = h[k].hist;
    out[k].energy_total = cb[k].etot;
    out[k].energy_squared_total = cb[k].esq;
    out[k].round_trips = cb[k].rt;
    out[k].visit_epoch = cb[k].vep;
    out[k].t_found = cb[k].tfound;
    out[k].wl_hist = wl[k];
  }
  return SADMC_OK;
}

extern "C" {

int sadmc_abi_version(void) { return SADMC_ABI_VERSION; }

int sadmc_init_coords(int32_t system, int32_t n, double geom, uint64_t seed, double *coords) {
  if (!coords || n < 2) return SADMC_E_ARG;
  if (system == SADMC_LJ) return lj_init_coords(n, geom, seed, coords);
  if (system == SADMC_WCA) return wca_init_coords(n, geom, seed, coords);
  return SADMC_E_ARG;
}

int sadmc_init_spins(int32_t n, int32_t rule, uint64_t seed, int8_t *spins) {
  if (!spins || n < 2) return SADMC_E_ARG;
  uint64_t s0, s1;
  rng_seed(rule == 0 ? 10137ULL : (seed ^ kInitTag), s0, s1);
  for (int64_t k = 0; k < (int64_t)n * n; k++) spins[k] = (int8_t)(((int8_t)rng_next(s0, s1) & 1) * 2 - 1);
  return SADMC_OK;
}

int sadmc_bin_window(const sadmc_set_params *p, double geom, double *e_lo, double *e_hi, int64_t *cap) {
  if (!p) return SADMC_E_ARG;
  const double w = p->energy_bin > 0 ? p->energy_bin : (p->system == SADMC_ISING ? 4.0 : 1.0);
  double lo, hi;
  if (p->e_lo_bound < p->e_hi_bound) {
    lo = p->e_lo_bound;
    hi = p->e_hi_bound;
  } else if (p->system == SADMC_ISING) {
    lo = -2.0 * p->n * p->n;
    hi = 2.0 * p->n * p->n;
  } else {
    lo = p->system == SADMC_LJ ? lj_lowest(p->n) - 1.0 : 0.0;
    if (p->has_min_e && p->min_e - w > lo) lo = p->min_e - w;
    hi = p->has_max_e ? p->max_e : lo + 16384.0 * w;
  }
  (void)geom;
  if (e_lo) *e_lo = lo;
  if (e_hi) *e_hi = hi;
  if (cap) *cap = (int64_t)std::ceil((hi - lo) / w) + 8;
  return SADMC_OK;
}

int sadmc_create(sadmc_ctx **out, int device) {
  if (!out) return SADMC_E_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return SADMC_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SADMC_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SADMC_E_NODEVICE;
  sadmc_ctx *c = new sadmc_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return SADMC_E_HIP;
  }
  *out = c;
  return SADMC_OK;
}

static void free_set(SetHost &s) {
  (void)hipFree(s.d_ws);
  (void)hipFree(s.d_pos);
  (void)hipFree(s.d_spins);
  (void)hipFree(s.d_hot);
  (void)hipFree(s.d_cold);
  (void)hipFree(s.d_wlh);
  (void)hipFree(s.d_extra);
  (void)hipFree(s.d_pending);
  (void)hipFree(s.d_flag);
  (void)hipFree(s.d_eout);
  s.d_eout = nullptr;
  s.d_ws = nullptr;
  s.d_pos = nullptr;
  s.d_spins = nullptr;
  s.d_hot = nullptr;
  s.d_cold = nullptr;
  s.d_wlh = nullptr;
  s.d_extra = nullptr;
  s.d_pending = nullptr;
  s.d_flag = nullptr;
}

void sadmc_destroy(sadmc_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &s : c->sets) free_set(s);
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  for (auto &g : c->graph_defs) (void)hipGraphDestroy(g);
  (void)hipFree(c->d_sets);
  (void)hipFree(c->d_tiles);
  for (void *p : {(void *)c->d_pw, (void *)c->d_pm, (void *)c->d_pv, (void *)c->d_pbf, (void *)c->d_rec,
                  (void *)c->d_partial, (void *)c->d_grad, (void *)c->d_loss})
    (void)hipFree(p);
  (void)hipFree(c->d_ent);
  for (auto &z : c->zmcs)
    for (void *p : {(void *)z.d_sim, (void *)z.d_rep, (void *)z.d_slot, (void *)z.d_med, (void *)z.d_pos,
                    (void *)z.d_geom, (void *)z.d_slot_of, (void *)z.d_spins})
      (void)hipFree(p);
  for (auto &b : c->bsets)
    for (void *p : {(void *)b.d_ws, (void *)b.d_bins, (void *)b.d_pos, (void *)b.d_spins, (void *)b.d_flag,
                    (void *)b.d_hrs, (void *)b.d_hr})
      (void)hipFree(p);
  for (auto &t : c->pts)
    for (void *p : {(void *)t.d_rep, (void *)t.d_slot_e, (void *)t.d_pos, (void *)t.d_geom, (void *)t.d_slot_of,
                    (void *)t.d_spins, (void *)t.d_lrng})
      (void)hipFree(p);
  for (int k = 0; k < kSideStreams; k++) {
    if (c->side[k]) (void)hipStreamSynchronize(c->side[k]);
    if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *sadmc_last_error(const sadmc_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int auto_lanes(int n) {
  int want = (n + 3) / 4;
  int g = 2;
  while (g < want && g < 64) g <<= 1;
  return g;
}

int sadmc_add_set(sadmc_ctx *c, const sadmc_set_params *p, int32_t W, const uint64_t *seeds, const double *geom,
                  const double *coords, const int8_t *spins, int32_t *set_id) {
  if (!c || !p || W <= 0 || !seeds) return fail(c, SADMC_E_ARG, "null argument or no walkers");
  if (c->inited) return fail(c, SADMC_E_STATE, "sets must be added before sadmc_init");
  SetHost s;
  s.p = *p;
  s.W = W;
  const int n = p->n;
  if (p->system == SADMC_ISING) {
    if (n < 2 || n > 8) return fail(c, SADMC_E_ARG, "Ising lattice width must be 2..8");
    s.G = 1;
    s.K = 1;
    s.KMAX = 1;
    s.npad = 0;
  } else if (p->system == SADMC_LJ || p->system == SADMC_WCA) {
    if (n < 2 || n > 1024) return fail(c, SADMC_E_ARG, "atom count must be 2..1024");
    if (!geom) return fail(c, SADMC_E_ARG, "LJ/WCA need per-walker geometry");
    int G = p->lanes > 0 ? p->lanes : auto_lanes(n);
    if (G < 2 || G > 64 || (G & (G - 1))) return fail(c, SADMC_E_ARG, "lanes must be a power of two in 2..64");
    s.G = G;
    s.K = (n + G - 1) / G;
    if (s.K > 16) return fail(c, SADMC_E_ARG, "too few lanes for this many atoms (max 16 atoms per lane)");
    s.KMAX = s.K <= 1 ? 1 : s.K <= 2 ? 2 : s.K <= 4 ? 4 : s.K <= 8 ? 8 : 16;
    s.npad = G * s.K;
    if (p->precision == SADMC_FP32_PAIRS && p->system != SADMC_LJ)
      return fail(c, SADMC_E_ARG, "FP32 pair terms are implemented for LJ only");
  } else {
    return fail(c, SADMC_E_ARG, "unknown system");
  }
  if (p->method < SADMC_SAD || p->method > SADMC_CANONICAL) return fail(c, SADMC_E_ARG, "unknown method");
  if (p->method == SADMC_CANONICAL && !(p->method_param > 0)) return fail(c, SADMC_E_ARG, "canonical T must be > 0");
  if (p->move_param <= 0 && p->move_plan == SADMC_TRANSLATION_SCALE && p->system != SADMC_ISING)
    return fail(c, SADMC_E_ARG, "translation scale must be > 0");
  if (p->move_kind == SADMC_MOVE_LEARNED) {
    if (p->system != SADMC_LJ || (p->method != SADMC_SAD && p->method != SADMC_CANONICAL))
      return fail(c, SADMC_E_ARG, "the learned move is implemented for LJ walkers with SAD or canonical sampling");
  } else if (p->move_kind != SADMC_MOVE_CUBE) {
    return fail(c, SADMC_E_ARG, "unknown move kind");
  }
  s.width = p->energy_bin > 0 ? p->energy_bin : (p->system == SADMC_ISING ? 4.0 : 1.0);
  sadmc_bin_window(p, geom ? geom[0] : 0.0, &s.e_lo, &s.e_hi, &s.cap);
  s.seeds.assign(seeds, seeds + W);
  if (p->system == SADMC_ISING) {
    s.spins.resize((size_t)W * n * n);
    for (int w = 0; w < W; w++) {
      if (spins)
        std::memcpy(&s.spins[(size_t)w * n * n], spins + (size_t)w * n * n, (size_t)n * n);
      else
        sadmc_init_spins(n, p->ising_init, seeds[w], &s.spins[(size_t)w * n * n]);
    }
  } else {
    s.geom.assign(geom, geom + W);
    s.coords.resize((size_t)W * 3 * n);
    for (int w = 0; w < W; w++) {
      double *dst = &s.coords[(size_t)w * 3 * n];
      if (coords) {
        std::memcpy(dst, coords + (size_t)w * 3 * n, sizeof(double) * 3 * n);
      } else {
        int rc = sadmc_init_coords(p->system, n, geom[w], seeds[w], dst);
        if (rc) return fail(c, SADMC_E_INIT, "initial configuration failed");
      }
    }
  }
  c->sets.push_back(std::move(s));
  if (set_id) *set_id = (int32_t)c->sets.size() - 1;
  return SADMC_OK;
}

static void fill_args(SetHost &s) {
  SetArgs &a = s.args;
  a.ws = s.d_ws;
  a.pos = s.d_pos;
  a.spins = s.d_spins;
  a.hot = s.d_hot;
  a.cold = s.d_cold;
  a.wlh = s.d_wlh;
  a.extra = s.d_extra;
  a.cap = s.cap;
  a.W = s.W;
  a.n = s.p.n;
  a.npad = s.npad;
  a.G = s.G;
  a.K = s.K;
  a.method = s.p.method;
  a.move_plan = s.p.move_plan;
  a.has_min_e = s.p.has_min_e;
  a.has_max_e = s.p.has_max_e;
  a.has_min_gamma = s.p.has_min_gamma;
  a.refresh = s.p.refresh;
  a.collect_pressure = s.p.collect_pressure;
  a.precision = s.p.precision;
  a.move_kind = s.p.move_kind;
  a.method_param = s.p.method_param;
  a.move_param = s.p.move_param;
  a.min_e = s.p.min_e;
  a.max_e = s.p.max_e;
  a.rc2 = s.p.rc2;
  a.width = s.width;
  a.inv_width = 1.0 / s.width;
  a.flag = s.d_flag;
  a.mlp = nullptr;
  const uint64_t nn = (uint64_t)s.p.n;
  a.zone_n = ~0ULL - ((~0ULL - nn + 1) % nn);
  a.zone_site = (nn << __builtin_clzll(nn)) - 1;
}

static int check_flags(sadmc_ctx *c) {
  for (size_t k = 0; k < c->sets.size(); k++) {
    SetHost &s = c->sets[k];
    int32_t flag = 0;
    HIPCHECK(c, hipMemcpy(&flag, s.d_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (!flag) continue;
    std::vector<sadmc_walker_state> ws(s.W);
    HIPCHECK(c, hipMemcpy(ws.data(), s.d_ws, sizeof(sadmc_walker_state) * s.W, hipMemcpyDeviceToHost));
    for (int w = 0; w < s.W; w++) {
      int st = ws[w].status;
      if (!st) continue;
      char buf[160];
      std::snprintf(buf, sizeof buf, "set %zu walker %d stopped with status %d at move %llu", k, w, st,
                    (unsigned long long)ws[w].moves);
      int code = st == SADMC_WS_CAPACITY ? SADMC_E_CAPACITY
                 : st == SADMC_WS_VERIFY ? SADMC_E_VERIFY
                 : st == SADMC_WS_INIT   ? SADMC_E_INIT
                                         : SADMC_E_INDEX;
      return fail(c, code, buf);
    }
  }
  return SADMC_OK;
}

// Default network: W1 ~ N(0, 0.25^2), W2 ~ N(0, 0.25^2), w3 ~ N(0, 0.05^2),
// hidden biases 0.1, output bias ln 0.1 (sigma near 0.1 everywhere).
static void default_weights(uint64_t seed, float *w) {
  uint64_t s0, s1;
  rng_seed(seed ^ 0x5EEDC0DE5EEDC0DEULL, s0, s1);
  auto normal = [&]() {
    double u1 = gen_f64(s0, s1), u2 = gen_f64(s0, s1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int i = 0; i < kPropNP; i++) w[i] = 0.0f;
  for (int i = kW1; i < kB1; i++) w[i] = (float)(0.25 * normal());
  for (int i = kB1; i < kW2; i++) w[i] = 0.1f;
  for (int i = kW2; i < kB2; i++) w[i] = (float)(0.25 * normal());
  for (int i = kB2; i < kW3; i++) w[i] = 0.1f;
  for (int i = kW3; i < kB3; i++) w[i] = (float)(0.05 * normal());
  w[kB3] = (float)std::log(0.1);
}

static int upload_weights(sadmc_ctx *c, const float *w) {
  std::vector<float> m(kPropLds, 0.0f), z(kPropLds, 0.0f);
  std::vector<uint16_t> b(kPropLds, 0);
  for (int i = 0; i < kPropNP; i++) {
    if (!std::isfinite(w[i])) return fail(c, SADMC_E_ARG, "non-finite weight");
    m[i] = w[i];
    b[i] = f32_to_bf16(w[i]);
  }
  HIPCHECK(c, hipMemcpyAsync(c->d_pw, m.data(), sizeof(float) * kPropLds, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_pm, z.data(), sizeof(float) * kPropLds, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_pv, z.data(), sizeof(float) * kPropLds, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_pbf, b.data(), sizeof(uint16_t) * kPropLds, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->prop_step = 0;
  return SADMC_OK;
}

static int ensure_prop(sadmc_ctx *c) {
  if (c->d_pw) return SADMC_OK;
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipMalloc(&c->d_pw, sizeof(float) * kPropLds));
  HIPCHECK(c, hipMalloc(&c->d_pm, sizeof(float) * kPropLds));
  HIPCHECK(c, hipMalloc(&c->d_pv, sizeof(float) * kPropLds));
  HIPCHECK(c, hipMalloc(&c->d_pbf, sizeof(uint16_t) * kPropLds));
  HIPCHECK(c, hipMalloc(&c->d_grad, sizeof(float) * kPropLds));
  std::vector<float> w(kPropNP);
  default_weights(0, w.data());
  return upload_weights(c, w.data());
}

int sadmc_init(sadmc_ctx *c) {
  if (!c) return SADMC_E_ARG;
  if (c->inited) return fail(c, SADMC_E_STATE, "already initialised");
  if (c->sets.empty()) return fail(c, SADMC_E_STATE, "no walker sets");
  HIPCHECK(c, hipSetDevice(c->device));
  for (auto &s : c->sets) {
    const int W = s.W, n = s.p.n;
    HIPCHECK(c, hipMalloc(&s.d_ws, sizeof(sadmc_walker_state) * W));
    const size_t nb = (size_t)W * s.cap;
    HIPCHECK(c, hipMalloc(&s.d_hot, sizeof(HotBin) * nb));
    HIPCHECK(c, hipMemsetAsync(s.d_hot, 0, sizeof(HotBin) * nb, c->stream));
    HIPCHECK(c, hipMalloc(&s.d_cold, sizeof(ColdBin) * nb));
    HIPCHECK(c, hipMemsetAsync(s.d_cold, 0, sizeof(ColdBin) * nb, c->stream));
    if (s.p.method == SADMC_WL || s.p.method == SADMC_INV_T_WL) {
      HIPCHECK(c, hipMalloc(&s.d_wlh, sizeof(uint64_t) * nb));
      HIPCHECK(c, hipMemsetAsync(s.d_wlh, 0, sizeof(uint64_t) * nb, c->stream));
    }
    HIPCHECK(c, hipMalloc(&s.d_pending, sizeof(Pending)));
    HIPCHECK(c, hipMalloc(&s.d_eout, sizeof(double)));
    HIPCHECK(c, hipMalloc(&s.d_flag, sizeof(int32_t)));
    HIPCHECK(c, hipMemsetAsync(s.d_flag, 0, sizeof(int32_t), c->stream));
    std::vector<sadmc_walker_state> ws(W);
    std::memset(ws.data(), 0, sizeof(sadmc_walker_state) * W);
    for (int w = 0; w < W; w++) {
      rng_seed(s.seeds[w], ws[w].rng[0], ws[w].rng[1]);
      ws[w].geom = s.p.system == SADMC_ISING ? 0.0 : s.geom[w];
      ws[w].cap = s.cap;
    }
    HIPCHECK(c, hipMemcpyAsync(s.d_ws, ws.data(), sizeof(sadmc_walker_state) * W, hipMemcpyHostToDevice, c->stream));
    if (s.p.system == SADMC_ISING) {
      std::vector<uint64_t> bits(W, 0);
      for (int w = 0; w < W; w++)
        for (int b = 0; b < n * n; b++)
          if (s.spins[(size_t)w * n * n + b] > 0) bits[w] |= 1ULL << b;
      HIPCHECK(c, hipMalloc(&s.d_spins, sizeof(uint64_t) * W));
      HIPCHECK(c, hipMemcpyAsync(s.d_spins, bits.data(), sizeof(uint64_t) * W, hipMemcpyHostToDevice, c->stream));
    } else {
      std::vector<double> pos((size_t)W * 3 * s.npad, 0.0);
      for (int w = 0; w < W; w++)
        for (int d = 0; d < 3; d++)
          for (int j = 0; j < n; j++) pos[((size_t)w * 3 + d) * s.npad + j] = s.coords[((size_t)w * 3 + d) * n + j];
      HIPCHECK(c, hipMalloc(&s.d_pos, sizeof(double) * pos.size()));
      HIPCHECK(c, hipMemcpyAsync(s.d_pos, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice, c->stream));
      if (s.p.system == SADMC_WCA && s.p.collect_pressure) {
        HIPCHECK(c, hipMalloc(&s.d_extra, sizeof(PExtra) * (size_t)W * s.cap));
        HIPCHECK(c, hipMemsetAsync(s.d_extra, 0, sizeof(PExtra) * (size_t)W * s.cap, c->stream));
      }
    }
    fill_args(s);
    if (s.p.move_kind == SADMC_MOVE_LEARNED) {
      int rc = ensure_prop(c);
      if (rc) return rc;
      s.args.mlp = c->d_pbf;
    }
    LaunchArgs la{};
    la.e_lo = s.e_lo;
    HIPCHECK(c, launch(s.p.system, s.p.precision, s.p.method, s.KMAX, KK_INIT, s.args, la, c->stream));
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->inited = true;
  return check_flags(c);
}

// The side streams and their fork/join events, created on first use (a
// call after a failed creation creates only what is missing).
static int ensure_side_streams(sadmc_ctx *c) {
  if (c->ev_fork) return SADMC_OK;
  for (int k = 0; k < kSideStreams; k++) {
    if (!c->side[k]) HIPCHECK(c, hipStreamCreate(&c->side[k]));
    if (!c->ev_join[k]) HIPCHECK(c, hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
  }
  HIPCHECK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  return SADMC_OK;
}

int sadmc_run(sadmc_ctx *c, uint64_t moves) {
  if (!c) return SADMC_E_ARG;
  if (!c->inited) return fail(c, SADMC_E_STATE, "sadmc_init first");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  const int nsets = (int)c->sets.size();
  const int ns = nsets < kSideStreams ? nsets : kSideStreams;
  if (ns > 1) {
    // sets are independent: fork after everything already on the stream,
    // set i on side stream i mod ns, and join before the closing event
    int rc = ensure_side_streams(c);
    if (rc) return rc;
    HIPCHECK(c, hipEventRecord(c->ev_fork, c->stream));
    for (int k = 0; k < ns; k++) HIPCHECK(c, hipStreamWaitEvent(c->side[k], c->ev_fork, 0));
  }
  hipError_t le = hipSuccess;
  for (int i = 0; i < nsets && le == hipSuccess; i++) {
    SetHost &s = c->sets[i];
    LaunchArgs la{};
    la.moves = moves;
    la.waves = s.waves;
    hipStream_t st = ns > 1 ? c->side[i % ns] : c->stream;
    le = launch(s.p.system, s.p.precision, s.p.method, s.KMAX, KK_RUN, s.args, la, st);
  }
  if (ns > 1) {
    // joined even after a failed launch, so that the context's stream still
    // orders everything already queued on the side streams
    for (int k = 0; k < ns; k++) {
      HIPCHECK(c, hipEventRecord(c->ev_join[k], c->side[k]));
      HIPCHECK(c, hipStreamWaitEvent(c->stream, c->ev_join[k], 0));
    }
  }
  if (le != hipSuccess) return fail(c, SADMC_E_HIP, std::string("k_run launch: ") + hipGetErrorString(le));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return SADMC_OK;
}

int sadmc_run_set(sadmc_ctx *c, int32_t set, uint64_t moves) {
  if (!c) return SADMC_E_ARG;
  if (!c->inited) return fail(c, SADMC_E_STATE, "sadmc_init first");
  if (set < 0 || set >= (int32_t)c->sets.size()) return fail(c, SADMC_E_ARG, "bad set");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  SetHost &s = c->sets[set];
  LaunchArgs la{};
  la.moves = moves;
  la.waves = s.waves;
  HIPCHECK(c, launch(s.p.system, s.p.precision, s.p.method, s.KMAX, KK_RUN, s.args, la, c->stream));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return SADMC_OK;
}

int sadmc_check_recip_f32(int device, int32_t what, uint64_t lo, uint64_t hi, uint64_t *mismatches) {
  if (!mismatches || (what != 0 && what != 1) || lo > hi || hi > 0x100000000ULL) return SADMC_E_ARG;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return SADMC_E_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return SADMC_E_HIP;
  unsigned long long *d_bad = nullptr, bad = 0;
  if (hipMalloc(&d_bad, sizeof bad) != hipSuccess) return SADMC_E_HIP;
  hipError_t e = hipMemset(d_bad, 0, sizeof bad);
  if (e == hipSuccess) e = launch_check_recip(what, lo, hi, d_bad, nullptr);
  if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost);
  (void)hipFree(d_bad);
  if (e != hipSuccess) return SADMC_E_HIP;
  *mismatches = bad;
  return SADMC_OK;
}

int sadmc_sync(sadmc_ctx *c) {
  if (!c) return SADMC_E_ARG;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int sadmc_last_run_ms(const sadmc_ctx *c, double *ms) {
  if (!c || !ms || !c->timed) return SADMC_E_STATE;
  float f = 0.0f;
  if (hipEventElapsedTime(&f, c->ev0, c->ev1) != hipSuccess) return SADMC_E_HIP;
  *ms = f;
  return SADMC_OK;
}

// Tiles of the batched launch, heaviest first: a tile's cost is its walkers
// times atoms (the pair work of one move).
static int build_batch(sadmc_ctx *c) {
  if (c->d_tiles) return SADMC_OK;
  struct T {
    double cost;
    BatchTile t;
  };
  std::vector<T> tiles;
  size_t lds = 0;
  for (size_t k = 0; k < c->sets.size(); k++) {
    const SetHost &s = c->sets[k];
    if (s.p.method != SADMC_SAD || s.p.precision != SADMC_FP64 || s.p.system == SADMC_ISING ||
        s.p.move_kind != SADMC_MOVE_CUBE)
      return fail(c, SADMC_E_ARG, "the batched launch runs SAD walkers of LJ/WCA sets with FP64 pair terms");
    const int per_block = 64 / s.G;
    lds = std::max(lds, (size_t)per_block * 3 * s.npad * sizeof(double));
    for (int w = 0; w < s.W; w += per_block) {
      T t;
      t.cost = (double)std::min(per_block, s.W - w) * s.p.n * (s.p.system == SADMC_WCA ? 1.6 : 1.0);
      t.t.set = (int32_t)k;
      t.t.first_walker = w;
      t.t.sys = s.p.system;
      t.t.pad = 0;
      tiles.push_back(t);
    }
  }
  std::stable_sort(tiles.begin(), tiles.end(), [](const T &a, const T &b) { return a.cost > b.cost; });
  std::vector<BatchTile> bt(tiles.size());
  for (size_t k = 0; k < tiles.size(); k++) bt[k] = tiles[k].t;
  std::vector<SetArgs> sa(c->sets.size());
  for (size_t k = 0; k < c->sets.size(); k++) sa[k] = c->sets[k].args;
  // both tables or neither: a failure part way leaves nothing allocated
  hipError_t e = hipMalloc(&c->d_sets, sizeof(SetArgs) * sa.size());
  if (e == hipSuccess) e = hipMemcpy(c->d_sets, sa.data(), sizeof(SetArgs) * sa.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_tiles, sizeof(BatchTile) * bt.size());
  if (e == hipSuccess) e = hipMemcpy(c->d_tiles, bt.data(), sizeof(BatchTile) * bt.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(c->d_sets);
    (void)hipFree(c->d_tiles);
    c->d_sets = nullptr;
    c->d_tiles = nullptr;
    return fail(c, SADMC_E_HIP, std::string("batch tables: ") + hipGetErrorString(e));
  }
  c->ntiles = (int)bt.size();
  c->batch_lds = lds;
  return SADMC_OK;
}

int sadmc_run_batched(sadmc_ctx *c, uint64_t moves) {
  if (!c) return SADMC_E_ARG;
  if (!c->inited) return fail(c, SADMC_E_STATE, "sadmc_init first");
  HIPCHECK(c, hipSetDevice(c->device));
  int rc = build_batch(c);
  if (rc) return rc;
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHECK(c, launch_batched(c->d_sets, c->d_tiles, c->ntiles, c->batch_lds, c->batch_waves, moves, c->stream));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return SADMC_OK;
}

int sadmc_graph_run(sadmc_ctx *c, uint64_t moves) {
  if (!c) return SADMC_E_ARG;
  if (!c->inited) return fail(c, SADMC_E_STATE, "sadmc_init first");
  HIPCHECK(c, hipSetDevice(c->device));
  int rc = build_batch(c);
  if (rc) return rc;
  auto it = c->graphs.find(moves);
  if (it == c->graphs.end()) {
    hipGraph_t g;
    hipGraphExec_t ge;
    HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = launch_batched(c->d_sets, c->d_tiles, c->ntiles, c->batch_lds, c->batch_waves, moves,
                                   c->stream);
    HIPCHECK(c, hipStreamEndCapture(c->stream, &g));
    HIPCHECK(c, le);
    HIPCHECK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    c->graph_defs.push_back(g);
    it = c->graphs.emplace(moves, ge).first;
  }
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHECK(c, hipGraphLaunch(it->second, c->stream));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return SADMC_OK;
}

int sadmc_autotune_lanes(int device, const sadmc_set_params *p, int32_t n_walkers, const uint64_t *seeds,
                         const double *geom, uint64_t warm_moves, uint64_t trial_moves, int32_t *best_lanes,
                         double *rates) {
  if (!p || !seeds || !best_lanes || n_walkers <= 0) return SADMC_E_ARG;
  if (rates)
    for (int k = 0; k < 7; k++) rates[k] = 0.0;
  if (p->system == SADMC_ISING) {
    *best_lanes = 1;
    return SADMC_OK;
  }
  double best = -1.0;
  int bestg = 0, last = SADMC_OK;
  for (int k = 1; k < 7; k++) {
    const int G = 1 << k;
    if ((p->n + G - 1) / G > 16) continue;
    sadmc_set_params q = *p;
    q.lanes = G;
    sadmc_ctx *c = nullptr;
    int rc = sadmc_create(&c, device);
    if (rc) return rc;
    rc = sadmc_add_set(c, &q, n_walkers, seeds, geom, nullptr, nullptr, nullptr);
    if (!rc) rc = sadmc_init(c);
    if (!rc && warm_moves) rc = sadmc_run(c, warm_moves);
    if (!rc) rc = sadmc_sync(c);
    if (!rc) rc = sadmc_run(c, trial_moves);
    if (!rc) rc = sadmc_sync(c);
    double ms = 0.0;
    if (!rc) rc = sadmc_last_run_ms(c, &ms);
    sadmc_destroy(c);
    if (rc) {
      last = rc;
      continue;
    }
    const double rate = (double)n_walkers * (double)trial_moves / (ms * 1e-3);
    if (rates) rates[k] = rate;
    if (rate > best) {
      best = rate;
      bestg = G;
    }
  }
  if (!bestg) return last ? last : SADMC_E_ARG;
  *best_lanes = bestg;
  return SADMC_OK;
}

int sadmc_set_block_waves(sadmc_ctx *c, int32_t set, int32_t waves) {
  if (!c) return SADMC_E_ARG;
  if (waves != 1 && waves != 2 && waves != 4) return fail(c, SADMC_E_ARG, "waves per workgroup must be 1, 2 or 4");
  const size_t lds_max = 65536;
  if (set == -1) {
    size_t wave_lds = 0;
    for (const SetHost &s : c->sets)
      if (s.p.system != SADMC_ISING)
        wave_lds = std::max(wave_lds, (size_t)(64 / s.G) * 3 * s.npad * sizeof(double));
    if (wave_lds * waves > lds_max) return fail(c, SADMC_E_ARG, "the batched workgroup would exceed 64 KiB of LDS");
    if (waves != c->batch_waves) {
      // captured graphs hold the old launch shape
      for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
      c->graphs.clear();
      c->batch_waves = waves;
    }
    return SADMC_OK;
  }
  if (set < 0 || set >= (int32_t)c->sets.size()) return fail(c, SADMC_E_ARG, "bad set");
  SetHost &s = c->sets[set];
  const size_t wave_lds = s.p.system == SADMC_ISING ? 0 : (size_t)(64 / s.G) * 3 * s.npad * sizeof(double);
  if (wave_lds * waves > lds_max) return fail(c, SADMC_E_ARG, "the workgroup would exceed 64 KiB of LDS");
  s.waves = waves;
  return SADMC_OK;
}

// One timed trial of a mixed batch: every set with its lanes, the batched
// launch with `waves` wavefronts per workgroup, graph-captured; moves/s.
static int time_batch(int device, int32_t nsets, const sadmc_set_params *ps, const int32_t *nw,
                      const uint64_t *seeds, const double *geom, const int32_t *lanes, int waves, uint64_t warm,
                      uint64_t trial, double *rate) {
  sadmc_ctx *c = nullptr;
  int rc = sadmc_create(&c, device);
  if (rc) return rc;
  int64_t off = 0, total = 0;
  for (int32_t k = 0; k < nsets && !rc; k++) {
    sadmc_set_params q = ps[k];
    q.lanes = lanes[k];
    rc = sadmc_add_set(c, &q, nw[k], seeds + off, geom + off, nullptr, nullptr, nullptr);
    off += nw[k];
    total += nw[k];
  }
  if (!rc) rc = sadmc_init(c);
  if (!rc) rc = sadmc_set_block_waves(c, -1, waves);
  if (!rc && warm) rc = sadmc_graph_run(c, warm);
  if (!rc) rc = sadmc_sync(c);
  if (!rc) rc = sadmc_graph_run(c, trial);  // captures the trial's graph
  if (!rc) rc = sadmc_sync(c);
  if (!rc) rc = sadmc_graph_run(c, trial);
  if (!rc) rc = sadmc_sync(c);
  double ms = 0.0;
  if (!rc) rc = sadmc_last_run_ms(c, &ms);
  sadmc_destroy(c);
  if (rc) return rc;
  *rate = (double)total * (double)trial / (ms * 1e-3);
  return SADMC_OK;
}

int sadmc_autotune_batch(int device, int32_t nsets, const sadmc_set_params *ps, const int32_t *n_walkers,
                         const uint64_t *seeds, const double *geom, uint64_t warm_moves, uint64_t trial_moves,
                         int32_t *lanes_out, int32_t *waves_out, double *rate_out) {
  if (nsets <= 0 || !ps || !n_walkers || !seeds || !geom || !lanes_out || !waves_out || !trial_moves)
    return SADMC_E_ARG;
  std::vector<int32_t> lanes(nsets);
  std::vector<std::vector<int32_t>> cand(nsets);
  for (int32_t k = 0; k < nsets; k++) {
    if (ps[k].system == SADMC_ISING || n_walkers[k] <= 0) return SADMC_E_ARG;
    for (int G = 2; G <= 64; G <<= 1)
      if ((ps[k].n + G - 1) / G <= 16) cand[k].push_back(G);
    lanes[k] = ps[k].lanes > 0 ? ps[k].lanes : auto_lanes(ps[k].n);
  }
  int waves = 1;
  double best = 0.0;
  int rc = time_batch(device, nsets, ps, n_walkers, seeds, geom, lanes.data(), waves, warm_moves, trial_moves, &best);
  if (rc) return rc;
  // coordinate descent over each set's lanes, the whole batch timed each time
  for (int32_t k = 0; k < nsets; k++) {
    const int32_t keep = lanes[k];
    int32_t pick = keep;
    for (int32_t G : cand[k]) {
      if (G == keep) continue;
      lanes[k] = G;
      double r = 0.0;
      if (time_batch(device, nsets, ps, n_walkers, seeds, geom, lanes.data(), waves, warm_moves, trial_moves, &r) ==
              SADMC_OK &&
          r > best) {
        best = r;
        pick = G;
      }
    }
    lanes[k] = pick;
  }
  // then the wavefronts per workgroup
  for (int wv : {2, 4}) {
    double r = 0.0;
    if (time_batch(device, nsets, ps, n_walkers, seeds, geom, lanes.data(), wv, warm_moves, trial_moves, &r) ==
            SADMC_OK &&
        r > best) {
      best = r;
      waves = wv;
    }
  }
  for (int32_t k = 0; k < nsets; k++) lanes_out[k] = lanes[k];
  *waves_out = waves;
  if (rate_out) *rate_out = best;
  return SADMC_OK;
}

static SetHost *get_set(sadmc_ctx *c, int32_t set, int32_t walker) {
  if (!c || set < 0 || set >= (int32_t)c->sets.size()) return nullptr;
  SetHost *s = &c->sets[set];
  if (walker < 0 || walker >= s->W) return nullptr;
  return s;
}

int sadmc_plan_move(sadmc_ctx *c, int32_t set, int32_t walker, double scale, double *e_new, int32_t *some) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  LaunchArgs la{};
  la.walker = walker;
  la.scale = scale;
  la.pending = s->d_pending;
  HIPCHECK(c, launch(s->p.system, s->p.precision, s->p.method, s->KMAX, KK_PLAN, s->args, la, c->stream));
  Pending p;
  HIPCHECK(c, hipMemcpyAsync(&p, s->d_pending, sizeof p, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (e_new) *e_new = p.e;
  if (some) *some = p.some;
  return SADMC_OK;
}

int sadmc_confirm(sadmc_ctx *c, int32_t set, int32_t walker) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  LaunchArgs la{};
  la.walker = walker;
  la.pending = s->d_pending;
  HIPCHECK(c, launch(s->p.system, s->p.precision, s->p.method, s->KMAX, KK_CONFIRM, s->args, la, c->stream));
  Pending none{};
  HIPCHECK(c, hipMemcpyAsync(s->d_pending, &none, sizeof none, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_plan_learned(sadmc_ctx *c, int32_t set, int32_t walker, double *e_new, double *lnh, double *to,
                       int32_t *some) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  if (s->p.move_kind != SADMC_MOVE_LEARNED) return fail(c, SADMC_E_ARG, "the set does not use the learned move");
  LaunchArgs la{};
  la.walker = walker;
  la.pending = s->d_pending;
  HIPCHECK(c, launch(s->p.system, s->p.precision, s->p.method, s->KMAX, KK_PLAN, s->args, la, c->stream));
  Pending p;
  HIPCHECK(c, hipMemcpyAsync(&p, s->d_pending, sizeof p, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (e_new) *e_new = p.e;
  if (lnh) *lnh = p.dabse;
  if (to) {
    to[0] = p.tx;
    to[1] = p.ty;
    to[2] = p.tz;
  }
  if (some) *some = p.some;
  return SADMC_OK;
}

// ---------------------------------------------------------- learned proposal
int sadmc_prop_init(sadmc_ctx *c, uint64_t seed) {
  if (!c) return SADMC_E_ARG;
  int rc = ensure_prop(c);
  if (rc) return rc;
  std::vector<float> w(kPropNP);
  default_weights(seed, w.data());
  return upload_weights(c, w.data());
}

int sadmc_prop_set_weights(sadmc_ctx *c, const float *w, int32_t count) {
  if (!c || !w || count != kPropNP) return fail(c, SADMC_E_ARG, "expected 1377 weights");
  int rc = ensure_prop(c);
  if (rc) return rc;
  return upload_weights(c, w);
}

int sadmc_prop_get_weights(sadmc_ctx *c, float *master, uint16_t *bf16) {
  if (!c) return SADMC_E_ARG;
  int rc = ensure_prop(c);
  if (rc) return rc;
  if (master) HIPCHECK(c, hipMemcpyAsync(master, c->d_pw, sizeof(float) * kPropNP, hipMemcpyDeviceToHost, c->stream));
  if (bf16) HIPCHECK(c, hipMemcpyAsync(bf16, c->d_pbf, sizeof(uint16_t) * kPropNP, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

static int reserve_records(sadmc_ctx *c, int64_t need) {
  if (need <= c->rec_cap) return SADMC_OK;
  int64_t cap = std::max<int64_t>(need, c->rec_cap * 2);
  float *p = nullptr;
  HIPCHECK(c, hipMalloc(&p, sizeof(float) * kPropRec * cap));
  if (c->nrec)
    HIPCHECK(c, hipMemcpyAsync(p, c->d_rec, sizeof(float) * kPropRec * c->nrec, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_rec);
  c->d_rec = p;
  c->rec_cap = cap;
  return SADMC_OK;
}

int sadmc_prop_collect(sadmc_ctx *c, int32_t set, uint64_t seed, int32_t per_walker) {
  SetHost *s = get_set(c, set, 0);
  if (!s || !c->inited || per_walker <= 0) return fail(c, SADMC_E_ARG, "bad set or count, or not initialised");
  if (s->p.system != SADMC_LJ || s->p.method != SADMC_SAD)
    return fail(c, SADMC_E_ARG, "probe records come from LJ SAD walkers");
  int rc = reserve_records(c, c->nrec + (int64_t)s->W * per_walker);
  if (rc) return rc;
  LaunchArgs la{};
  la.seed = seed;
  la.count = per_walker;
  la.out = c->d_rec + c->nrec * kPropRec;
  HIPCHECK(c, launch(s->p.system, s->p.precision, s->p.method, s->KMAX, KK_PROBE, s->args, la, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->nrec += (int64_t)s->W * per_walker;
  return SADMC_OK;
}

int sadmc_prop_records(sadmc_ctx *c, float *out, int64_t max_records, int64_t *count) {
  if (!c) return SADMC_E_ARG;
  if (count) *count = c->nrec;
  const int64_t k = std::min(max_records, c->nrec);
  if (out && k > 0) {
    HIPCHECK(c, hipMemcpyAsync(out, c->d_rec, sizeof(float) * kPropRec * k, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  return SADMC_OK;
}

int sadmc_prop_load_records(sadmc_ctx *c, const float *rec, int64_t count) {
  if (!c || count < 0 || (count && !rec)) return fail(c, SADMC_E_ARG, "bad records");
  HIPCHECK(c, hipSetDevice(c->device));
  c->nrec = 0;
  c->prop_cursor = 0;
  int rc = reserve_records(c, count);
  if (rc) return rc;
  if (count) {
    HIPCHECK(c, hipMemcpyAsync(c->d_rec, rec, sizeof(float) * kPropRec * count, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  c->nrec = count;
  return SADMC_OK;
}

static int reserve_partial(sadmc_ctx *c, int blocks) {
  if (blocks <= c->partial_blocks) return SADMC_OK;
  (void)hipFree(c->d_partial);
  c->d_partial = nullptr;
  HIPCHECK(c, hipMalloc(&c->d_partial, sizeof(float) * (kPropNP + 1) * blocks));
  c->partial_blocks = blocks;
  return SADMC_OK;
}

int sadmc_prop_gradient(sadmc_ctx *c, int64_t first, int32_t count, float *grad, float *loss) {
  if (!c || count <= 0 || first < 0) return fail(c, SADMC_E_ARG, "bad range");
  int rc = ensure_prop(c);
  if (rc) return rc;
  if (c->nrec == 0) return fail(c, SADMC_E_STATE, "no probe records");
  const int blocks = (count + 63) / 64;
  rc = reserve_partial(c, blocks);
  if (rc) return rc;
  if (c->loss_cap < 1) {
    HIPCHECK(c, hipMalloc(&c->d_loss, sizeof(float) * 1024));
    c->loss_cap = 1024;
  }
  HIPCHECK(c, launch_prop_grad(c->d_rec, c->nrec, first, count, c->d_pbf, c->d_partial, c->stream));
  HIPCHECK(c, launch_prop_update(c->d_partial, blocks, count, c->d_pw, c->d_pm, c->d_pv, c->d_pbf, 0.0f, 1,
                                 c->d_grad, c->d_loss, c->stream));
  if (grad) HIPCHECK(c, hipMemcpyAsync(grad, c->d_grad, sizeof(float) * kPropNP, hipMemcpyDeviceToHost, c->stream));
  if (loss) HIPCHECK(c, hipMemcpyAsync(loss, c->d_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_prop_train(sadmc_ctx *c, int32_t steps, int32_t batch, float lr, float *losses) {
  if (!c || steps <= 0 || batch <= 0 || !(lr > 0)) return fail(c, SADMC_E_ARG, "bad training arguments");
  int rc = ensure_prop(c);
  if (rc) return rc;
  if (c->nrec == 0) return fail(c, SADMC_E_STATE, "no probe records");
  const int blocks = (batch + 63) / 64;
  rc = reserve_partial(c, blocks);
  if (rc) return rc;
  if (c->loss_cap < steps) {
    (void)hipFree(c->d_loss);
    c->d_loss = nullptr;
    HIPCHECK(c, hipMalloc(&c->d_loss, sizeof(float) * steps));
    c->loss_cap = steps;
  }
  for (int k = 0; k < steps; k++) {
    const int64_t first = c->prop_cursor % c->nrec;
    c->prop_cursor = (c->prop_cursor + batch) % c->nrec;
    c->prop_step += 1;
    HIPCHECK(c, launch_prop_grad(c->d_rec, c->nrec, first, batch, c->d_pbf, c->d_partial, c->stream));
    HIPCHECK(c, launch_prop_update(c->d_partial, blocks, batch, c->d_pw, c->d_pm, c->d_pv, c->d_pbf, lr,
                                   c->prop_step, nullptr, c->d_loss + k, c->stream));
  }
  if (losses)
    HIPCHECK(c, hipMemcpyAsync(losses, c->d_loss, sizeof(float) * steps, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

}  // extern "C"

int ctx_device(const sadmc_ctx *c) { return c->device; }
int ctx_fail(sadmc_ctx *c, int code, const char *msg) { return fail(c, code, msg); }

// The walker-set entropy reduction with its scratch either grown on demand
// in the context (fixed_base null) or taken from a fixed device buffer of
// `fixed_cap` bytes at `fixed_base` (no allocation; sizes beyond it fail
// with SADMC_E_CAPACITY).
// Scratch layout: the grid summary and its partial records, the per-walker
// constants, S and counts, the chunks' partial sums.  Every kernel writes
// each scratch entry it later reads, so nothing needs clearing first.
int mean_entropy_impl(sadmc_ctx *c, int32_t set, double e_a, double e_b, int64_t max_bins, double *e_first,
                      int64_t *nbins, double *S, int64_t *count, char *fixed_base, size_t fixed_cap) {
  SetHost *s = get_set(c, set, 0);
  if (!s || !c->inited) return fail(c, SADMC_E_ARG, "bad set or not initialised");
  if (s->p.method != SADMC_SAD) return fail(c, SADMC_E_ARG, "the entropy reduction reads SAD walkers");
  HIPCHECK(c, hipSetDevice(c->device));
  const double w = s->width;
  const size_t al = 256;
  auto up = [&](size_t b) { return (b + al - 1) / al * al; };
  const size_t ew_off = up(sizeof(EntropyRange) * (1 + kEntropyRangeBlocks));
  const size_t fixed = ew_off + up(entropy_walker_bytes() * s->W);
  // room for `bytes` of scratch; returns its base or null (error set)
  auto room = [&](size_t bytes) -> char * {
    if (fixed_base) {
      if (bytes > fixed_cap) {
        char buf[200];
        std::snprintf(buf, sizeof buf,
                      "the reduction needs %zu bytes of scratch; the allocation-free entry point holds %zu", bytes,
                      fixed_cap);
        fail(c, SADMC_E_CAPACITY, buf);
        return nullptr;
      }
      return fixed_base;
    }
    if (c->ent_bytes < bytes) {
      (void)hipFree(c->d_ent);
      c->d_ent = nullptr;
      c->ent_bytes = 0;
      hipError_t e = hipMalloc(&c->d_ent, bytes);
      if (e != hipSuccess) {
        fail(c, SADMC_E_HIP, std::string("entropy scratch: ") + hipGetErrorString(e));
        return nullptr;
      }
      c->ent_bytes = bytes;
    }
    return (char *)c->d_ent;
  };
  char *base = room(fixed);
  if (!base) return fixed_base ? SADMC_E_CAPACITY : SADMC_E_HIP;
  EntropyRange range;
  HIPCHECK(c, launch_entropy_range(s->d_ws, s->W, w, (EntropyRange *)base, (EntropyRange *)base + 1, c->stream));
  HIPCHECK(c, hipMemcpyAsync(&range, base, sizeof range, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (!range.any) return fail(c, SADMC_E_STATE, "no running walkers");
  const double gmin = range.gmin;
  const int64_t ng = range.ng;
  if (std::isnan(e_a) || std::isnan(e_b)) {
    e_a = range.lo;
    e_b = range.hi;
  } else if (!(e_a < e_b)) {
    return fail(c, SADMC_E_ARG, "empty energy window (e_a >= e_b)");
  }
  const int64_t ga = (int64_t)std::ceil((e_a - gmin) / w - 0.5), gb = (int64_t)std::floor((e_b - gmin) / w - 0.5) + 1;
  if (!(gb > ga)) return fail(c, SADMC_E_STATE, "no common energy window for the walkers");
  if (e_first) *e_first = gmin + 0.5 * w;
  if (nbins) *nbins = ng;
  if (ng > max_bins) return fail(c, SADMC_E_CAPACITY, "output holds fewer bins than the set covers");
  const int chunks = entropy_chunks(s->W, ng);
  const size_t s_off = fixed, c_off = s_off + up(sizeof(double) * ng), ps_off = c_off + up(sizeof(int64_t) * ng),
               pc_off = ps_off + up(sizeof(double) * ng * chunks), total = pc_off + up(sizeof(int64_t) * ng * chunks);
  base = room(total);
  if (!base) return fixed_base ? SADMC_E_CAPACITY : SADMC_E_HIP;
  double *dS = (double *)(base + s_off), *dPS = (double *)(base + ps_off);
  int64_t *dC = (int64_t *)(base + c_off), *dPC = (int64_t *)(base + pc_off);
  int rc = SADMC_OK;
  if (launch_entropy(s->d_ws, s->d_hot, s->W, s->cap, w, gmin, ga, gb, ng, base + ew_off, chunks, dPS, dPC, dS, dC,
                     c->stream) != hipSuccess ||
      (S && hipMemcpyAsync(S, dS, sizeof(double) * ng, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
      (count && hipMemcpyAsync(count, dC, sizeof(int64_t) * ng, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    rc = fail(c, SADMC_E_HIP, "entropy reduction failed");
  }
  return rc;
}

extern "C" {

int sadmc_mean_entropy(sadmc_ctx *c, int32_t set, double e_a, double e_b, int64_t max_bins, double *e_first,
                       int64_t *nbins, double *S, int64_t *count) {
  if (!c) return SADMC_E_ARG;
  return mean_entropy_impl(c, set, e_a, e_b, max_bins, e_first, nbins, S, count, nullptr, 0);
}

int sadmc_compute_energy(sadmc_ctx *c, int32_t set, int32_t walker, double *e) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !e || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  HIPCHECK(c, hipSetDevice(c->device));
  LaunchArgs la{};
  la.walker = walker;
  la.eout = s->d_eout;
  HIPCHECK(c, launch(s->p.system, s->p.precision, s->p.method, s->KMAX, KK_ENERGY, s->args, la, c->stream));
  HIPCHECK(c, hipMemcpyAsync(e, s->d_eout, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_verify_energy(sadmc_ctx *c, int32_t set, int32_t walker, double *egood, int32_t *ok) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  double eg = 0.0;
  int rc = sadmc_compute_energy(c, set, walker, &eg);
  if (rc) return rc;
  sadmc_walker_state st;
  rc = sadmc_get_walker_state(c, set, walker, &st);
  if (rc) return rc;
  if (egood) *egood = eg;
  // lj.rs:106-108 / wca.rs:178-180 expected_accuracy; Ising keeps the
  // trait's default verify_energy, which checks nothing
  const double nn = (double)s->p.n;
  const double expected = s->p.system == SADMC_LJ    ? std::fabs(st.E) * 1e-14 * nn * nn
                          : s->p.system == SADMC_WCA ? std::fabs(st.E) * 1e-13 * nn * nn
                                                     : INFINITY;
  const bool good = !(std::fabs(eg - st.E) > expected);
  if (ok) *ok = good ? 1 : 0;
  if (!good) {
    char buf[200];
    std::snprintf(buf, sizeof buf, "set %d walker %d: error in E is %.17g when it should be < %.17g", set, walker,
                  eg - st.E, expected);
    return fail(c, SADMC_E_VERIFY, buf);
  }
  return SADMC_OK;
}

int sadmc_system_info(const sadmc_set_params *p, double geom, sadmc_system_info_t *out) {
  if (!p || !out) return SADMC_E_ARG;
  std::memset(out, 0, sizeof *out);
  const int n = p->n;
  if (p->system == SADMC_ISING) {
    if (n < 2 || n > 8) return SADMC_E_ARG;
    out->has_delta_energy = 1;  // ising.rs:76-78
    out->delta_energy = 4.0;
    out->min_moves_to_randomize = (uint64_t)n * (uint64_t)n;  // ising.rs:86-91
    out->dimensionality = out->min_moves_to_randomize;
    out->max_size = 0.5;  // ising.rs:120-122
  } else if (p->system == SADMC_LJ || p->system == SADMC_WCA) {
    if (n < 2) return SADMC_E_ARG;
    out->has_lowest_possible_energy = 1;  // lj.rs:245-248, wca.rs:234-236
    out->lowest_possible_energy = p->system == SADMC_LJ ? -0.5 * (double)n * ((double)n - 1.0) : 0.0;
    out->min_moves_to_randomize = (uint64_t)n;  // lj.rs:280-285, wca.rs:268-273
    out->dimensionality = 3 * (uint64_t)n;
    // lj.rs:375-377 (the sphere radius), wca.rs:354-356 (the box diagonal)
    out->max_size = p->system == SADMC_LJ ? geom : std::sqrt(3.0 * geom * geom);
  } else {
    return SADMC_E_ARG;
  }
  return SADMC_OK;
}

int sadmc_upload_pressure(sadmc_ctx *c, int32_t set, const uint64_t *count, const double *total) {
  if (!c || set < 0 || set >= (int32_t)c->sets.size() || !c->inited || !count || !total)
    return fail(c, SADMC_E_ARG, "bad arguments");
  SetHost &s = c->sets[set];
  if (!s.d_extra) return fail(c, SADMC_E_ARG, "the set does not collect the pressure");
  const size_t nb = (size_t)s.W * s.cap;
  std::vector<PExtra> x(nb);
  for (size_t k = 0; k < nb; k++) {
    x[k].count = count[k];
    x[k].total = total[k];
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(s.d_extra, x.data(), sizeof(PExtra) * nb, hipMemcpyHostToDevice));
  return SADMC_OK;
}

int sadmc_walker_count(const sadmc_ctx *c, int32_t set, int32_t *n, int64_t *cap) {
  if (!c || set < 0 || set >= (int32_t)c->sets.size()) return SADMC_E_ARG;
  if (n) *n = c->sets[set].W;
  if (cap) *cap = c->sets[set].cap;
  return SADMC_OK;
}

int sadmc_get_walker_state(sadmc_ctx *c, int32_t set, int32_t walker, sadmc_walker_state *out) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !out || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(out, s->d_ws + walker, sizeof *out, hipMemcpyDeviceToHost));
  return SADMC_OK;
}

int sadmc_walker_bins(sadmc_ctx *c, int32_t set, int32_t walker, int64_t cap, double *lnw, uint64_t *hist,
                      uint64_t *tf, double *et, double *es, uint8_t *vis, uint64_t *rt, uint64_t *wlh,
                      uint64_t *pc, double *pt) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !c->inited) return fail(c, SADMC_E_ARG, "bad set/walker or not initialised");
  sadmc_walker_state st;
  int rc = sadmc_get_walker_state(c, set, walker, &st);
  if (rc) return rc;
  if (st.len > cap) return fail(c, SADMC_E_ARG, "output arrays too short for len");
  std::vector<sadmc_bin> b(st.len);
  {
    int rc2 = gather_bins(c, *s, walker, st.base, st.len, b.data());
    if (rc2) return rc2;
  }
  std::vector<PExtra> x;
  if (s->d_extra) {
    x.resize(st.len);
    HIPCHECK(c, hipMemcpy(x.data(), s->d_extra + (size_t)walker * s->cap + st.base, sizeof(PExtra) * st.len,
                          hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < st.len; i++) {
    if (lnw) lnw[i] = b[i].lnw;
    if (hist) hist[i] = b[i].histogram;
    if (tf) tf[i] = b[i].t_found;
    if (et) et[i] = b[i].energy_total;
    if (es) es[i] = b[i].energy_squared_total;
    if (vis) vis[i] = (st.bulk_visited || b[i].visit_epoch == st.epoch) ? 1 : 0;
    if (rt) rt[i] = b[i].round_trips;
    if (wlh) wlh[i] = i < st.wl_len ? b[i].wl_hist : 0;
    if (pc) pc[i] = x.empty() ? 0 : x[i].count;
    if (pt) pt[i] = x.empty() ? 0.0 : x[i].total;
  }
  return SADMC_OK;
}

int sadmc_walker_coords(sadmc_ctx *c, int32_t set, int32_t walker, double *coords) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !coords || !c->inited || s->p.system == SADMC_ISING) return fail(c, SADMC_E_ARG, "bad arguments");
  std::vector<double> p((size_t)3 * s->npad);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(p.data(), s->d_pos + (size_t)walker * 3 * s->npad, sizeof(double) * p.size(),
                        hipMemcpyDeviceToHost));
  const int n = s->p.n;
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < n; j++) coords[d * n + j] = p[(size_t)d * s->npad + j];
  return SADMC_OK;
}

int sadmc_walker_spins(sadmc_ctx *c, int32_t set, int32_t walker, int8_t *spins) {
  SetHost *s = get_set(c, set, walker);
  if (!s || !spins || !c->inited || s->p.system != SADMC_ISING) return fail(c, SADMC_E_ARG, "bad arguments");
  uint64_t bits = 0;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(&bits, s->d_spins + walker, sizeof bits, hipMemcpyDeviceToHost));
  const int n = s->p.n;
  for (int b = 0; b < n * n; b++) spins[b] = ((bits >> b) & 1ULL) ? 1 : -1;
  return SADMC_OK;
}

int sadmc_download_raw(sadmc_ctx *c, int32_t set, sadmc_walker_state *states, sadmc_bin *bins) {
  if (!c || set < 0 || set >= (int32_t)c->sets.size() || !c->inited) return fail(c, SADMC_E_ARG, "bad set");
  SetHost &s = c->sets[set];
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (states)
    HIPCHECK(c, hipMemcpy(states, s.d_ws, sizeof(sadmc_walker_state) * s.W, hipMemcpyDeviceToHost));
  if (bins) {
    for (int w = 0; w < s.W; w++) {
      int rc = gather_bins(c, s, w, 0, s.cap, bins + (size_t)w * s.cap);
      if (rc) return rc;
    }
  }
  return SADMC_OK;
}

int sadmc_upload_raw(sadmc_ctx *c, int32_t set, const sadmc_walker_state *states, const sadmc_bin *bins,
                     const double *coords, const int8_t *spins) {
  if (!c || set < 0 || set >= (int32_t)c->sets.size() || !c->inited) return fail(c, SADMC_E_ARG, "bad set");
  SetHost &s = c->sets[set];
  const int n = s.p.n;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (states)
    HIPCHECK(c, hipMemcpy(s.d_ws, states, sizeof(sadmc_walker_state) * s.W, hipMemcpyHostToDevice));
  if (bins) {
    const size_t nb = (size_t)s.W * s.cap;
    std::vector<HotBin> h(nb);
    std::vector<ColdBin> cb(nb);
    std::vector<uint64_t> wl(nb);
    for (size_t k = 0; k < nb; k++) {
      h[k].lnw = bins[k].lnw;
      h[k].hist = bins[k].histogram;
      cb[k].etot = bins[k].energy_total;
      cb[k].esq = bins[k].energy_squared_total;
      cb[k].rt = bins[k].round_trips;
      cb[k].vep = bins[k].visit_epoch;
      cb[k].tfound = bins[k].t_found;
      cb[k].pad = 0;
      wl[k] = bins[k].wl_hist;
    }
    HIPCHECK(c, hipMemcpy(s.d_hot, h.data(), sizeof(HotBin) * nb, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(s.d_cold, cb.data(), sizeof(ColdBin) * nb, hipMemcpyHostToDevice));
    if (s.d_wlh) HIPCHECK(c, hipMemcpy(s.d_wlh, wl.data(), sizeof(uint64_t) * nb, hipMemcpyHostToDevice));
  }
  if (coords && s.p.system != SADMC_ISING) {
    std::vector<double> pos((size_t)s.W * 3 * s.npad, 0.0);
    for (int w = 0; w < s.W; w++)
      for (int d = 0; d < 3; d++)
        for (int j = 0; j < n; j++) pos[((size_t)w * 3 + d) * s.npad + j] = coords[((size_t)w * 3 + d) * n + j];
    HIPCHECK(c, hipMemcpy(s.d_pos, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice));
  }
  if (spins && s.p.system == SADMC_ISING) {
    std::vector<uint64_t> bits(s.W, 0);
    for (int w = 0; w < s.W; w++)
      for (int b = 0; b < n * n; b++)
        if (spins[(size_t)w * n * n + b] > 0) bits[w] |= 1ULL << b;
    HIPCHECK(c, hipMemcpy(s.d_spins, bits.data(), sizeof(uint64_t) * s.W, hipMemcpyHostToDevice));
  }
  int32_t zero = 0;
  HIPCHECK(c, hipMemcpy(s.d_flag, &zero, sizeof zero, hipMemcpyHostToDevice));
  return SADMC_OK;
}

// ------------------------------------------------------ parallel tempering

int sadmc_pt_add(sadmc_ctx *c, const sadmc_set_params *p, int32_t L, int32_t R, const double *temps,
                 const uint64_t *seeds, const double *geom, const double *coords, const int8_t *spins,
                 int32_t canonical_steps, int32_t *pt_id) {
  AllocGuard guard;
  if (!c || !p || !temps || !seeds || L <= 0 || R <= 0) return fail(c, SADMC_E_ARG, "null argument or empty ladder");
  if (canonical_steps <= 0) return fail(c, SADMC_E_ARG, "canonical_steps must be > 0");
  if (p->precision != SADMC_FP64 || p->refresh != 0 || p->move_kind != SADMC_MOVE_CUBE)
    return fail(c, SADMC_E_ARG, "parallel tempering runs the cube move in FP64 with the error-estimate energy rule");
  for (int r = 0; r < R; r++)
    if (!(temps[r] > 0)) return fail(c, SADMC_E_ARG, "replica temperatures must be > 0");
  PtHost t;
  t.L = L;
  t.R = R;
  const int n = p->n;
  if (p->system == SADMC_ISING) {
    if (n < 2 || n > 8) return fail(c, SADMC_E_ARG, "Ising lattice width must be 2..8");
    t.sys = SYS_ISING;
  } else if (p->system == SADMC_LJ || p->system == SADMC_WCA) {
    if (n < 2 || n > 1024) return fail(c, SADMC_E_ARG, "atom count must be 2..1024");
    if (!geom) return fail(c, SADMC_E_ARG, "LJ/WCA need per-ladder geometry");
    if (!(p->move_param > 0)) return fail(c, SADMC_E_ARG, "the move half-width must be > 0");
    t.sys = p->system == SADMC_LJ ? SYS_LJ : SYS_WCA;
    int G = p->lanes > 0 ? p->lanes : auto_lanes(n);
    if (G < 2 || G > 64 || (G & (G - 1))) return fail(c, SADMC_E_ARG, "lanes must be a power of two in 2..64");
    t.G = G;
    t.K = (n + G - 1) / G;
    if (t.K > 16) return fail(c, SADMC_E_ARG, "too few lanes for this many atoms (max 16 atoms per lane)");
    t.KMAX = t.K <= 1 ? 1 : t.K <= 2 ? 2 : t.K <= 4 ? 4 : t.K <= 8 ? 8 : 16;
    t.npad = G * t.K;
  } else {
    return fail(c, SADMC_E_ARG, "unknown system");
  }
  if ((int64_t)R * t.G > 512) return fail(c, SADMC_E_ARG, "replicas * lanes must be <= 512 (one workgroup per ladder)");
  if (pt_lds_bytes(R, t.npad) > 65536) return fail(c, SADMC_E_ARG, "a ladder's systems do not fit 64 KiB of LDS");
  t.threads = ((R * t.G + 63) / 64) * 64;
  const int64_t S = (int64_t)L * R;
  // initial configurations: one per ladder, copied to every replica
  std::vector<uint64_t> bits;
  std::vector<double> pos;
  if (t.sys == SYS_ISING) {
    bits.assign(S, 0);
    std::vector<int8_t> sp((size_t)n * n);
    for (int l = 0; l < L; l++) {
      if (spins)
        std::memcpy(sp.data(), spins + (size_t)l * n * n, (size_t)n * n);
      else
        sadmc_init_spins(n, p->ising_init, seeds[l], sp.data());
      uint64_t b = 0;
      for (int k = 0; k < n * n; k++)
        if (sp[k] > 0) b |= 1ULL << k;
      for (int r = 0; r < R; r++) bits[(size_t)l * R + r] = b;
    }
  } else {
    pos.assign((size_t)S * 3 * t.npad, 0.0);
    std::vector<double> cf((size_t)3 * n);
    for (int l = 0; l < L; l++) {
      if (coords) {
        std::memcpy(cf.data(), coords + (size_t)l * 3 * n, sizeof(double) * 3 * n);
      } else if (sadmc_init_coords(p->system, n, geom[l], seeds[l], cf.data())) {
        return fail(c, SADMC_E_INIT, "initial configuration failed");
      }
      for (int r = 0; r < R; r++)
        for (int d = 0; d < 3; d++)
          for (int j = 0; j < n; j++) pos[(((size_t)l * R + r) * 3 + d) * t.npad + j] = cf[(size_t)d * n + j];
    }
  }
  // replicas: clones of seed_from_u64(seed); the ladder's generator jumped
  std::vector<sadmc_replica> rep(S);
  std::vector<int32_t> slot_of(S);
  std::vector<uint64_t> lrng(2 * (size_t)L);
  for (int l = 0; l < L; l++) {
    uint64_t s0, s1;
    rng_seed(seeds[l], s0, s1);
    for (int r = 0; r < R; r++) {
      sadmc_replica &q = rep[(size_t)l * R + r];
      std::memset(&q, 0, sizeof q);
      q.rng[0] = s0;
      q.rng[1] = s1;
      q.T = temps[r];
      q.geom = geom ? geom[l] : 0.0;
      slot_of[(size_t)l * R + r] = r;
    }
    rng_jump(s0, s1);
    lrng[2 * l] = s0;
    lrng[2 * l + 1] = s1;
  }
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, guard.alloc(&t.d_rep, sizeof(sadmc_replica) * S));
  HIPCHECK(c, guard.alloc(&t.d_slot_e, sizeof(double) * 2 * S));
  HIPCHECK(c, guard.alloc(&t.d_slot_of, sizeof(int32_t) * S));
  HIPCHECK(c, guard.alloc(&t.d_lrng, sizeof(uint64_t) * 2 * L));
  HIPCHECK(c, hipMemcpyAsync(t.d_rep, rep.data(), sizeof(sadmc_replica) * S, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(t.d_slot_of, slot_of.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(t.d_lrng, lrng.data(), sizeof(uint64_t) * 2 * L, hipMemcpyHostToDevice, c->stream));
  if (t.sys == SYS_ISING) {
    HIPCHECK(c, guard.alloc(&t.d_spins, sizeof(uint64_t) * S));
    HIPCHECK(c, hipMemcpyAsync(t.d_spins, bits.data(), sizeof(uint64_t) * S, hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHECK(c, guard.alloc(&t.d_pos, sizeof(double) * pos.size()));
    HIPCHECK(c, hipMemcpyAsync(t.d_pos, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, guard.alloc(&t.d_geom, sizeof(double) * L));
    HIPCHECK(c, hipMemcpyAsync(t.d_geom, geom, sizeof(double) * L, hipMemcpyHostToDevice, c->stream));
  }
  SetArgs &a = t.args;
  std::memset(&a, 0, sizeof a);
  a.pos = t.d_pos;
  a.spins = t.d_spins;
  a.W = (int32_t)S;
  a.n = n;
  a.npad = t.npad;
  a.G = t.G;
  a.K = t.K;
  a.method = SADMC_CANONICAL;
  a.precision = SADMC_FP64;
  a.move_kind = SADMC_MOVE_CUBE;
  a.rc2 = p->rc2;
  a.width = 1.0;
  a.inv_width = 1.0;
  const uint64_t nn = (uint64_t)n;
  a.zone_n = ~0ULL - ((~0ULL - nn + 1) % nn);
  a.zone_site = (nn << __builtin_clzll(nn)) - 1;
  PtArgs &pa = t.pa;
  pa.rep = t.d_rep;
  pa.slot_e = t.d_slot_e;
  pa.slot_of = t.d_slot_of;
  pa.lrng = t.d_lrng;
  pa.geom = t.d_geom;
  pa.R = R;
  const int64_t steps = (t.sys == SYS_ISING ? (int64_t)n * n : (int64_t)n) * canonical_steps;
  if (steps > INT32_MAX) return fail(c, SADMC_E_ARG, "too many moves per round");
  pa.steps = (int32_t)steps;
  pa.d = p->move_param;
  pa.sc = t.sys == SYS_ISING ? 0.0 : uniform_scale(-p->move_param, p->move_param);
  HIPCHECK(c, launch_pt(t.sys, t.KMAX, a, pa, L, t.threads, 0, 1, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  guard.release();
  c->pts.push_back(t);
  if (pt_id) *pt_id = (int32_t)c->pts.size() - 1;
  return SADMC_OK;
}

static PtHost *pt_get(sadmc_ctx *c, int32_t pt) {
  if (!c || pt < 0 || pt >= (int32_t)c->pts.size()) return nullptr;
  return &c->pts[pt];
}

int sadmc_pt_run(sadmc_ctx *c, int32_t pt, uint64_t rounds) {
  PtHost *t = pt_get(c, pt);
  if (!t) return fail(c, SADMC_E_ARG, "no such tempering ladder set");
  if (!rounds) return SADMC_OK;
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHECK(c, launch_pt(t->sys, t->KMAX, t->args, t->pa, t->L, t->threads, rounds, 0, c->stream));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return SADMC_OK;
}

int sadmc_pt_replicas(sadmc_ctx *c, int32_t pt, sadmc_replica *out) {
  PtHost *t = pt_get(c, pt);
  if (!t || !out) return fail(c, SADMC_E_ARG, "no such tempering ladder set or null output");
  HIPCHECK(c, hipMemcpyAsync(out, t->d_rep, sizeof(sadmc_replica) * t->L * t->R, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

static int pt_slot(sadmc_ctx *c, PtHost *t, int32_t l, int32_t r, int64_t *slot) {
  if (l < 0 || l >= t->L || r < 0 || r >= t->R) return fail(c, SADMC_E_ARG, "ladder or replica out of range");
  int32_t s = 0;
  HIPCHECK(c, hipMemcpyAsync(&s, t->d_slot_of + (int64_t)l * t->R + r, sizeof s, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  *slot = (int64_t)l * t->R + s;
  return SADMC_OK;
}

int sadmc_pt_coords(sadmc_ctx *c, int32_t pt, int32_t l, int32_t r, double *coords) {
  PtHost *t = pt_get(c, pt);
  if (!t || !coords) return fail(c, SADMC_E_ARG, "no such tempering ladder set or null output");
  if (t->sys == SYS_ISING) return fail(c, SADMC_E_ARG, "Ising replicas have spins, not coordinates");
  int64_t slot;
  int rc = pt_slot(c, t, l, r, &slot);
  if (rc) return rc;
  const int n = t->args.n;
  for (int d = 0; d < 3; d++)
    HIPCHECK(c, hipMemcpyAsync(coords + (size_t)d * n, t->d_pos + (slot * 3 + d) * t->npad, sizeof(double) * n,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_pt_spins(sadmc_ctx *c, int32_t pt, int32_t l, int32_t r, int8_t *spins) {
  PtHost *t = pt_get(c, pt);
  if (!t || !spins) return fail(c, SADMC_E_ARG, "no such tempering ladder set or null output");
  if (t->sys != SYS_ISING) return fail(c, SADMC_E_ARG, "only Ising replicas have spins");
  int64_t slot;
  int rc = pt_slot(c, t, l, r, &slot);
  if (rc) return rc;
  uint64_t b = 0;
  HIPCHECK(c, hipMemcpyAsync(&b, t->d_spins + slot, sizeof b, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const int n = t->args.n;
  for (int k = 0; k < n * n; k++) spins[k] = ((b >> k) & 1ULL) ? 1 : -1;
  return SADMC_OK;
}

int sadmc_pt_ladder_rng(sadmc_ctx *c, int32_t pt, uint64_t *out) {
  PtHost *t = pt_get(c, pt);
  if (!t || !out) return fail(c, SADMC_E_ARG, "no such tempering ladder set or null output");
  HIPCHECK(c, hipMemcpyAsync(out, t->d_lrng, sizeof(uint64_t) * 2 * t->L, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

// --------------------------------------------------- binning-path walkers

int sadmc_bin_add_set(sadmc_ctx *c, const sadmc_set_params *p, int32_t W, const uint64_t *seeds, const double *geom,
                      const double *coords, const int8_t *spins, int32_t *set_id) {
  AllocGuard guard;
  if (!c || !p || W <= 0 || !seeds) return fail(c, SADMC_E_ARG, "null argument or no walkers");
  if (p->method != SADMC_SAD && p->method != SADMC_SAMC && p->method != SADMC_WL && p->method != SADMC_INV_T_WL)
    return fail(c, SADMC_E_ARG, "the binning walker runs SAD, SAMC, WL or 1/t-WL");
  if (p->precision != SADMC_FP64 || p->refresh != 0 || p->move_kind != SADMC_MOVE_CUBE)
    return fail(c, SADMC_E_ARG, "the binning walker runs the cube move in FP64 with the error-estimate energy rule");
  if (p->binning != SADMC_BIN_HISTOGRAM && p->binning != SADMC_BIN_LINEAR)
    return fail(c, SADMC_E_ARG, "unknown binning");
  BinSetHost b;
  b.p = *p;
  b.W = W;
  const int n = p->n;
  if (p->system == SADMC_ISING) {
    if (n < 2 || n > 8) return fail(c, SADMC_E_ARG, "Ising lattice width must be 2..8");
    b.sys = SYS_ISING;
  } else if (p->system == SADMC_LJ || p->system == SADMC_WCA) {
    if (n < 2 || n > 1024) return fail(c, SADMC_E_ARG, "atom count must be 2..1024");
    if (!geom) return fail(c, SADMC_E_ARG, "LJ/WCA need per-walker geometry");
    if (p->move_plan == SADMC_TRANSLATION_SCALE && !(p->move_param > 0))
      return fail(c, SADMC_E_ARG, "translation scale must be > 0");
    b.sys = p->system == SADMC_LJ ? SYS_LJ : SYS_WCA;
    int G = p->lanes > 0 ? p->lanes : auto_lanes(n);
    if (G < 2 || G > 64 || (G & (G - 1))) return fail(c, SADMC_E_ARG, "lanes must be a power of two in 2..64");
    b.G = G;
    b.K = (n + G - 1) / G;
    if (b.K > 16) return fail(c, SADMC_E_ARG, "too few lanes for this many atoms (max 16 atoms per lane)");
    b.KMAX = b.K <= 1 ? 1 : b.K <= 2 ? 2 : b.K <= 4 ? 4 : b.K <= 8 ? 8 : 16;
    b.npad = G * b.K;
  } else {
    return fail(c, SADMC_E_ARG, "unknown system");
  }
  // the window: BinningParams' default width 1.0, bounds as for the main sets
  sadmc_set_params q = *p;
  const double width = p->energy_bin > 0 ? p->energy_bin : 1.0;
  q.energy_bin = width;
  double e_lo, e_hi;
  sadmc_bin_window(&q, geom ? geom[0] : 0.0, &e_lo, &e_hi, &b.cap);
  if (p->high_resolution_de < 0 || std::isnan(p->high_resolution_de))
    return fail(c, SADMC_E_ARG, "high_resolution_de must be >= 0");
  if (p->high_resolution_de > 0) {
    const double nb = std::ceil((e_hi - e_lo) / p->high_resolution_de) + 8;
    if (!(nb < 2147483648.0)) return fail(c, SADMC_E_ARG, "high_resolution_de is too fine for the energy window");
    b.hr_cap = (int64_t)nb;
  }
  std::vector<uint64_t> bits;
  std::vector<double> pos;
  if (b.sys == SYS_ISING) {
    bits.assign(W, 0);
    std::vector<int8_t> sp((size_t)n * n);
    for (int w = 0; w < W; w++) {
      if (spins)
        std::memcpy(sp.data(), spins + (size_t)w * n * n, (size_t)n * n);
      else
        sadmc_init_spins(n, p->ising_init, seeds[w], sp.data());
      for (int k = 0; k < n * n; k++)
        if (sp[k] > 0) bits[w] |= 1ULL << k;
    }
  } else {
    pos.assign((size_t)W * 3 * b.npad, 0.0);
    std::vector<double> cf((size_t)3 * n);
    for (int w = 0; w < W; w++) {
      if (coords) {
        std::memcpy(cf.data(), coords + (size_t)w * 3 * n, sizeof(double) * 3 * n);
      } else if (sadmc_init_coords(p->system, n, geom[w], seeds[w], cf.data())) {
        return fail(c, SADMC_E_INIT, "initial configuration failed");
      }
      for (int d = 0; d < 3; d++)
        for (int j = 0; j < n; j++) pos[((size_t)w * 3 + d) * b.npad + j] = cf[(size_t)d * n + j];
    }
  }
  std::vector<sadmc_bwalker_state> ws(W);
  for (int w = 0; w < W; w++) {
    sadmc_bwalker_state &s = ws[w];
    std::memset(&s, 0, sizeof s);
    rng_seed(seeds[w], s.rng[0], s.rng[1]);
    s.width = width;
    s.cap = b.cap;
    s.geom = geom ? geom[w] : 0.0;
    s.method = p->method == SADMC_INV_T_WL ? SADMC_WL : p->method;
    s.inv_t = p->method == SADMC_INV_T_WL;
    s.has_min_gamma = p->method == SADMC_WL && p->has_min_gamma;
    s.min_gamma = p->method_param;
    s.min_T = p->method_param;
    s.t0 = p->method_param;
    s.gamma = 1.0;
  }
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, guard.alloc(&b.d_ws, sizeof(sadmc_bwalker_state) * W));
  HIPCHECK(c, guard.alloc(&b.d_bins, sizeof(sadmc_bbin) * (size_t)W * b.cap));
  HIPCHECK(c, guard.alloc(&b.d_flag, sizeof(int32_t)));
  HIPCHECK(c, hipMemsetAsync(b.d_bins, 0, sizeof(sadmc_bbin) * (size_t)W * b.cap, c->stream));
  HIPCHECK(c, hipMemsetAsync(b.d_flag, 0, sizeof(int32_t), c->stream));
  HIPCHECK(c, hipMemcpyAsync(b.d_ws, ws.data(), sizeof(sadmc_bwalker_state) * W, hipMemcpyHostToDevice, c->stream));
  if (b.hr_cap) {
    HIPCHECK(c, guard.alloc(&b.d_hrs, sizeof(sadmc_hires) * W));
    HIPCHECK(c, guard.alloc(&b.d_hr, sizeof(uint64_t) * (size_t)W * b.hr_cap));
    HIPCHECK(c, hipMemsetAsync(b.d_hrs, 0, sizeof(sadmc_hires) * W, c->stream));
    HIPCHECK(c, hipMemsetAsync(b.d_hr, 0, sizeof(uint64_t) * (size_t)W * b.hr_cap, c->stream));
  }
  if (b.sys == SYS_ISING) {
    HIPCHECK(c, guard.alloc(&b.d_spins, sizeof(uint64_t) * W));
    HIPCHECK(c, hipMemcpyAsync(b.d_spins, bits.data(), sizeof(uint64_t) * W, hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHECK(c, guard.alloc(&b.d_pos, sizeof(double) * pos.size()));
    HIPCHECK(c, hipMemcpyAsync(b.d_pos, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice, c->stream));
  }
  SetArgs &a = b.args;
  std::memset(&a, 0, sizeof a);
  a.pos = b.d_pos;
  a.spins = b.d_spins;
  a.flag = b.d_flag;
  a.W = W;
  a.n = n;
  a.npad = b.npad;
  a.G = b.G;
  a.K = b.K;
  a.method = SADMC_CANONICAL;
  a.precision = SADMC_FP64;
  a.move_kind = SADMC_MOVE_CUBE;
  a.collect_pressure = p->collect_pressure;
  a.rc2 = p->rc2;
  a.width = 1.0;
  a.inv_width = 1.0;
  const uint64_t nn = (uint64_t)n;
  a.zone_n = ~0ULL - ((~0ULL - nn + 1) % nn);
  a.zone_site = (nn << __builtin_clzll(nn)) - 1;
  BinArgs &ba = b.ba;
  ba.ws = b.d_ws;
  ba.bins = b.d_bins;
  ba.flag = b.d_flag;
  ba.cap = b.cap;
  ba.e_lo = e_lo;
  ba.move_plan = p->move_plan;
  ba.has_min_e = p->has_min_e;
  ba.has_max_e = p->has_max_e;
  ba.collect_pressure = p->collect_pressure;
  ba.move_param = p->move_param;
  ba.min_e = p->min_e;
  ba.max_e = p->max_e;
  ba.linear = p->binning == SADMC_BIN_LINEAR;
  ba.hrs = b.d_hrs;
  ba.hr = b.d_hr;
  ba.hr_cap = b.hr_cap;
  ba.hr_de = p->high_resolution_de;
  HIPCHECK(c, launch_binning(b.sys, b.KMAX, ba, a, 0, 1, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  guard.release();
  c->bsets.push_back(b);
  if (set_id) *set_id = (int32_t)c->bsets.size() - 1;
  return SADMC_OK;
}

static BinSetHost *bset_get(sadmc_ctx *c, int32_t set) {
  if (!c || set < 0 || set >= (int32_t)c->bsets.size()) return nullptr;
  return &c->bsets[set];
}

int sadmc_bin_run(sadmc_ctx *c, int32_t set, uint64_t moves) {
  BinSetHost *b = bset_get(c, set);
  if (!b) return fail(c, SADMC_E_ARG, "no such binning walker set");
  HIPCHECK(c, hipSetDevice(c->device));
  if (moves) {
    HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHECK(c, launch_binning(b->sys, b->KMAX, b->ba, b->args, moves, 0, c->stream));
    HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
  }
  int32_t flag = 0;
  HIPCHECK(c, hipMemcpyAsync(&flag, b->d_flag, sizeof flag, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (!flag) return SADMC_OK;
  std::vector<sadmc_bwalker_state> ws(b->W);
  HIPCHECK(c, hipMemcpy(ws.data(), b->d_ws, sizeof(sadmc_bwalker_state) * b->W, hipMemcpyDeviceToHost));
  for (int w = 0; w < b->W; w++) {
    const int st = ws[w].status;
    if (!st) continue;
    char buf[160];
    std::snprintf(buf, sizeof buf, "binning set %d walker %d stopped with status %d at move %llu", set, w, st,
                  (unsigned long long)ws[w].moves);
    return fail(c, st == SADMC_WS_CAPACITY ? SADMC_E_CAPACITY : st == SADMC_WS_VERIFY ? SADMC_E_VERIFY : SADMC_E_INDEX,
                buf);
  }
  return SADMC_OK;
}

int sadmc_bin_state(sadmc_ctx *c, int32_t set, int32_t w, sadmc_bwalker_state *out) {
  BinSetHost *b = bset_get(c, set);
  if (!b || !out || w < 0 || w >= b->W) return fail(c, SADMC_E_ARG, "no such binning walker or null output");
  HIPCHECK(c, hipMemcpyAsync(out, b->d_ws + w, sizeof *out, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_bin_bins(sadmc_ctx *c, int32_t set, int32_t w, int64_t max_bins, sadmc_bbin *out) {
  sadmc_bwalker_state st;
  int rc = sadmc_bin_state(c, set, w, &st);
  if (rc) return rc;
  if (st.len > max_bins || !out) return fail(c, SADMC_E_CAPACITY, "output holds fewer bins than the walker has");
  BinSetHost *b = bset_get(c, set);
  if (st.len > 0)
    HIPCHECK(c, hipMemcpyAsync(out, b->d_bins + (int64_t)w * b->cap + st.base, sizeof(sadmc_bbin) * st.len,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_bin_hires(sadmc_ctx *c, int32_t set, int32_t w, sadmc_hires *out, int64_t max_bins, uint64_t *counts) {
  BinSetHost *b = bset_get(c, set);
  if (!b || !out || w < 0 || w >= b->W) return fail(c, SADMC_E_ARG, "no such binning walker or null output");
  if (!b->hr_cap) return fail(c, SADMC_E_ARG, "the set keeps no high_resolution histogram");
  HIPCHECK(c, hipMemcpyAsync(out, b->d_hrs + w, sizeof *out, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (!counts) return SADMC_OK;
  if (out->len > max_bins) return fail(c, SADMC_E_CAPACITY, "output holds fewer bins than the side histogram has");
  if (out->len > 0)
    HIPCHECK(c, hipMemcpyAsync(counts, b->d_hr + (int64_t)w * b->hr_cap + out->base, sizeof(uint64_t) * out->len,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_bin_coords(sadmc_ctx *c, int32_t set, int32_t w, double *coords) {
  BinSetHost *b = bset_get(c, set);
  if (!b || !coords || w < 0 || w >= b->W) return fail(c, SADMC_E_ARG, "no such binning walker or null output");
  if (b->sys == SYS_ISING) return fail(c, SADMC_E_ARG, "Ising walkers have spins, not coordinates");
  const int n = b->p.n;
  for (int d = 0; d < 3; d++)
    HIPCHECK(c, hipMemcpyAsync(coords + (size_t)d * n, b->d_pos + ((int64_t)w * 3 + d) * b->npad, sizeof(double) * n,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_bin_spins(sadmc_ctx *c, int32_t set, int32_t w, int8_t *spins) {
  BinSetHost *b = bset_get(c, set);
  if (!b || !spins || w < 0 || w >= b->W) return fail(c, SADMC_E_ARG, "no such binning walker or null output");
  if (b->sys != SYS_ISING) return fail(c, SADMC_E_ARG, "only Ising walkers have spins");
  uint64_t bits = 0;
  HIPCHECK(c, hipMemcpyAsync(&bits, b->d_spins + w, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const int n = b->p.n;
  for (int k = 0; k < n * n; k++) spins[k] = ((bits >> k) & 1ULL) ? 1 : -1;
  return SADMC_OK;
}

// ------------------------------------------------ energy-ceiling replicas

namespace {

// A system on the host for from_params: positions (SoA) or spins, and the
// energy summed in the device's lane-group order.
struct HostSys {
  int sys = 0, n = 0, G = 1;
  double radius = 0.0;     // LJ sphere radius, or the WCA box length
  double rc2 = 0.0;        // WCA cutoff squared
  std::vector<double> c;   // 3n
  std::vector<int8_t> s;   // n*n
  double E = 0.0, err = 0.0;

  // lane l sums the pairs (i, j < i) of its atoms i = l, l + G, ... < m; the
  // lanes meet in an xor butterfly
  double pair_energy(int m) const {
    const double *x = c.data(), *y = x + n, *z = x + 2 * n;
    const double L = radius, h = 0.5 * radius;
    double acc[64];
    for (int l = 0; l < G; l++) acc[l] = 0.0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < i; j++)
        acc[i % G] += sys == SYS_LJ ? lj_pair(norm2(x[i] - x[j], y[i] - y[j], z[i] - z[j]))
                                    : wca_pair(norm2(min_image(x[i] - x[j], L, h), min_image(y[i] - y[j], L, h),
                                                     min_image(z[i] - z[j], L, h)),
                                               rc2);
    return butterfly(acc);
  }
  double butterfly(double *acc) const {
    for (int d = 1; d < G; d <<= 1) {
      double t[64];
      for (int l = 0; l < G; l++) t[l] = acc[l] + acc[l ^ d];
      for (int l = 0; l < G; l++) acc[l] = t[l];
    }
    return acc[0];
  }

  double energy() const {
    if (sys == SYS_ISING) {
      const int N = n;
      double e = 0.0;
      for (int i1 = 0; i1 < N; i1++)
        for (int j1 = 0; j1 < N; j1++) {
          const int j2 = j1 + 1 == N ? 0 : j1 + 1;
          const int i2 = i1 + 1 == N ? 0 : i1 + 1;
          const int nt = s[i1 + j2 * N] + s[i2 + j1 * N];
          e += (double)nt * (double)s[i1 + j1 * N];
        }
      return e;
    }
    return pair_energy(n);
  }
  // System::randomize (ising.rs:79-85, lj.rs:262-279, wca.rs:251-267)
  double randomize(uint64_t &s0, uint64_t &s1) {
    if (sys == SYS_ISING) {
      for (int k = 0; k < n * n; k++) s[k] = (int8_t)(((rng_next(s0, s1) & 1ULL) ? 1 : 0) * 2 - 1);
    } else if (sys == SYS_WCA) {
      // the atoms are added back one at a time, each through set_energy
      // (wca.rs:105-116, 164-180) with the atom count after the addition
      const double L = radius, h = 0.5 * radius, scL = uniform_scale(0.0, L);
      double *x = c.data(), *y = x + n, *z = x + 2 * n;
      for (int i = 0; i < n; i++) {
        const double px = put_in_cell(uniform_sample(s0, s1, 0.0, scL), L);
        const double py = put_in_cell(uniform_sample(s0, s1, 0.0, scL), L);
        const double pz = put_in_cell(uniform_sample(s0, s1, 0.0, scL), L);
        double acc[64];
        for (int l = 0; l < G; l++) acc[l] = 0.0;
        for (int j = 0; j < i; j++)
          acc[j % G] += wca_pair(norm2(min_image(x[j] - px, L, h), min_image(y[j] - py, L, h),
                                       min_image(z[j] - pz, L, h)),
                                 rc2);
        const double dabse = butterfly(acc);
        x[i] = px;
        y[i] = py;
        z[i] = pz;
        const double new_e = E + dabse, cnt = (double)(i + 1);
        const double single = dabse > std::fabs(new_e) ? 1e-14 * dabse * cnt : 1e-14 * std::fabs(new_e) * cnt;
        err += single * cnt;
        if (err > std::fabs(new_e) * 1e-13 * cnt * cnt) {
          E = pair_energy(i + 1);
          err = 1e-15 * E * cnt;
        } else {
          E = new_e;
        }
      }
    } else {
      const double sc = uniform_scale(-1.0, 1.0);
      double *x = c.data(), *y = x + n, *z = x + 2 * n;
      for (int i = 0; i < n; i++) {
        double a, b, d;
        for (;;) {
          a = uniform_sample(s0, s1, -1.0, sc);
          b = uniform_sample(s0, s1, -1.0, sc);
          d = uniform_sample(s0, s1, -1.0, sc);
          if (norm2(a, b, d) < 1.0) break;
        }
        x[i] = a * radius;
        y[i] = b * radius;
        z[i] = d * radius;
      }
    }
    E = energy();
    return E;
  }
};

}  // namespace

int sadmc_zmc_add(sadmc_ctx *c, const sadmc_set_params *p, int32_t S, const uint64_t *seeds, const double *geom,
                  const double *coords, const int8_t *spins, double min_T, int64_t isbnb, int32_t RM,
                  int32_t *zmc_id) {
  AllocGuard guard;
  if (!c || !p || S <= 0 || !seeds) return fail(c, SADMC_E_ARG, "null argument or no runs");
  if (RM < 2) return fail(c, SADMC_E_ARG, "max_replicas must be >= 2");
  if (p->precision != SADMC_FP64 || p->refresh != 0 || p->move_kind != SADMC_MOVE_CUBE)
    return fail(c, SADMC_E_ARG, "energy-ceiling replicas run the cube move in FP64 with the error-estimate rule");
  ZmcHost z;
  z.S = S;
  z.RM = RM;
  const int n = p->n;
  z.n = n;
  if (p->system == SADMC_ISING) {
    if (n < 2 || n > 8) return fail(c, SADMC_E_ARG, "Ising lattice width must be 2..8");
    z.sys = SYS_ISING;
  } else if (p->system == SADMC_LJ || p->system == SADMC_WCA) {
    if (n < 2 || n > 1024) return fail(c, SADMC_E_ARG, "atom count must be 2..1024");
    if (!geom) return fail(c, SADMC_E_ARG, "LJ/WCA runs need a sphere radius or box length per run");
    z.sys = p->system == SADMC_LJ ? SYS_LJ : SYS_WCA;
    if (z.sys == SYS_WCA)
      for (int s = 0; s < S; s++)
        if (!(geom[s] > 0.0) || geom[s] * geom[s] < p->rc2)
          return fail(c, SADMC_E_ARG, "each WCA box must be at least the cutoff wide");
    int G = p->lanes > 0 ? p->lanes : auto_lanes(n);
    if (G < 2 || G > 64 || (G & (G - 1))) return fail(c, SADMC_E_ARG, "lanes must be a power of two in 2..64");
    z.G = G;
    z.K = (n + G - 1) / G;
    if (z.K > 16) return fail(c, SADMC_E_ARG, "too few lanes for this many atoms (max 16 atoms per lane)");
    z.KMAX = z.K <= 1 ? 1 : z.K <= 2 ? 2 : z.K <= 4 ? 4 : z.K <= 8 ? 8 : 16;
    z.npad = G * z.K;
  } else {
    return fail(c, SADMC_E_ARG, "unknown system");
  }
  if (zmc_lds_bytes(RM, z.npad) > 65536) return fail(c, SADMC_E_ARG, "a run's systems do not fit 64 KiB of LDS");
  z.threads = 64;  // one wave per run; its lane groups take the slots in turn
  const int steps = z.sys == SYS_ISING ? n * n : n;
  const double dim = z.sys == SYS_ISING ? (double)n * n : 3.0 * n;
  const int64_t SR = (int64_t)S * RM;
  std::vector<ZSim> sim(S);
  std::vector<sadmc_zreplica> rep(SR);
  std::vector<double> slot(3 * SR, 0.0), med((size_t)S * 4096, 0.0);
  std::vector<int32_t> slot_of(SR);
  std::vector<double> pos(z.sys == SYS_ISING ? 0 : (size_t)SR * 3 * z.npad, 0.0);
  std::vector<uint64_t> bits(z.sys == SYS_ISING ? SR : 0, 0);
  std::memset(rep.data(), 0, sizeof(sadmc_zreplica) * SR);
  const int MAX_INIT = 1 << 15;
  std::vector<int> init_failed(S, 0);
  // MC::from_params (energy_replicas.rs:344-395) for run s; runs are
  // independent, so they are spread over host threads
  auto init_run = [&](int s) {
    std::vector<double> energies(MAX_INIT);
    HostSys sys;
    sys.sys = z.sys;
    sys.n = n;
    sys.G = z.G;
    sys.radius = geom ? geom[s] : 0.0;
    sys.rc2 = p->rc2;
    if (z.sys == SYS_ISING) {
      sys.s.resize((size_t)n * n);
      if (spins)
        std::memcpy(sys.s.data(), spins + (size_t)s * n * n, (size_t)n * n);
      else
        sadmc_init_spins(n, p->ising_init, seeds[s], sys.s.data());
    } else {
      sys.c.resize((size_t)3 * n);
      if (coords)
        std::memcpy(sys.c.data(), coords + (size_t)s * 3 * n, sizeof(double) * 3 * n);
      else if (sadmc_init_coords(p->system, n, sys.radius, seeds[s], sys.c.data())) {
        init_failed[s] = 1;
        return;
      }
    }
    sys.E = sys.energy();
    uint64_t s0, s1;
    rng_seed(seeds[s], s0, s1);
    for (int k = 0; k < MAX_INIT; k++) energies[k] = sys.randomize(s0, s1);
    HostSys high = sys;
    high.randomize(s0, s1);
    while (sys.E > energies[MAX_INIT / 2]) sys.randomize(s0, s1);
    std::sort(energies.begin(), energies.end());
    const double e2 = energies[MAX_INIT / 2], e4 = energies[MAX_INIT / 4];
    // max_size: ising.rs:120-122, lj.rs:375-377, wca.rs:354-356
    const double max_size = z.sys == SYS_ISING ? 0.5
                            : z.sys == SYS_LJ  ? sys.radius
                                               : std::sqrt(norm2(sys.radius, sys.radius, sys.radius));
    for (int r = 0; r < 2; r++) {  // Replica::new (energy_replicas.rs:145-168)
      sadmc_zreplica &q = rep[(size_t)s * RM + r];
      q.max_energy = r == 0 ? INFINITY : e2;
      q.cutoff_energy = r == 0 ? e2 : e4;
      q.lowest_max_energy = q.max_energy;
      q.unique_visitors = 1;
      q.collecting_data = 1;
      q.rng[0] = s0;
      q.rng[1] = s1;
      q.translation_scale = max_size;
    }
    rng_jump(s0, s1);
    sim[s].rng[0] = s0;
    sim[s].rng[1] = s1;
    sim[s].moves = 0;
    sim[s].R = 2;
    sim[s].status = 0;
    sim[s].med_len = 1;
    med[(size_t)s * 4096] = e4;
    for (int r = 0; r < RM; r++) slot_of[(size_t)s * RM + r] = r;
    const HostSys *systems[2] = {&high, &sys};
    for (int r = 0; r < 2; r++) {
      const size_t w = (size_t)s * RM + r;
      slot[3 * w] = systems[r]->E;
      slot[3 * w + 1] = systems[r]->err;
      slot[3 * w + 2] = rep[w].lowest_max_energy;
      if (z.sys == SYS_ISING) {
        for (int k = 0; k < n * n; k++)
          if (systems[r]->s[k] > 0) bits[w] |= 1ULL << k;
      } else {
        for (int d = 0; d < 3; d++)
          for (int j = 0; j < n; j++) pos[(w * 3 + d) * z.npad + j] = systems[r]->c[(size_t)d * n + j];
      }
    }
  };
  {
    const int nt = std::max(1, std::min<int>(S, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t]() {
        for (int s = t; s < S; s += nt) init_run(s);
      });
    for (auto &th : pool) th.join();
  }
  for (int s = 0; s < S; s++)
    if (init_failed[s]) return fail(c, SADMC_E_INIT, "initial configuration failed");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, guard.alloc(&z.d_sim, sizeof(ZSim) * S));
  HIPCHECK(c, guard.alloc(&z.d_rep, sizeof(sadmc_zreplica) * SR));
  HIPCHECK(c, guard.alloc(&z.d_slot, sizeof(double) * 3 * SR));
  HIPCHECK(c, guard.alloc(&z.d_med, sizeof(double) * med.size()));
  HIPCHECK(c, guard.alloc(&z.d_slot_of, sizeof(int32_t) * SR));
  HIPCHECK(c, hipMemcpyAsync(z.d_sim, sim.data(), sizeof(ZSim) * S, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(z.d_rep, rep.data(), sizeof(sadmc_zreplica) * SR, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(z.d_slot, slot.data(), sizeof(double) * 3 * SR, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(z.d_med, med.data(), sizeof(double) * med.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(z.d_slot_of, slot_of.data(), sizeof(int32_t) * SR, hipMemcpyHostToDevice, c->stream));
  if (z.sys == SYS_ISING) {
    HIPCHECK(c, guard.alloc(&z.d_spins, sizeof(uint64_t) * SR));
    HIPCHECK(c, hipMemcpyAsync(z.d_spins, bits.data(), sizeof(uint64_t) * SR, hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHECK(c, guard.alloc(&z.d_pos, sizeof(double) * pos.size()));
    HIPCHECK(c, hipMemcpyAsync(z.d_pos, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, guard.alloc(&z.d_geom, sizeof(double) * S));
    HIPCHECK(c, hipMemcpyAsync(z.d_geom, geom, sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  SetArgs &a = z.args;
  std::memset(&a, 0, sizeof a);
  a.pos = z.d_pos;
  a.spins = z.d_spins;
  a.W = (int32_t)SR;
  a.n = n;
  a.npad = z.npad;
  a.G = z.G;
  a.K = z.K;
  a.method = SADMC_CANONICAL;
  a.precision = SADMC_FP64;
  a.move_kind = SADMC_MOVE_CUBE;
  a.rc2 = p->rc2;
  a.width = 1.0;
  a.inv_width = 1.0;
  const uint64_t nn = (uint64_t)n;
  a.zone_n = ~0ULL - ((~0ULL - nn + 1) % nn);
  a.zone_site = (nn << __builtin_clzll(nn)) - 1;
  ZArgs &za = z.za;
  za.sim = z.d_sim;
  za.rep = z.d_rep;
  za.slot = z.d_slot;
  za.slot_of = z.d_slot_of;
  za.med = z.d_med;
  za.geom = z.d_geom;
  za.RMAX = RM;
  za.steps = steps;
  za.isbnb = isbnb > 0 ? (uint64_t)isbnb : 64;
  za.min_T = min_T;
  za.shrink = std::pow(0.5, 1.0 / dim);
  guard.release();
  c->zmcs.push_back(z);
  if (zmc_id) *zmc_id = (int32_t)c->zmcs.size() - 1;
  return SADMC_OK;
}

static ZmcHost *zmc_get(sadmc_ctx *c, int32_t id) {
  if (!c || id < 0 || id >= (int32_t)c->zmcs.size()) return nullptr;
  return &c->zmcs[id];
}

int sadmc_zmc_run(sadmc_ctx *c, int32_t id, uint64_t rounds) {
  ZmcHost *z = zmc_get(c, id);
  if (!z) return fail(c, SADMC_E_ARG, "no such replica set");
  HIPCHECK(c, hipSetDevice(c->device));
  if (rounds) {
    HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHECK(c, launch_zmc(z->sys, z->KMAX, z->za, z->args, z->S, z->threads, rounds, c->stream));
    HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
  }
  std::vector<ZSim> sim(z->S);
  HIPCHECK(c, hipMemcpyAsync(sim.data(), z->d_sim, sizeof(ZSim) * z->S, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < z->S; s++)
    if (sim[s].status) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "replica run %d stopped: its lowest replica needs to split beyond %d replicas",
                    s, z->RM);
      return fail(c, SADMC_E_CAPACITY, buf);
    }
  return SADMC_OK;
}

int sadmc_zmc_replicas(sadmc_ctx *c, int32_t id, int32_t s, int32_t max_out, sadmc_zreplica *out, int32_t *count) {
  ZmcHost *z = zmc_get(c, id);
  if (!z || s < 0 || s >= z->S) return fail(c, SADMC_E_ARG, "no such replica run");
  ZSim sim;
  HIPCHECK(c, hipMemcpyAsync(&sim, z->d_sim + s, sizeof sim, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (count) *count = sim.R;
  if (out) {
    const int k = std::min(max_out, sim.R);
    if (k > 0)
      HIPCHECK(c, hipMemcpyAsync(out, z->d_rep + (int64_t)s * z->RM, sizeof(sadmc_zreplica) * k,
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  return SADMC_OK;
}

static int zmc_slot(sadmc_ctx *c, ZmcHost *z, int32_t s, int32_t r, int64_t *w) {
  if (s < 0 || s >= z->S || r < 0 || r >= z->RM) return fail(c, SADMC_E_ARG, "run or replica out of range");
  int32_t slot = 0;
  HIPCHECK(c, hipMemcpyAsync(&slot, z->d_slot_of + (int64_t)s * z->RM + r, sizeof slot, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  *w = (int64_t)s * z->RM + slot;
  return SADMC_OK;
}

int sadmc_zmc_coords(sadmc_ctx *c, int32_t id, int32_t s, int32_t r, double *coords) {
  ZmcHost *z = zmc_get(c, id);
  if (!z || !coords) return fail(c, SADMC_E_ARG, "no such replica set or null output");
  if (z->sys == SYS_ISING) return fail(c, SADMC_E_ARG, "Ising replicas have spins, not coordinates");
  int64_t w;
  int rc = zmc_slot(c, z, s, r, &w);
  if (rc) return rc;
  for (int d = 0; d < 3; d++)
    HIPCHECK(c, hipMemcpyAsync(coords + (size_t)d * z->n, z->d_pos + (w * 3 + d) * z->npad, sizeof(double) * z->n,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return SADMC_OK;
}

int sadmc_zmc_spins(sadmc_ctx *c, int32_t id, int32_t s, int32_t r, int8_t *spins) {
  ZmcHost *z = zmc_get(c, id);
  if (!z || !spins) return fail(c, SADMC_E_ARG, "no such replica set or null output");
  if (z->sys != SYS_ISING) return fail(c, SADMC_E_ARG, "only Ising replicas have spins");
  int64_t w;
  int rc = zmc_slot(c, z, s, r, &w);
  if (rc) return rc;
  uint64_t b = 0;
  HIPCHECK(c, hipMemcpyAsync(&b, z->d_spins + w, sizeof b, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < z->n * z->n; k++) spins[k] = ((b >> k) & 1ULL) ? 1 : -1;
  return SADMC_OK;
}

int sadmc_zmc_state(sadmc_ctx *c, int32_t id, int32_t s, uint64_t *rng, uint64_t *moves, double *median,
                    int32_t *len) {
  ZmcHost *z = zmc_get(c, id);
  if (!z || s < 0 || s >= z->S) return fail(c, SADMC_E_ARG, "no such replica run");
  ZSim sim;
  HIPCHECK(c, hipMemcpyAsync(&sim, z->d_sim + s, sizeof sim, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (rng) {
    rng[0] = sim.rng[0];
    rng[1] = sim.rng[1];
  }
  if (moves) *moves = sim.moves;
  if (len) *len = sim.med_len;
  if (median && sim.med_len > 0) {
    HIPCHECK(c, hipMemcpyAsync(median, z->d_med + (int64_t)s * 4096, sizeof(double) * sim.med_len,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  return SADMC_OK;
}

}  // extern "C"
