"""Python host side of the sadmc walker engine (ctypes over include/sadmc.h).

Mirrors the reference's driver-facing API: a `WalkerSet` is a batch of
independent `EnergyMC` walkers (src/mc/energy.rs:167-210) whose
`move_once` (energy.rs:904-974) runs on the GPU, K moves per launch.  The
product path is the compiled HIP library `libsadmc_amd.so` next to this file;
there is no CPU fallback: importing works anywhere, but creating a context
raises unless a gfx950 device is present and the library is built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SADMC_LIB: another build of the same sources (build_lib.py --host-asan)
LIB_PATH = os.environ.get("SADMC_LIB") or os.path.join(_HERE, "libsadmc_amd.so")
_LIB = None

ISING, LJ, WCA = 0, 1, 2
SAD, SAMC, WL, INV_T_WL, CANONICAL = 0, 1, 2, 3, 4
TRANSLATION_SCALE, ACCEPTANCE_RATE = 0, 1
BIN_HISTOGRAM, BIN_LINEAR = 0, 1
# walker status codes (sadmc_walker_state.status, sadmc_bwalker_state.status)
WS_OK, WS_INDEX, WS_VERIFY, WS_INIT, WS_CAPACITY = 0, 1, 2, 3, 4
FP64, FP32_PAIRS = 0, 1
MOVE_CUBE, MOVE_LEARNED = 0, 1

WCA_RC2 = 2.0 ** (1.0 / 3.0)  # (2^(1/6) sigma)^2, wca.rs:61-63

ERRORS = {
    -1: "argument", -2: "HIP", -3: "capacity", -4: "verify", -5: "init", -6: "state", -7: "no device",
    -8: "index",
}


class SadmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sadmc error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class SetParams(C.Structure):
    _fields_ = [
        ("system", C.c_int32), ("n", C.c_int32), ("method", C.c_int32), ("move_plan", C.c_int32),
        ("method_param", C.c_double), ("has_min_gamma", C.c_int32), ("precision", C.c_int32),
        ("move_param", C.c_double), ("energy_bin", C.c_double),
        ("has_min_e", C.c_int32), ("has_max_e", C.c_int32), ("min_e", C.c_double), ("max_e", C.c_double),
        ("rc2", C.c_double), ("refresh", C.c_int32), ("ising_init", C.c_int32),
        ("collect_pressure", C.c_int32), ("lanes", C.c_int32),
        ("e_lo_bound", C.c_double), ("e_hi_bound", C.c_double),
        ("move_kind", C.c_int32), ("binning", C.c_int32), ("high_resolution_de", C.c_double),
        ("reserved", C.c_double * 3),
    ]


WALKER_STATE_DTYPE = np.dtype([
    ("rng", "<u8", 2), ("E", "<f8"), ("error", "<f8"), ("moves", "<u8"), ("accepted_moves", "<u8"),
    ("acceptance_rate", "<f8"), ("translation_scale", "<f8"), ("bins_min", "<f8"), ("bins_width", "<f8"),
    ("base", "<i8"), ("len", "<i8"), ("max_S", "<f8"), ("max_S_index", "<i8"), ("epoch", "<u8"),
    ("bulk_visited", "<u8"), ("method", "<i4"), ("status", "<i4"),
    ("too_lo", "<f8"), ("too_hi", "<f8"), ("latest_parameter", "<f8"), ("min_T", "<f8"),
    ("tL", "<u8"), ("tF", "<u8"), ("num_states", "<u8"), ("highest_hist", "<u8"), ("tmax", "<u8"),
    ("t0", "<f8"), ("wl_gamma", "<f8"), ("wl_num_states", "<f8"), ("wl_min_energy", "<f8"),
    ("wl_lowest_hist", "<u8"), ("wl_highest_hist", "<u8"), ("wl_total_hist", "<u8"), ("wl_len", "<i8"),
    ("has_pressure", "<i4"), ("reserved0", "<i4"), ("geom", "<f8"), ("cap", "<i8"), ("reserved", "<u8", 3),
])

BIN_DTYPE = np.dtype([
    ("lnw", "<f8"), ("energy_total", "<f8"), ("energy_squared_total", "<f8"), ("histogram", "<u8"),
    ("t_found", "<u8"), ("round_trips", "<u8"), ("visit_epoch", "<u8"), ("wl_hist", "<u8"),
])


# sadmc_replica: one replica of a tempering ladder.
REPLICA_DTYPE = np.dtype([
    ("rng", "<u8", 2), ("E", "<f8"), ("error", "<f8"), ("T", "<f8"), ("total_energy", "<f8"),
    ("total_energy_squared", "<f8"), ("accepted", "<u8"), ("rejected", "<u8"), ("ignored", "<u8"),
    ("accepted_swaps", "<u8"), ("rejected_swaps", "<u8"), ("geom", "<f8"), ("status", "<i4"), ("reserved0", "<i4"),
])

# sadmc_zreplica: one replica of an energy-ceiling run.
ZREPLICA_DTYPE = np.dtype([
    ("max_energy", "<f8"), ("cutoff_energy", "<f8"), ("rejected_count", "<u8"), ("accepted_count", "<u8"),
    ("above_count", "<u8"), ("below_count", "<u8"), ("upwelling_count", "<u8"), ("above_total", "<f8"),
    ("below_total", "<f8"), ("above_total_squared", "<f8"), ("below_total_squared", "<f8"),
    ("lowest_max_energy", "<f8"), ("unique_visitors", "<u8"), ("collecting_data", "<i4"), ("reserved", "<i4"),
    ("rng", "<u8", 2), ("translation_scale", "<f8"), ("E", "<f8"), ("error", "<f8"), ("pressure_total", "<f8"),
    ("pressure_count", "<u8"),
])

# sadmc_bcounts / sadmc_bwalker_state / sadmc_bbin: walkers of the binning path.
BCOUNTS_DTYPE = np.dtype([
    ("min_total", "<f8"), ("max_total", "<f8"), ("e_max_total", "<f8"), ("e_max_count", "<f8"),
    ("min_count", "<u8"), ("max_count", "<u8"), ("total_count", "<u8"), ("exists", "<i4"), ("reserved", "<i4"),
])
BWALKER_STATE_DTYPE = np.dtype([
    ("rng", "<u8", 2), ("E", "<f8"), ("error", "<f8"), ("moves", "<u8"), ("accepted_moves", "<u8"),
    ("acceptance_rate", "<f8"), ("translation_scale", "<f8"), ("min", "<f8"), ("width", "<f8"), ("min_e", "<f8"),
    ("max_e", "<f8"), ("len", "<i8"), ("method", "<i4"), ("status", "<i4"), ("num_states", "<u8"), ("tL", "<u8"),
    ("min_T", "<f8"), ("too_lo", "<f8"), ("too_hi", "<f8"), ("tF", "<f8"), ("latest_parameter", "<f8"),
    ("t0", "<f8"), ("gamma", "<f8"),
    ("lnw", BCOUNTS_DTYPE), ("energy", BCOUNTS_DTYPE), ("t_found", BCOUNTS_DTYPE), ("hist", BCOUNTS_DTYPE),
    ("pressure", BCOUNTS_DTYPE),
    ("base", "<i8"), ("cap", "<i8"), ("geom", "<f8"), ("inv_t", "<i4"), ("has_min_gamma", "<i4"),
    ("min_gamma", "<f8"), ("reserved", "<u8", 3),
])
BBIN_DTYPE = np.dtype([
    ("lnw_total", "<f8"), ("lnw_count", "<u8"), ("energy_total", "<f8"), ("energy_count", "<u8"),
    ("t_found_total", "<f8"), ("t_found_count", "<u8"), ("hist_count", "<u8"), ("pressure_total", "<f8"),
    ("pressure_count", "<u8"), ("reserved", "<u8"),
])
# sadmc_hires: the high_resolution side histogram's statistics.
HIRES_DTYPE = np.dtype([
    ("min", "<f8"), ("width", "<f8"), ("min_e", "<f8"), ("max_e", "<f8"), ("len", "<i8"), ("base", "<i8"),
    ("total_count", "<u8"), ("min_count", "<u8"), ("max_count", "<u8"), ("e_max_count", "<f8"),
])


class SystemInfo(C.Structure):
    """sadmc_system_info_t: the constant System trait queries."""
    _fields_ = [
        ("has_lowest_possible_energy", C.c_int32), ("has_delta_energy", C.c_int32),
        ("lowest_possible_energy", C.c_double), ("delta_energy", C.c_double),
        ("min_moves_to_randomize", C.c_uint64), ("dimensionality", C.c_uint64), ("max_size", C.c_double),
    ]


class WalkerStateC(C.Structure):
    _fields_ = [(name, C.c_uint8 * WALKER_STATE_DTYPE.itemsize) for name in ("raw",)]


# Every entry point the header declares (checked against the library by the
# CPU test suite).
EXPORTS = [
    "sadmc_abi_version", "sadmc_create", "sadmc_destroy", "sadmc_last_error", "sadmc_add_set", "sadmc_init",
    "sadmc_run", "sadmc_run_set", "sadmc_sync", "sadmc_last_run_ms",
    "sadmc_compute_energy", "sadmc_verify_energy", "sadmc_system_info", "sadmc_upload_pressure", "sadmc_plan_move", "sadmc_confirm", "sadmc_walker_count",
    "sadmc_check_recip_f32",
    "sadmc_get_walker_state", "sadmc_walker_bins", "sadmc_walker_coords", "sadmc_walker_spins",
    "sadmc_download_raw", "sadmc_upload_raw", "sadmc_init_coords", "sadmc_init_spins", "sadmc_bin_window",
    "sadmc_run_batched", "sadmc_graph_run", "sadmc_autotune_lanes", "sadmc_set_block_waves",
    "sadmc_autotune_batch",
    "sadmc_plan_learned", "sadmc_prop_init", "sadmc_prop_set_weights", "sadmc_prop_get_weights",
    "sadmc_prop_collect", "sadmc_prop_records", "sadmc_prop_load_records", "sadmc_prop_gradient",
    "sadmc_prop_train", "sadmc_mean_entropy", "sadmc_mean_entropy_noalloc",
    "sadmc_pt_add", "sadmc_pt_run", "sadmc_pt_replicas", "sadmc_pt_coords", "sadmc_pt_spins", "sadmc_pt_ladder_rng",
    "sadmc_bin_add_set", "sadmc_bin_run", "sadmc_bin_state", "sadmc_bin_bins", "sadmc_bin_coords", "sadmc_bin_spins",
    "sadmc_bin_hires",
    "sadmc_zmc_add", "sadmc_zmc_run", "sadmc_zmc_replicas", "sadmc_zmc_coords", "sadmc_zmc_spins", "sadmc_zmc_state",
]

PROP_NP, PROP_REC, PROP_NS = 1377, 24, 16
PROP_SMIN, PROP_SMAX = -4.605170186, 0.0


def lib():
    """Load libsadmc_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    L.sadmc_abi_version.restype = C.c_int
    L.sadmc_create.argtypes = [C.POINTER(vp), C.c_int]
    L.sadmc_destroy.argtypes = [vp]
    L.sadmc_last_error.argtypes = [vp]; L.sadmc_last_error.restype = C.c_char_p
    L.sadmc_add_set.argtypes = [vp, C.POINTER(SetParams), i32, vp, vp, vp, vp, C.POINTER(i32)]
    L.sadmc_init.argtypes = [vp]
    L.sadmc_run.argtypes = [vp, u64]
    L.sadmc_sync.argtypes = [vp]
    L.sadmc_run_set.argtypes = [vp, i32, u64]
    L.sadmc_compute_energy.argtypes = [vp, i32, i32, C.POINTER(d)]
    L.sadmc_verify_energy.argtypes = [vp, i32, i32, C.POINTER(d), C.POINTER(i32)]
    L.sadmc_system_info.argtypes = [C.POINTER(SetParams), d, C.POINTER(SystemInfo)]
    L.sadmc_check_recip_f32.argtypes = [C.c_int, i32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
    L.sadmc_upload_pressure.argtypes = [vp, i32, vp, vp]
    L.sadmc_last_run_ms.argtypes = [vp, C.POINTER(d)]
    L.sadmc_plan_move.argtypes = [vp, i32, i32, d, C.POINTER(d), C.POINTER(i32)]
    L.sadmc_confirm.argtypes = [vp, i32, i32]
    L.sadmc_walker_count.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(i64)]
    L.sadmc_get_walker_state.argtypes = [vp, i32, i32, vp]
    L.sadmc_walker_bins.argtypes = [vp, i32, i32, i64] + [vp] * 10
    L.sadmc_walker_coords.argtypes = [vp, i32, i32, vp]
    L.sadmc_walker_spins.argtypes = [vp, i32, i32, vp]
    L.sadmc_download_raw.argtypes = [vp, i32, vp, vp]
    L.sadmc_upload_raw.argtypes = [vp, i32, vp, vp, vp, vp]
    L.sadmc_init_coords.argtypes = [i32, i32, d, u64, vp]
    L.sadmc_init_spins.argtypes = [i32, i32, u64, vp]
    L.sadmc_bin_window.argtypes = [C.POINTER(SetParams), d, C.POINTER(d), C.POINTER(d), C.POINTER(i64)]
    L.sadmc_run_batched.argtypes = [vp, u64]
    L.sadmc_graph_run.argtypes = [vp, u64]
    L.sadmc_autotune_lanes.argtypes = [C.c_int, C.POINTER(SetParams), i32, vp, vp, u64, u64, C.POINTER(i32),
                                       C.POINTER(d)]
    L.sadmc_set_block_waves.argtypes = [vp, i32, i32]
    L.sadmc_autotune_batch.argtypes = [C.c_int, i32, vp, vp, vp, vp, u64, u64, vp, C.POINTER(i32), C.POINTER(d)]
    L.sadmc_plan_learned.argtypes = [vp, i32, i32, C.POINTER(d), C.POINTER(d), vp, C.POINTER(i32)]
    L.sadmc_prop_init.argtypes = [vp, u64]
    L.sadmc_prop_set_weights.argtypes = [vp, vp, i32]
    L.sadmc_prop_get_weights.argtypes = [vp, vp, vp]
    L.sadmc_prop_collect.argtypes = [vp, i32, u64, i32]
    L.sadmc_prop_records.argtypes = [vp, vp, i64, C.POINTER(i64)]
    L.sadmc_prop_load_records.argtypes = [vp, vp, i64]
    L.sadmc_prop_gradient.argtypes = [vp, i64, i32, vp, C.POINTER(C.c_float)]
    L.sadmc_prop_train.argtypes = [vp, i32, i32, C.c_float, vp]
    L.sadmc_mean_entropy.argtypes = [vp, i32, d, d, i64, C.POINTER(d), C.POINTER(i64), vp, vp]
    L.sadmc_pt_add.argtypes = [vp, C.POINTER(SetParams), i32, i32, vp, vp, vp, vp, vp, i32, C.POINTER(i32)]
    L.sadmc_pt_run.argtypes = [vp, i32, u64]
    L.sadmc_pt_replicas.argtypes = [vp, i32, vp]
    L.sadmc_pt_coords.argtypes = [vp, i32, i32, i32, vp]
    L.sadmc_pt_spins.argtypes = [vp, i32, i32, i32, vp]
    L.sadmc_pt_ladder_rng.argtypes = [vp, i32, vp]
    L.sadmc_bin_add_set.argtypes = [vp, C.POINTER(SetParams), i32, vp, vp, vp, vp, C.POINTER(i32)]
    L.sadmc_bin_run.argtypes = [vp, i32, u64]
    L.sadmc_bin_state.argtypes = [vp, i32, i32, vp]
    L.sadmc_bin_bins.argtypes = [vp, i32, i32, i64, vp]
    L.sadmc_bin_coords.argtypes = [vp, i32, i32, vp]
    L.sadmc_bin_spins.argtypes = [vp, i32, i32, vp]
    L.sadmc_bin_hires.argtypes = [vp, i32, i32, vp, i64, vp]
    L.sadmc_zmc_add.argtypes = [vp, C.POINTER(SetParams), i32, vp, vp, vp, vp, d, i64, i32, C.POINTER(i32)]
    L.sadmc_zmc_run.argtypes = [vp, i32, u64]
    L.sadmc_zmc_replicas.argtypes = [vp, i32, i32, i32, vp, C.POINTER(i32)]
    L.sadmc_zmc_coords.argtypes = [vp, i32, i32, i32, vp]
    L.sadmc_zmc_spins.argtypes = [vp, i32, i32, i32, vp]
    L.sadmc_zmc_state.argtypes = [vp, i32, i32, vp, C.POINTER(u64), vp, C.POINTER(i32)]
    for name in EXPORTS:
        if name not in ("sadmc_last_error", "sadmc_destroy", "sadmc_abi_version"):
            getattr(L, name).restype = C.c_int
    _LIB = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_params(**kw):
    p = SetParams()
    defaults = dict(rc2=WCA_RC2, move_param=0.05)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def init_coords(system, n, geom, seed):
    c = np.zeros(3 * n)
    rc = lib().sadmc_init_coords(system, n, geom, seed, _p(c))
    if rc:
        raise SadmcError(rc, "init_coords")
    return c


def init_spins(n, rule, seed):
    s = np.zeros(n * n, np.int8)
    rc = lib().sadmc_init_spins(n, rule, seed, _p(s))
    if rc:
        raise SadmcError(rc, "init_spins")
    return s


def autotune_lanes(params, seeds, geom, warm_moves=500, trial_moves=500, device=0):
    """Fastest lanes-per-walker for a walker set (see sadmc_autotune_lanes)."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    g = np.ascontiguousarray(np.broadcast_to(geom, seeds.shape), dtype=np.float64)
    best = C.c_int32()
    rates = (C.c_double * 7)()
    rc = lib().sadmc_autotune_lanes(device, C.byref(params), len(seeds), _p(seeds), _p(g), warm_moves,
                                    trial_moves, C.byref(best), rates)
    if rc:
        raise SadmcError(rc, "autotune_lanes")
    return best.value, {1 << k: rates[k] for k in range(7) if rates[k] > 0}


def bin_window(params, geom=0.0):
    lo, hi, cap = C.c_double(), C.c_double(), C.c_int64()
    lib().sadmc_bin_window(C.byref(params), geom, C.byref(lo), C.byref(hi), C.byref(cap))
    return lo.value, hi.value, cap.value


class Context:
    """One device context (one HIP stream) holding walker sets."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.sadmc_create(C.byref(h), device)
        if rc:
            raise SadmcError(rc, "sadmc_create: no gfx950 device or HIP failure")
        self.h = h
        self.sets = []

    def _check(self, rc, what=""):
        if rc:
            raise SadmcError(rc, f"{what}: {lib().sadmc_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            lib().sadmc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_set(self, params, seeds, geom=None, coords=None, spins=None):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        g = None if geom is None else np.ascontiguousarray(np.broadcast_to(geom, seeds.shape), dtype=np.float64)
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        s = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        sid = C.c_int32()
        self._check(lib().sadmc_add_set(self.h, C.byref(params), len(seeds), _p(seeds), _p(g), _p(c), _p(s),
                                        C.byref(sid)), "add_set")
        self.sets.append((params, len(seeds)))
        return sid.value

    def init(self):
        self._check(lib().sadmc_init(self.h), "init")

    def run(self, moves, sync=True):
        self._check(lib().sadmc_run(self.h, moves), "run")
        if sync:
            self.sync()

    def run_set(self, s, moves, sync=True):
        """Advance only the walkers of set s by `moves`."""
        self._check(lib().sadmc_run_set(self.h, s, moves), "run_set")
        if sync:
            self.sync()

    def compute_energy(self, s, w):
        """System::compute_energy of walker w, recomputed on the device."""
        e = C.c_double()
        self._check(lib().sadmc_compute_energy(self.h, s, w, C.byref(e)), "compute_energy")
        return e.value

    def verify_energy(self, s, w):
        """System::verify_energy: (ok, recomputed energy); a failed check
        raises SadmcError (code -4), as the reference panics."""
        e, ok = C.c_double(), C.c_int32()
        self._check(lib().sadmc_verify_energy(self.h, s, w, C.byref(e), C.byref(ok)), "verify_energy")
        return bool(ok.value), e.value

    def upload_pressure(self, s, count, total):
        """The WCA pressure extra of every walker, [W][cap] physical bins."""
        self._check(lib().sadmc_upload_pressure(self.h, s, _p(np.ascontiguousarray(count, np.uint64)),
                                                _p(np.ascontiguousarray(total, np.float64))), "upload_pressure")

    def set_block_waves(self, s, waves):
        """Wavefronts per workgroup (1, 2, 4) of set s's run launch; s = -1:
        of the batched launch."""
        self._check(lib().sadmc_set_block_waves(self.h, s, waves), "set_block_waves")

    def run_batched(self, moves, sync=True, graph=False):
        fn = lib().sadmc_graph_run if graph else lib().sadmc_run_batched
        self._check(fn(self.h, moves), "graph_run" if graph else "run_batched")
        if sync:
            self.sync()

    def mean_entropy(self, s, window=None):
        """(E, S, count) of a SAD set reduced on the device (bins no walker covers dropped)."""
        ea, eb = window if window is not None else (float("nan"), float("nan"))
        e0, nb = C.c_double(), C.c_int64()
        # every walker's bins lie in the set's window of `cap` slots, so the
        # common grid has at most `cap` bins: one call normally suffices
        n, cap = C.c_int32(), C.c_int64()
        self._check(lib().sadmc_walker_count(self.h, s, C.byref(n), C.byref(cap)), "walker_count")
        S = np.zeros(max(cap.value, 1))
        cnt = np.zeros(max(cap.value, 1), np.int64)
        rc = lib().sadmc_mean_entropy(self.h, s, ea, eb, len(S), C.byref(e0), C.byref(nb), _p(S), _p(cnt))
        if rc == -3:
            S = np.zeros(nb.value)
            cnt = np.zeros(nb.value, np.int64)
            rc = lib().sadmc_mean_entropy(self.h, s, ea, eb, nb.value, C.byref(e0), C.byref(nb), _p(S), _p(cnt))
        self._check(rc, "mean_entropy")
        S, cnt = S[:nb.value], cnt[:nb.value]
        w = self.sets[s][0].energy_bin if self.sets[s][0].energy_bin > 0 else (4.0 if self.sets[s][0].system == ISING
                                                                               else 1.0)
        E = e0.value + np.arange(nb.value) * w
        keep = cnt > 0
        return E[keep], S[keep], cnt[keep]

    # ---- walkers of the binning path (energy_binning.rs)
    def bin_add_set(self, params, seeds, geom=None, coords=None, spins=None):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        g = None if geom is None else np.ascontiguousarray(np.broadcast_to(geom, seeds.shape), dtype=np.float64)
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        sid = C.c_int32()
        self._check(lib().sadmc_bin_add_set(self.h, C.byref(params), len(seeds), _p(seeds), _p(g), _p(c), _p(sp),
                                            C.byref(sid)), "bin_add_set")
        if not hasattr(self, "bsets"):
            self.bsets = []
        self.bsets.append((params, len(seeds)))
        return sid.value

    def bin_run(self, s, moves):
        self._check(lib().sadmc_bin_run(self.h, s, moves), "bin_run")

    def bin_state(self, s, w):
        st = np.zeros((), BWALKER_STATE_DTYPE)
        self._check(lib().sadmc_bin_state(self.h, s, w, _p(st)), "bin_state")
        return st

    def bin_bins(self, s, w):
        n = int(self.bin_state(s, w)["len"])
        out = np.zeros(max(n, 1), BBIN_DTYPE)
        self._check(lib().sadmc_bin_bins(self.h, s, w, len(out), _p(out)), "bin_bins")
        return out[:n]

    def bin_coords(self, s, w):
        c = np.zeros(3 * self.bsets[s][0].n)
        self._check(lib().sadmc_bin_coords(self.h, s, w, _p(c)), "bin_coords")
        return c

    def bin_spins(self, s, w):
        n = self.bsets[s][0].n
        a = np.zeros(n * n, np.int8)
        self._check(lib().sadmc_bin_spins(self.h, s, w, _p(a)), "bin_spins")
        return a

    def bin_hires(self, s, w):
        """The high_resolution side histogram of walker w: (statistics, counts)."""
        st = np.zeros((), HIRES_DTYPE)
        self._check(lib().sadmc_bin_hires(self.h, s, w, _p(st), 0, None), "bin_hires")
        n = int(st["len"])
        counts = np.zeros(max(n, 1), np.uint64)
        self._check(lib().sadmc_bin_hires(self.h, s, w, _p(st), len(counts), _p(counts)), "bin_hires")
        return st, counts[:n]

    # ---- energy-ceiling replicas (energy_replicas.rs)
    def zmc_add(self, params, seeds, geom=None, coords=None, spins=None, min_T=0.2, isbnb=64, max_replicas=32):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        g = None if geom is None else np.ascontiguousarray(np.broadcast_to(geom, seeds.shape), dtype=np.float64)
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        zid = C.c_int32()
        self._check(lib().sadmc_zmc_add(self.h, C.byref(params), len(seeds), _p(seeds), _p(g), _p(c), _p(sp), min_T,
                                        isbnb, max_replicas, C.byref(zid)), "zmc_add")
        if not hasattr(self, "zmcs"):
            self.zmcs = []
        self.zmcs.append((params, len(seeds), max_replicas))
        return zid.value

    def zmc_run(self, z, rounds):
        self._check(lib().sadmc_zmc_run(self.h, z, rounds), "zmc_run")

    def zmc_replicas(self, z, s):
        rm = self.zmcs[z][2]
        out = np.zeros(rm, ZREPLICA_DTYPE)
        cnt = C.c_int32()
        self._check(lib().sadmc_zmc_replicas(self.h, z, s, rm, _p(out), C.byref(cnt)), "zmc_replicas")
        return out[:cnt.value]

    def zmc_coords(self, z, s, r):
        c = np.zeros(3 * self.zmcs[z][0].n)
        self._check(lib().sadmc_zmc_coords(self.h, z, s, r, _p(c)), "zmc_coords")
        return c

    def zmc_spins(self, z, s, r):
        n = self.zmcs[z][0].n
        a = np.zeros(n * n, np.int8)
        self._check(lib().sadmc_zmc_spins(self.h, z, s, r, _p(a)), "zmc_spins")
        return a

    def zmc_state(self, z, s):
        """(generator, moves, median estimator energies) of run s."""
        rng = np.zeros(2, np.uint64)
        mv = C.c_uint64()
        med = np.zeros(4096)
        ln = C.c_int32()
        self._check(lib().sadmc_zmc_state(self.h, z, s, _p(rng), C.byref(mv), _p(med), C.byref(ln)), "zmc_state")
        return rng, mv.value, med[:ln.value]

    # ---- parallel tempering
    def pt_add(self, params, temps, seeds, geom=None, coords=None, spins=None, canonical_steps=1):
        """Ladders (one per seed) of len(temps) replicas; returns the ladder-set id."""
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        T = np.ascontiguousarray(temps, dtype=np.float64)
        g = None if geom is None else np.ascontiguousarray(np.broadcast_to(geom, seeds.shape), dtype=np.float64)
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        pid = C.c_int32()
        self._check(lib().sadmc_pt_add(self.h, C.byref(params), len(seeds), len(T), _p(T), _p(seeds), _p(g), _p(c),
                                       _p(sp), canonical_steps, C.byref(pid)), "pt_add")
        if not hasattr(self, "pts"):
            self.pts = []
        self.pts.append((params, len(seeds), len(T)))
        return pid.value

    def pt_run(self, pt, rounds, sync=True):
        self._check(lib().sadmc_pt_run(self.h, pt, rounds), "pt_run")
        if sync:
            self.sync()

    def pt_replicas(self, pt):
        """Replica records, shape (ladders, replicas)."""
        _, L, R = self.pts[pt]
        out = np.zeros(L * R, REPLICA_DTYPE)
        self._check(lib().sadmc_pt_replicas(self.h, pt, _p(out)), "pt_replicas")
        return out.reshape(L, R)

    def pt_coords(self, pt, ladder, r):
        c = np.zeros(3 * self.pts[pt][0].n)
        self._check(lib().sadmc_pt_coords(self.h, pt, ladder, r, _p(c)), "pt_coords")
        return c

    def pt_spins(self, pt, ladder, r):
        n = self.pts[pt][0].n
        a = np.zeros(n * n, np.int8)
        self._check(lib().sadmc_pt_spins(self.h, pt, ladder, r, _p(a)), "pt_spins")
        return a

    def pt_ladder_rng(self, pt):
        out = np.zeros(2 * self.pts[pt][1], np.uint64)
        self._check(lib().sadmc_pt_ladder_rng(self.h, pt, _p(out)), "pt_ladder_rng")
        return out.reshape(-1, 2)

    # ---- learned proposal
    def plan_learned(self, s, w):
        """(e_new, ln H, new position) of walker w's learned move, or None if the wall rejects it."""
        e, h, some = C.c_double(), C.c_double(), C.c_int32()
        to = np.zeros(3)
        self._check(lib().sadmc_plan_learned(self.h, s, w, C.byref(e), C.byref(h), _p(to), C.byref(some)),
                    "plan_learned")
        return (e.value, h.value, to) if some.value else None

    def prop_init(self, seed=0):
        self._check(lib().sadmc_prop_init(self.h, seed), "prop_init")

    def prop_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(lib().sadmc_prop_set_weights(self.h, _p(w), len(w)), "prop_set_weights")

    def prop_get_weights(self):
        """(FP32 master weights, the BF16 copy as float32)."""
        m = np.zeros(PROP_NP, np.float32)
        b = np.zeros(PROP_NP, np.uint16)
        self._check(lib().sadmc_prop_get_weights(self.h, _p(m), _p(b)), "prop_get_weights")
        return m, (b.astype(np.uint32) << 16).view(np.float32)

    def prop_collect(self, s, seed, per_walker):
        self._check(lib().sadmc_prop_collect(self.h, s, seed, per_walker), "prop_collect")

    def prop_records(self):
        n = C.c_int64()
        self._check(lib().sadmc_prop_records(self.h, None, 0, C.byref(n)), "prop_records")
        out = np.zeros((n.value, PROP_REC), np.float32)
        self._check(lib().sadmc_prop_records(self.h, _p(out), n.value, C.byref(n)), "prop_records")
        return out

    def prop_load_records(self, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float32).reshape(-1, PROP_REC)
        self._check(lib().sadmc_prop_load_records(self.h, _p(rec), len(rec)), "prop_load_records")

    def prop_gradient(self, first, count):
        g = np.zeros(PROP_NP, np.float32)
        loss = C.c_float()
        self._check(lib().sadmc_prop_gradient(self.h, first, count, _p(g), C.byref(loss)), "prop_gradient")
        return loss.value, g

    def prop_train(self, steps, batch, lr):
        losses = np.zeros(steps, np.float32)
        self._check(lib().sadmc_prop_train(self.h, steps, batch, lr, _p(losses)), "prop_train")
        return losses

    def sync(self):
        self._check(lib().sadmc_sync(self.h), "sync")

    def last_run_ms(self):
        ms = C.c_double()
        self._check(lib().sadmc_last_run_ms(self.h, C.byref(ms)), "last_run_ms")
        return ms.value

    def plan_move(self, s, w, scale):
        e, some = C.c_double(), C.c_int32()
        self._check(lib().sadmc_plan_move(self.h, s, w, scale, C.byref(e), C.byref(some)), "plan_move")
        return e.value if some.value else None

    def confirm(self, s, w):
        self._check(lib().sadmc_confirm(self.h, s, w), "confirm")

    def count(self, s):
        n, cap = C.c_int32(), C.c_int64()
        self._check(lib().sadmc_walker_count(self.h, s, C.byref(n), C.byref(cap)), "count")
        return n.value, cap.value

    def state(self, s, w):
        a = np.zeros(1, WALKER_STATE_DTYPE)
        self._check(lib().sadmc_get_walker_state(self.h, s, w, _p(a)), "state")
        return a[0]

    def bins(self, s, w):
        n = int(self.state(s, w)["len"])
        out = dict(
            lnw=np.zeros(n), histogram=np.zeros(n, np.uint64), t_found=np.zeros(n, np.uint64),
            energy_total=np.zeros(n), energy_squared_total=np.zeros(n), visited=np.zeros(n, np.uint8),
            round_trips=np.zeros(n, np.uint64), wl_hist=np.zeros(n, np.uint64),
            pressure_count=np.zeros(n, np.uint64), pressure_total=np.zeros(n),
        )
        self._check(lib().sadmc_walker_bins(self.h, s, w, n, *[_p(out[k]) for k in (
            "lnw", "histogram", "t_found", "energy_total", "energy_squared_total", "visited", "round_trips",
            "wl_hist", "pressure_count", "pressure_total")]), "bins")
        return out

    def coords(self, s, w):
        n = self.sets[s][0].n
        c = np.zeros(3 * n)
        self._check(lib().sadmc_walker_coords(self.h, s, w, _p(c)), "coords")
        return c

    def spins(self, s, w):
        n = self.sets[s][0].n
        a = np.zeros(n * n, np.int8)
        self._check(lib().sadmc_walker_spins(self.h, s, w, _p(a)), "spins")
        return a

    def download_raw(self, s, with_bins=True):
        W, cap = self.count(s)
        st = np.zeros(W, WALKER_STATE_DTYPE)
        b = np.zeros(W * cap, BIN_DTYPE) if with_bins else None
        self._check(lib().sadmc_download_raw(self.h, s, _p(st), _p(b)), "download_raw")
        return st, (b.reshape(W, cap) if with_bins else None)

    def upload_raw(self, s, states, bins, coords=None, spins=None):
        self._check(lib().sadmc_upload_raw(self.h, s, _p(np.ascontiguousarray(states)),
                                           _p(np.ascontiguousarray(bins)),
                                           _p(None if coords is None else np.ascontiguousarray(coords, np.float64)),
                                           _p(None if spins is None else np.ascontiguousarray(spins, np.int8))),
                    "upload_raw")


def autotune_batch(sets, warm_moves=20000, trial_moves=2000, device=0):
    """Joint tuning of a mixed batch: `sets` is a list of (params, seeds,
    geom); returns (lanes per set, wavefronts per workgroup, moves/s)."""
    n = len(sets)
    ps = (SetParams * n)(*[p for p, _, _ in sets])
    nw = np.array([len(sd) for _, sd, _ in sets], np.int32)
    seeds = np.ascontiguousarray(np.concatenate([np.asarray(sd, np.uint64) for _, sd, _ in sets]))
    geom = np.ascontiguousarray(np.concatenate([np.broadcast_to(np.asarray(g, np.float64), (len(sd),))
                                                for _, sd, g in sets]))
    lanes = np.zeros(n, np.int32)
    waves, rate = C.c_int32(), C.c_double()
    rc = lib().sadmc_autotune_batch(device, n, C.cast(ps, C.c_void_p), _p(nw), _p(seeds), _p(geom), warm_moves,
                                    trial_moves, _p(lanes), C.byref(waves), C.byref(rate))
    if rc:
        raise SadmcError(rc, "autotune_batch")
    return [int(x) for x in lanes], waves.value, rate.value


def check_recip_f32(lo, hi, what=0, device=0):
    """what = 0: how many binary32 patterns in [lo, hi) get a different 1/x
    from the FP32 pair terms' reciprocal than from the IEEE division (0 over
    [2^-126, 2^126)); what = 1: how many get a fast pair term that is neither
    NaN nor equal to the division's (0 over all non-negative patterns)."""
    out = C.c_uint64()
    rc = lib().sadmc_check_recip_f32(device, what, C.c_uint64(lo), C.c_uint64(hi), C.byref(out))
    if rc:
        raise SadmcError(rc, "check_recip_f32")
    return int(out.value)


def system_info(params, geom=0.0):
    """The System trait's constant queries for a parameter set: a dict with
    lowest_possible_energy / delta_energy (None where the reference returns
    None), min_moves_to_randomize, dimensionality and max_size."""
    out = SystemInfo()
    rc = lib().sadmc_system_info(C.byref(params), geom, C.byref(out))
    if rc:
        raise SadmcError(rc, "system_info")
    return dict(lowest_possible_energy=out.lowest_possible_energy if out.has_lowest_possible_energy else None,
                delta_energy=out.delta_energy if out.has_delta_energy else None,
                min_moves_to_randomize=out.min_moves_to_randomize, dimensionality=out.dimensionality,
                max_size=out.max_size)


def logical_bins(state, bins_row):
    """Slice one walker's window to its logical bins [0, len)."""
    b, n = int(state["base"]), int(state["len"])
    return bins_row[b:b + n]
