"""ctypes wrapper of the C oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker.  See sadmc_oracle.h for what each
function restates from the reference.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ISING, LJ, WCA, SCRIPTED = 0, 1, 2, 3
SAD, SAMC, WL, INV_T_WL, CANONICAL = 0, 1, 2, 3, 4


class Params(C.Structure):
    _fields_ = [
        ("system", C.c_int32), ("n", C.c_int32), ("method", C.c_int32), ("move_plan", C.c_int32),
        ("method_param", C.c_double), ("has_min_gamma", C.c_int32), ("precision", C.c_int32),
        ("move_param", C.c_double), ("energy_bin", C.c_double),
        ("has_min_e", C.c_int32), ("has_max_e", C.c_int32), ("min_e", C.c_double), ("max_e", C.c_double),
        ("radius", C.c_double), ("box", C.c_double), ("rc2", C.c_double),
        ("refresh", C.c_int32), ("order_g", C.c_int32), ("ising_init", C.c_int32),
        ("collect_pressure", C.c_int32), ("seed", C.c_uint64),
        ("move_kind", C.c_int32), ("drop_hastings", C.c_int32), ("mlp", C.c_void_p),
        ("binning", C.c_int32), ("high_resolution_de", C.c_double),
    ]


class Replica(C.Structure):
    _fields_ = [("E", C.c_double), ("error", C.c_double), ("T", C.c_double), ("total_energy", C.c_double),
                ("total_energy_squared", C.c_double), ("accepted", C.c_uint64), ("rejected", C.c_uint64),
                ("ignored", C.c_uint64), ("accepted_swaps", C.c_uint64), ("rejected_swaps", C.c_uint64),
                ("rng", C.c_uint64 * 2)]


class Scalars(C.Structure):
    _fields_ = [
        ("E", C.c_double), ("error", C.c_double), ("moves", C.c_uint64), ("accepted", C.c_uint64),
        ("acceptance_rate", C.c_double), ("translation_scale", C.c_double), ("rng", C.c_uint64 * 2),
        ("bins_min", C.c_double), ("bins_width", C.c_double), ("len", C.c_int64),
        ("max_S", C.c_double), ("max_S_index", C.c_int64), ("method", C.c_int32), ("status", C.c_int32),
        ("too_lo", C.c_double), ("too_hi", C.c_double), ("lambda_", C.c_double), ("min_T", C.c_double),
        ("tL", C.c_uint64), ("tF", C.c_uint64), ("num_states", C.c_uint64), ("highest_hist", C.c_uint64),
        ("t0", C.c_double),
        ("wl_gamma", C.c_double), ("wl_num_states", C.c_double), ("wl_min_energy", C.c_double),
        ("wl_lowest", C.c_uint64), ("wl_highest", C.c_uint64), ("wl_total", C.c_uint64),
        ("wl_len", C.c_int64), ("has_pressure", C.c_int32),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "rng" else v
        return d


def build():
    """Compile liboracle.so from the C sources (make)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    # SADMC_ORACLE_LIB: another build of the same sources (the sanitizer
    # build of the Makefile's `asan` target)
    path = os.environ.get("SADMC_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    u64p = C.POINTER(C.c_uint64)
    dp = C.POINTER(C.c_double)
    L.or_splitmix64_next.argtypes = [u64p]; L.or_splitmix64_next.restype = C.c_uint64
    L.or_rng_seed.argtypes = [C.c_uint64, u64p]
    L.or_rng_next.argtypes = [u64p]; L.or_rng_next.restype = C.c_uint64
    L.or_gen_range.argtypes = [u64p, C.c_uint64, C.c_uint64]; L.or_gen_range.restype = C.c_uint64
    L.or_uniform_index.argtypes = [u64p, C.c_uint64]; L.or_uniform_index.restype = C.c_uint64
    L.or_gen_f64.argtypes = [u64p]; L.or_gen_f64.restype = C.c_double
    L.or_uniform_f64.argtypes = [u64p, C.c_double, C.c_double]; L.or_uniform_f64.restype = C.c_double
    L.or_lj_pair.argtypes = [C.c_double]; L.or_lj_pair.restype = C.c_double
    L.or_lj_pair_f32.argtypes = [C.c_float]; L.or_lj_pair_f32.restype = C.c_float
    L.or_wca_pair.argtypes = [C.c_double, C.c_double]; L.or_wca_pair.restype = C.c_double
    L.or_mc_new.argtypes = [C.POINTER(Params), C.c_void_p, C.c_void_p]; L.or_mc_new.restype = C.c_void_p
    L.or_mc_new_scripted.argtypes = [C.POINTER(Params), C.c_double, dp, C.c_int64]
    L.or_mc_new_scripted.restype = C.c_void_p
    L.or_mc_free.argtypes = [C.c_void_p]
    L.or_mc_run.argtypes = [C.c_void_p, C.c_uint64]; L.or_mc_run.restype = C.c_int
    L.or_mc_scalars.argtypes = [C.c_void_p, C.POINTER(Scalars)]
    L.or_mc_bins.argtypes = [C.c_void_p] + [C.c_void_p] * 10
    L.or_mc_coords.argtypes = [C.c_void_p, C.c_void_p]
    L.or_mc_spins.argtypes = [C.c_void_p, C.c_void_p]
    L.or_mc_compute_energy.argtypes = [C.c_void_p]; L.or_mc_compute_energy.restype = C.c_double
    L.or_mc_reference_energy.argtypes = [C.c_void_p]; L.or_mc_reference_energy.restype = C.c_double
    L.or_mc_plan_move.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]; L.or_mc_plan_move.restype = C.c_int
    L.or_mc_confirm.argtypes = [C.c_void_p]
    L.or_mc_energy.argtypes = [C.c_void_p]; L.or_mc_energy.restype = C.c_double
    L.or_mc_trace.argtypes = [C.c_void_p, C.c_uint64]
    L.or_mc_trace_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]; L.or_mc_trace_copy.restype = C.c_uint64
    L.or_mc_last_move.argtypes = [C.c_void_p, C.c_void_p]; L.or_mc_last_move.restype = None
    L.or_mc_load_state.argtypes = [C.c_void_p, C.POINTER(Scalars)] + [C.c_void_p] * 8
    L.or_mc_load_state.restype = C.c_int
    L.or_ising_init_spins.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.c_void_p]
    L.or_ising_energy.argtypes = [C.c_int32, C.c_void_p]; L.or_ising_energy.restype = C.c_double
    L.or_lj_init.argtypes = [C.c_int32, C.c_double, C.c_uint64, C.c_void_p]; L.or_lj_init.restype = C.c_int
    L.or_wca_init.argtypes = [C.c_int32, C.c_double, C.c_uint64, C.c_void_p]; L.or_wca_init.restype = C.c_int
    L.or_lj_energy.argtypes = [C.c_int32, C.c_void_p, C.c_int32]; L.or_lj_energy.restype = C.c_double
    L.or_lj_energy_ld.argtypes = [C.c_int32, C.c_void_p]; L.or_lj_energy_ld.restype = C.c_double
    L.or_wca_energy_celllist.argtypes = [C.c_int32, C.c_void_p, C.c_double, C.c_double]
    L.or_wca_energy_celllist.restype = C.c_double
    L.or_wca_energy.argtypes = [C.c_int32, C.c_void_p, C.c_double, C.c_double, C.c_int32]
    L.or_wca_energy.restype = C.c_double
    L.or_wca_energy_ld.argtypes = [C.c_int32, C.c_void_p, C.c_double, C.c_double]
    L.or_wca_energy_ld.restype = C.c_double
    L.or_sad_gamma.argtypes = [C.c_double] * 4; L.or_sad_gamma.restype = C.c_double
    L.or_bench.argtypes = [C.POINTER(Params), C.c_int32, C.c_uint64, C.c_uint64]; L.or_bench.restype = C.c_double
    vp = C.c_void_p
    L.or_rng_jump.argtypes = [u64p]
    L.or_pt_new.argtypes = [C.POINTER(Params), C.c_int32, vp, C.c_uint64, vp, vp, C.c_int32]
    L.or_pt_new.restype = vp
    L.or_pt_run.argtypes = [vp, C.c_uint64]
    L.or_pt_replica.argtypes = [vp, C.c_int32, C.POINTER(Replica)]
    L.or_pt_coords.argtypes = [vp, C.c_int32, vp]
    L.or_pt_rng.argtypes = [vp, u64p]
    L.or_pt_spins.argtypes = [vp, C.c_int32, vp]
    L.or_zmc_new.argtypes = [vp, C.c_uint64, C.c_double, C.c_uint64, vp, vp]; L.or_zmc_new.restype = vp
    L.or_zmc_run.argtypes = [vp, C.c_uint64]
    L.or_zmc_replicas.argtypes = [vp]; L.or_zmc_replicas.restype = C.c_int32
    L.or_zmc_replica.argtypes = [vp, C.c_int32, vp]
    L.or_zmc_coords.argtypes = [vp, C.c_int32, vp]
    L.or_zmc_spins.argtypes = [vp, C.c_int32, vp]
    L.or_zmc_rng.argtypes = [vp, u64p]
    L.or_zmc_median.argtypes = [vp, vp]; L.or_zmc_median.restype = C.c_int32
    L.or_zmc_moves.argtypes = [vp]; L.or_zmc_moves.restype = C.c_uint64
    L.or_zmc_free.argtypes = [vp]
    L.or_bmc_new.argtypes = [vp, vp, vp]; L.or_bmc_new.restype = vp
    L.or_bmc_new_scripted.argtypes = [vp, C.c_double, vp, C.c_int64]; L.or_bmc_new_scripted.restype = vp
    L.or_bmc_run.argtypes = [vp, C.c_uint64]; L.or_bmc_run.restype = C.c_int
    L.or_bmc_state.argtypes = [vp, vp]
    L.or_bmc_bins.argtypes = [vp] + [vp] * 9
    L.or_bmc_lnw_at.argtypes = [vp, C.c_double]; L.or_bmc_lnw_at.restype = C.c_double
    L.or_bmc_count_at.argtypes = [vp, C.c_double]; L.or_bmc_count_at.restype = C.c_double
    L.or_bmc_coords.argtypes = [vp, vp]
    L.or_bmc_spins.argtypes = [vp, vp]
    L.or_bmc_free.argtypes = [vp]
    L.or_bmc_hires.argtypes = [vp, vp, vp]; L.or_bmc_hires.restype = C.c_int64
    L.or_pt_free.argtypes = [vp]
    L.or_prop_features.argtypes = [C.c_int32, vp, C.c_int32, C.c_double, C.c_double, C.c_double, vp]
    L.or_prop_mlp.argtypes = [vp, vp]; L.or_prop_mlp.restype = C.c_float
    L.or_prop_gaussian.argtypes = [u64p, vp]
    L.or_prop_lnh.argtypes = [C.c_float, C.c_float, C.c_float]; L.or_prop_lnh.restype = C.c_double
    L.or_prop_loss_grad.argtypes = [vp, vp, C.c_int64, vp]; L.or_prop_loss_grad.restype = C.c_double
    L.or_prop_loss_grad_f64.argtypes = [vp, vp, C.c_int64, vp]; L.or_prop_loss_grad_f64.restype = C.c_double
    L.or_prop_probe.argtypes = [vp, C.c_uint64, C.c_int32, vp]
    L.or_mc_plan_learned.argtypes = [vp, dp, dp, vp]; L.or_mc_plan_learned.restype = C.c_int
    _LIB = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


PROP_NF, PROP_H, PROP_NP, PROP_NS, PROP_R, PROP_REC = 8, 32, 1377, 16, 2, 24
PROP_SMIN, PROP_SMAX = -4.605170186, 0.0
_KEEP = {}


def make_params(**kw):
    """Walker parameters; `mlp` (float32[1377]) selects the learned move."""
    p = Params()
    defaults = dict(energy_bin=0.0, move_param=0.05, rc2=2.0 ** (1.0 / 3.0))
    defaults.update(kw)
    mlp = defaults.pop("mlp", None)
    for k, v in defaults.items():
        setattr(p, k, v)
    if mlp is not None:
        w = np.ascontiguousarray(mlp, dtype=np.float32)
        assert w.shape == (PROP_NP,)
        _KEEP[id(p)] = w
        p.mlp = w.ctypes.data
        p.move_kind = 1
    return p


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def prop_features(coords, skip, pos):
    c = np.ascontiguousarray(coords, dtype=np.float64)
    f = np.zeros(PROP_NF, np.float32)
    lib().or_prop_features(len(c) // 3, _ptr(c), skip, pos[0], pos[1], pos[2], _ptr(f))
    return f


def prop_mlp(w, f):
    w = np.ascontiguousarray(w, dtype=np.float32)
    f = np.ascontiguousarray(f, dtype=np.float32)
    return lib().or_prop_mlp(_ptr(w), _ptr(f))


def prop_gaussian(rng):
    xi = np.zeros(3, np.float32)
    lib().or_prop_gaussian(rng.s, _ptr(xi))
    return xi


def prop_lnh(s, s_new, xi2):
    return lib().or_prop_lnh(s, s_new, xi2)


def prop_loss_grad(w, records, f64=False):
    r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, PROP_REC)
    g = np.zeros(PROP_NP)
    if f64:
        w = np.ascontiguousarray(w, dtype=np.float64)
        loss = lib().or_prop_loss_grad_f64(_ptr(w), _ptr(r), len(r), _ptr(g))
    else:
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = lib().or_prop_loss_grad(_ptr(w), _ptr(r), len(r), _ptr(g))
    return loss, g


class RNG:
    """Xoroshiro128Plus seeded by SplitMix64 (src/rng.rs:27, energy.rs:835)."""

    def __init__(self, seed=0, state=None):
        self.s = (C.c_uint64 * 2)()
        if state is not None:
            self.s[0], self.s[1] = state
        else:
            lib().or_rng_seed(seed, self.s)

    def next_u64(self):
        return lib().or_rng_next(self.s)

    def gen_range(self, lo, hi):
        return lib().or_gen_range(self.s, lo, hi)

    def uniform_index(self, n):
        return lib().or_uniform_index(self.s, n)

    def gen_f64(self):
        return lib().or_gen_f64(self.s)

    def uniform_f64(self, lo, hi):
        return lib().or_uniform_f64(self.s, lo, hi)

    @property
    def state(self):
        return (self.s[0], self.s[1])


BCOUNTS_FIELDS = [("min_total", "<f8"), ("max_total", "<f8"), ("e_max_total", "<f8"), ("e_max_count", "<f8"),
                  ("min_count", "<u8"), ("max_count", "<u8"), ("total_count", "<u8"), ("exists", "<i4"),
                  ("reserved", "<i4")]
BSTATE_DTYPE = np.dtype([
    ("rng", "<u8", 2), ("E", "<f8"), ("error", "<f8"), ("moves", "<u8"), ("accepted_moves", "<u8"),
    ("acceptance_rate", "<f8"), ("translation_scale", "<f8"), ("min", "<f8"), ("width", "<f8"), ("min_e", "<f8"),
    ("max_e", "<f8"), ("len", "<i8"), ("method", "<i4"), ("status", "<i4"), ("num_states", "<u8"), ("tL", "<u8"),
    ("min_T", "<f8"), ("too_lo", "<f8"), ("too_hi", "<f8"), ("tF", "<f8"), ("latest_parameter", "<f8"),
    ("t0", "<f8"), ("gamma", "<f8"),
] + [(name, np.dtype(BCOUNTS_FIELDS)) for name in ("lnw", "energy", "t_found", "hist", "pressure")])


class BWalker:
    """One walker of the binning path (energy_binning.rs over binning/histogram.rs)."""

    def __init__(self, params, coords=None, spins=None, script=None, e0=0.0):
        self.p = params
        if script is not None:
            sc = np.ascontiguousarray(script, dtype=np.float64)
            self.h = lib().or_bmc_new_scripted(C.byref(params), e0, _ptr(sc), len(sc))
            return
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        self.h = lib().or_bmc_new(C.byref(params), _ptr(c), _ptr(sp))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_bmc_free(self.h)
            self.h = None

    def run(self, moves):
        return lib().or_bmc_run(self.h, moves)

    def state(self):
        st = np.zeros((), BSTATE_DTYPE)
        lib().or_bmc_state(self.h, _ptr(st))
        return st

    def bins(self):
        n = int(self.state()["len"])
        out = {k: np.zeros(n) for k in ("lnw_total", "energy_total", "t_found_total", "pressure_total")}
        out.update({k: np.zeros(n, np.uint64) for k in ("lnw_count", "energy_count", "t_found_count", "hist_count",
                                                         "pressure_count")})
        lib().or_bmc_bins(self.h, *[_ptr(out[k]) for k in (
            "lnw_total", "lnw_count", "energy_total", "energy_count", "t_found_total", "t_found_count", "hist_count",
            "pressure_total", "pressure_count")])
        return out

    def lnw_at(self, e):
        return lib().or_bmc_lnw_at(self.h, e)

    def count_at(self, e):
        return lib().or_bmc_count_at(self.h, e)

    def hires(self):
        """The high_resolution side histogram: (statistics dict, counts), or None."""
        st = np.zeros(10)
        n = lib().or_bmc_hires(self.h, _ptr(st), None)
        if n < 0:
            return None
        counts = np.zeros(max(n, 1), np.uint64)
        lib().or_bmc_hires(self.h, _ptr(st), _ptr(counts))
        keys = ("min", "width", "min_e", "max_e", "len", "total_count", "min_count", "max_count", "e_max_count",
                "min_total")
        d = dict(zip(keys, st.tolist()))
        for k in ("len", "total_count", "min_count", "max_count"):
            d[k] = int(d[k])
        return d, counts[:n]

    def coords(self):
        c = np.zeros(3 * self.p.n)
        lib().or_bmc_coords(self.h, _ptr(c))
        return c

    def spins(self):
        s = np.zeros(self.p.n * self.p.n, np.int8)
        lib().or_bmc_spins(self.h, _ptr(s))
        return s


class Walker:
    """One EnergyMC walker (energy.rs:167-210) over a system."""

    def __init__(self, params, coords=None, spins=None, script=None, e0=0.0):
        L = lib()
        self.p = params
        if script is not None:
            s = np.ascontiguousarray(script, dtype=np.float64)
            self.h = L.or_mc_new_scripted(C.byref(params), e0, s.ctypes.data_as(C.POINTER(C.c_double)), len(s))
        else:
            c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
            sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
            self.h = L.or_mc_new(C.byref(params), _ptr(c), _ptr(sp))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mc_free(self.h)
            self.h = None

    def run(self, moves):
        return lib().or_mc_run(self.h, moves)

    def trace(self, cap):
        lib().or_mc_trace(self.h, cap)

    def trace_copy(self, cap):
        e = np.zeros(cap, np.float64)
        a = np.zeros(cap, np.uint8)
        n = lib().or_mc_trace_copy(self.h, _ptr(e), _ptr(a))
        return e[:n], a[:n]

    def last_move(self):
        """The last move's decision inputs: proposed energy (NaN: None), the
        learned move's log acceptance ratio and uniform (NaN: none drawn) and
        its wall margin |to|^2 - R^2."""
        out = np.zeros(4)
        lib().or_mc_last_move(self.h, _ptr(out))
        return dict(e2=out[0], lr=out[1], u=out[2], wall=out[3])

    def load_state(self, scalars, bins, coords):
        """Overwrite this (LJ/WCA SAD) walker with another engine's state:
        `scalars` a dict with the scalars() keys, `bins` a dict with the
        bins() keys, `coords` SoA positions."""
        need = {"E", "error", "moves", "accepted", "acceptance_rate", "translation_scale", "rng", "bins_min",
                "bins_width", "len", "max_S", "max_S_index", "method", "status", "too_lo", "too_hi", "lambda_",
                "min_T", "tL", "tF", "num_states", "highest_hist"}
        missing = need - set(scalars)
        assert not missing, f"load_state: missing {sorted(missing)}"
        sc = Scalars()
        for name, _ in Scalars._fields_:
            if name in scalars:
                v = scalars[name]
                if name == "rng":
                    sc.rng[0], sc.rng[1] = int(v[0]), int(v[1])
                else:
                    setattr(sc, name, type(getattr(sc, name))(v))
        arr = {k: np.ascontiguousarray(bins[k], dtype=dt) for k, dt in (
            ("lnw", np.float64), ("histogram", np.uint64), ("t_found", np.uint64), ("energy_total", np.float64),
            ("energy_squared_total", np.float64), ("visited", np.uint8), ("round_trips", np.uint64))}
        c = np.ascontiguousarray(coords, dtype=np.float64)
        rc = lib().or_mc_load_state(self.h, C.byref(sc), *[_ptr(arr[k]) for k in (
            "lnw", "histogram", "t_found", "energy_total", "energy_squared_total", "visited", "round_trips")],
            _ptr(c))
        assert rc == 0, "load_state: LJ/WCA SAD walkers only"

    def scalars(self):
        s = Scalars()
        lib().or_mc_scalars(self.h, C.byref(s))
        return s.as_dict()

    def bins(self):
        n = self.scalars()["len"]
        out = dict(
            lnw=np.zeros(n), histogram=np.zeros(n, np.uint64), t_found=np.zeros(n, np.uint64),
            energy_total=np.zeros(n), energy_squared_total=np.zeros(n), visited=np.zeros(n, np.uint8),
            round_trips=np.zeros(n, np.uint64), wl_hist=np.zeros(n, np.uint64),
            pressure_count=np.zeros(n, np.uint64), pressure_total=np.zeros(n),
        )
        lib().or_mc_bins(self.h, *[_ptr(out[k]) for k in (
            "lnw", "histogram", "t_found", "energy_total", "energy_squared_total", "visited",
            "round_trips", "wl_hist", "pressure_count", "pressure_total")])
        return out

    def coords(self):
        c = np.zeros(3 * self.p.n)
        lib().or_mc_coords(self.h, _ptr(c))
        return c

    def spins(self):
        s = np.zeros(self.p.n * self.p.n, np.int8)
        lib().or_mc_spins(self.h, _ptr(s))
        return s

    def plan_move(self, scale):
        e = C.c_double()
        return e.value if lib().or_mc_plan_move(self.h, scale, C.byref(e)) else None

    def confirm(self):
        lib().or_mc_confirm(self.h)

    def plan_learned(self):
        """(e_new, ln H, new position) of a learned move, or None if the wall rejects it."""
        e, h = C.c_double(), C.c_double()
        to = np.zeros(3)
        if not lib().or_mc_plan_learned(self.h, C.byref(e), C.byref(h), _ptr(to)):
            return None
        return e.value, h.value, to

    def probe(self, seed, count):
        """`count` training records (features, J on the scale grid) of this walker."""
        out = np.zeros((count, PROP_REC), np.float32)
        lib().or_prop_probe(self.h, seed, count, _ptr(out))
        return out

    def energy(self):
        return lib().or_mc_energy(self.h)

    def compute_energy(self):
        return lib().or_mc_compute_energy(self.h)

    def reference_energy(self):
        return lib().or_mc_reference_energy(self.h)


def lj_init(n, radius, seed):
    c = np.zeros(3 * n)
    rc = lib().or_lj_init(n, radius, seed, _ptr(c))
    assert rc == 0
    return c


def wca_init(n, box, seed):
    c = np.zeros(3 * n)
    lib().or_wca_init(n, box, seed, _ptr(c))
    return c


def ising_spins(n, rule, seed):
    s = np.zeros(n * n, np.int8)
    lib().or_ising_init_spins(n, rule, seed, _ptr(s))
    return s


def ising_energy(n, spins):
    s = np.ascontiguousarray(spins, dtype=np.int8)
    return lib().or_ising_energy(n, _ptr(s))


def lj_energy(coords, order_g=0):
    c = np.ascontiguousarray(coords, dtype=np.float64)
    return lib().or_lj_energy(len(c) // 3, _ptr(c), order_g)


def lj_energy_ld(coords):
    c = np.ascontiguousarray(coords, dtype=np.float64)
    return lib().or_lj_energy_ld(len(c) // 3, _ptr(c))


def wca_energy(coords, box, rc2, order_g=0):
    c = np.ascontiguousarray(coords, dtype=np.float64)
    return lib().or_wca_energy(len(c) // 3, _ptr(c), box, rc2, order_g)


def wca_energy_celllist(coords, box, rc2):
    c = np.ascontiguousarray(coords, dtype=np.float64)
    return lib().or_wca_energy_celllist(len(c) // 3, _ptr(c), box, rc2)


def wca_energy_ld(coords, box, rc2):
    c = np.ascontiguousarray(coords, dtype=np.float64)
    return lib().or_wca_energy_ld(len(c) // 3, _ptr(c), box, rc2)


def sad_gamma(lam, t, tF, ns):
    return lib().or_sad_gamma(lam, t, tF, ns)


def bench(params, threads, moves, seed0=0):
    return lib().or_bench(C.byref(params), threads, moves, seed0)


class Tempering:
    """A parallel-tempering ladder (src/mc/tempering.rs)."""

    def __init__(self, params, temperatures, seed, coords=None, spins=None, canonical_steps=1):
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        self.R = len(T)
        self.n = params.n
        self.system = params.system
        self.h = lib().or_pt_new(C.byref(params), self.R, _ptr(T), seed, _ptr(c), _ptr(sp), canonical_steps)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_pt_free(self.h)
            self.h = None

    def run(self, rounds):
        lib().or_pt_run(self.h, rounds)

    def replica(self, r):
        out = Replica()
        lib().or_pt_replica(self.h, r, C.byref(out))
        d = {name: getattr(out, name) for name, _ in Replica._fields_}
        d["rng"] = list(out.rng)
        return d

    def coords(self, r):
        c = np.zeros(3 * self.n)
        lib().or_pt_coords(self.h, r, _ptr(c))
        return c

    def spins(self, r):
        s = np.zeros(self.n * self.n, np.int8)
        lib().or_pt_spins(self.h, r, _ptr(s))
        return s

    def rng(self):
        s = (C.c_uint64 * 2)()
        lib().or_pt_rng(self.h, s)
        return list(s)


class ZRep(C.Structure):
    _fields_ = [("max_energy", C.c_double), ("cutoff_energy", C.c_double), ("rejected_count", C.c_uint64),
                ("accepted_count", C.c_uint64), ("above_count", C.c_uint64), ("below_count", C.c_uint64),
                ("upwelling_count", C.c_uint64), ("above_total", C.c_double), ("below_total", C.c_double),
                ("above_total_squared", C.c_double), ("below_total_squared", C.c_double),
                ("lowest_max_energy", C.c_double), ("unique_visitors", C.c_uint64), ("collecting_data", C.c_int32),
                ("reserved", C.c_int32), ("rng", C.c_uint64 * 2), ("translation_scale", C.c_double),
                ("E", C.c_double), ("error", C.c_double), ("pressure_total", C.c_double),
                ("pressure_count", C.c_uint64)]


class Zmc:
    """Energy-ceiling replicas (src/mc/energy_replicas.rs)."""

    def __init__(self, params, seed, min_T=0.2, isbnb=64, coords=None, spins=None):
        c = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        sp = None if spins is None else np.ascontiguousarray(spins, dtype=np.int8)
        self.n = params.n
        self.h = lib().or_zmc_new(C.byref(params), seed, min_T, isbnb, _ptr(c), _ptr(sp))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_zmc_free(self.h)
            self.h = None

    def run(self, rounds):
        lib().or_zmc_run(self.h, rounds)

    def replicas(self):
        out = []
        for k in range(lib().or_zmc_replicas(self.h)):
            r = ZRep()
            lib().or_zmc_replica(self.h, k, C.byref(r))
            d = {name: getattr(r, name) for name, _ in ZRep._fields_ if name != "reserved"}
            d["rng"] = list(r.rng)
            out.append(d)
        return out

    def coords(self, k):
        c = np.zeros(3 * self.n)
        lib().or_zmc_coords(self.h, k, _ptr(c))
        return c

    def spins(self, k):
        s = np.zeros(self.n * self.n, np.int8)
        lib().or_zmc_spins(self.h, k, _ptr(s))
        return s

    def rng(self):
        s = (C.c_uint64 * 2)()
        lib().or_zmc_rng(self.h, s)
        return list(s)

    def median(self):
        n = lib().or_zmc_median(self.h, None)
        a = np.zeros(n)
        lib().or_zmc_median(self.h, _ptr(a))
        return a

    def moves(self):
        return lib().or_zmc_moves(self.h)


def rng_jump(state):
    s = (C.c_uint64 * 2)(*state)
    lib().or_rng_jump(s)
    return list(s)
