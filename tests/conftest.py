"""Shared test setup: import paths for the product package and the oracle,
the `gpu` marker, and helpers that map one parameter set onto both."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sad-monte-carlo_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built HIP library")


def oracle_params(sp, seed, geom=0.0, order_g=None):
    """The oracle parameters equivalent to one product SetParams + walker."""
    import oracle as O
    g = sp.lanes if order_g is None else order_g
    return O.make_params(
        system=sp.system, n=sp.n, method=sp.method, move_plan=sp.move_plan, method_param=sp.method_param,
        has_min_gamma=sp.has_min_gamma, precision=sp.precision, move_param=sp.move_param,
        energy_bin=sp.energy_bin, has_min_e=sp.has_min_e, has_max_e=sp.has_max_e, min_e=sp.min_e,
        max_e=sp.max_e, radius=geom if sp.system == 1 else 0.0, box=geom if sp.system == 2 else 0.0,
        rc2=sp.rc2, refresh=sp.refresh, order_g=g if sp.system != 0 else 0, ising_init=sp.ising_init,
        collect_pressure=sp.collect_pressure, binning=sp.binning, high_resolution_de=sp.high_resolution_de,
        seed=int(seed))


SCALAR_KEYS = [
    ("E", "E"), ("moves", "moves"), ("accepted_moves", "accepted"), ("acceptance_rate", "acceptance_rate"),
    ("translation_scale", "translation_scale"), ("bins_min", "bins_min"), ("bins_width", "bins_width"),
    ("len", "len"), ("max_S", "max_S"), ("max_S_index", "max_S_index"), ("method", "method"),
    ("status", "status"), ("too_lo", "too_lo"), ("too_hi", "too_hi"), ("latest_parameter", "lambda_"),
    ("tL", "tL"), ("tF", "tF"), ("num_states", "num_states"), ("highest_hist", "highest_hist"),
    ("t0", "t0"), ("wl_gamma", "wl_gamma"), ("wl_num_states", "wl_num_states"),
    ("wl_min_energy", "wl_min_energy"), ("wl_lowest_hist", "wl_lowest"), ("wl_highest_hist", "wl_highest"),
    ("wl_total_hist", "wl_total"), ("wl_len", "wl_len"), ("error", "error"), ("has_pressure", "has_pressure"),
]
SAD_ONLY = {"too_lo", "too_hi", "latest_parameter", "tL", "tF", "num_states", "highest_hist"}
WL_ONLY = {"wl_gamma", "wl_num_states", "wl_min_energy", "wl_lowest_hist", "wl_highest_hist", "wl_total_hist",
           "wl_len"}


def compare_walker(gpu_state, gpu_bins, orc, method, label=""):
    """Assert bit-identity of every field of one walker's EnergyMC state."""
    s = orc.scalars()
    for gk, ok in SCALAR_KEYS:
        if gk in SAD_ONLY and method != 0:
            continue
        if gk in WL_ONLY and method not in (2, 3):
            continue
        if gk == "t0" and method not in (1, 3):
            continue
        gv, ov = gpu_state[gk], s[ok]
        if isinstance(ov, float):
            assert np.float64(gv).tobytes() == np.float64(ov).tobytes(), f"{label} {gk}: gpu {gv!r} oracle {ov!r}"
        else:
            assert int(gv) == int(ov), f"{label} {gk}: gpu {gv} oracle {ov}"
    assert list(map(int, gpu_state["rng"])) == list(map(int, s["rng"])), f"{label} rng"
    ob = orc.bins()
    for k in ("lnw", "histogram", "t_found", "energy_total", "energy_squared_total", "visited", "round_trips"):
        np.testing.assert_array_equal(gpu_bins[k], ob[k], err_msg=f"{label} bins {k}")
    if method in (2, 3):
        np.testing.assert_array_equal(gpu_bins["wl_hist"], ob["wl_hist"], err_msg=f"{label} wl_hist")
    np.testing.assert_array_equal(gpu_bins["pressure_count"], ob["pressure_count"], err_msg=f"{label} p_count")
    np.testing.assert_array_equal(gpu_bins["pressure_total"], ob["pressure_total"], err_msg=f"{label} p_total")


def run_oracles(params, moves, threads=None):
    """Oracle walkers for `params` (a list of oracle Params), each run for
    `moves` moves, on parallel threads (ctypes releases the GIL inside the C
    call); returns the walkers in order."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle as O

    def one(p):
        w = O.Walker(p)
        w.run(moves)
        return w

    threads = threads or min(len(params), 16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, params))


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import sadmc_amd as S

    made = []

    def make():
        c = S.Context(0)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()
