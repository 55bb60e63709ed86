"""Pins of the oracle's binning walker (src/mc/energy_binning.rs over
src/mc/binning/histogram.rs).

- A three-move scripted trajectory worked out by hand from the reference's
  source: the top-edge index rule (histogram.rs:129-140), bins growing at
  both ends (histogram.rs:141-158), BinCounts' running statistics
  (histogram.rs:59-81, 219-245), the "t_found" extra, SAD's range
  extensions with set_lnw zeroing counts (energy_binning.rs:341-403,
  histogram.rs:182-190), count_states, gamma (energy_binning.rs:503-530).
- binning.rs:340-364's known answer: a first datum of 7 has total, mean and
  running maximum 7.
- Convergence of SAD, SAMC, WL, 1/t-WL and WL with min_gamma to the exact
  4x4 Ising density of states (transfer-matrix fixture).
- Bookkeeping invariants on LJ and WCA walkers (every move lands in the
  "energy" extra; the WCA pressure extra every N^2 moves).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _script_walker(script, e0=0.0, **kw):
    p = O.make_params(system=O.SCRIPTED, n=1, method=O.SAD, method_param=1.0, energy_bin=1.0, **kw)
    return O.BWalker(p, script=np.array(script, float), e0=e0)


def _bc(st, name):
    return {k: st[name][k].item() for k in ("min_total", "max_total", "e_max_total", "e_max_count", "min_count",
                                             "max_count", "total_count", "exists")}


def test_scripted_trajectory_by_hand():
    b = _script_walker([1.0, -1.0, np.nan])
    st = b.state()
    assert float(st["min"]) == -0.5 and int(st["len"]) == 0  # Bins::new: (round(E/w) - 0.5) w, empty
    b.run(1)
    st, bins = b.state(), b.bins()
    # move 1: the first prep sets min = floor(E/w) w = 0; E' = 1.0 lands on the top edge of the one bin
    # (index 0), then a second bin is pushed; set_lnw above too_hi zeroes that bin's count
    assert float(st["E"]) == 1.0 and int(st["len"]) == 2 and float(st["min"]) == 0.0
    assert float(st["acceptance_rate"]) == 1.0
    np.testing.assert_array_equal(bins["lnw_total"], [0.0, 0.0])
    np.testing.assert_array_equal(bins["lnw_count"], [0, 0])
    np.testing.assert_array_equal(bins["energy_count"], [1, 0])
    np.testing.assert_array_equal(bins["t_found_total"], [0.0, 1.0])
    np.testing.assert_array_equal(bins["t_found_count"], [0, 1])
    assert _bc(st, "lnw") == dict(min_total=0.0, max_total=0.0, e_max_total=-np.inf, e_max_count=1.0,
                                  min_count=0, max_count=1, total_count=1, exists=1)
    assert _bc(st, "energy") == dict(min_total=0.0, max_total=0.0, e_max_total=-np.inf, e_max_count=-np.inf,
                                     min_count=0, max_count=1, total_count=1, exists=1)
    assert _bc(st, "t_found")["max_total"] == 1.0 and _bc(st, "hist")["exists"] == 0
    assert (float(st["too_lo"]), float(st["too_hi"])) == (0.0, 1.0)
    assert int(st["num_states"]) == 1 and int(st["tL"]) == 1 and float(st["tF"]) == 1.0
    assert float(st["latest_parameter"]) == 1.0

    b.run(1)
    st, bins = b.state(), b.bins()
    # move 2: E' = -1 is accepted from below too_lo; one bin is inserted in front (min -= w);
    # gamma = (1 + 2/1) / (1 + 2/1 (2/1)) = 0.6
    assert float(st["E"]) == -1.0 and int(st["len"]) == 3 and float(st["min"]) == -1.0
    np.testing.assert_array_equal(bins["lnw_total"], [0.6, 0.0, 0.0])
    np.testing.assert_array_equal(bins["lnw_count"], [1, 0, 0])
    np.testing.assert_array_equal(bins["energy_total"], [0.0, 0.0, 1.0])
    np.testing.assert_array_equal(bins["energy_count"], [0, 1, 1])
    np.testing.assert_array_equal(bins["t_found_total"], [2.0, 0.0, 1.0])
    assert _bc(st, "lnw") == dict(min_total=0.0, max_total=0.6, e_max_total=-1.0, e_max_count=1.0,
                                  min_count=0, max_count=1, total_count=2, exists=1)
    assert _bc(st, "energy") == dict(min_total=0.0, max_total=1.0, e_max_total=-np.inf, e_max_count=-np.inf,
                                     min_count=0, max_count=1, total_count=2, exists=1)
    assert _bc(st, "t_found")["max_total"] == 2.0 and float(st["tF"]) == 1.0

    b.run(1)
    st, bins = b.state(), b.bins()
    # move 3: no move; gamma = (1 + 3)/(1 + 9) = 0.4 takes ln w(-1) to 1.0 and its count to 2 > 1, so
    # too_lo extends to -1: set_lnw below the old too_lo gives -0.5 (and count 0); num_states 2
    assert float(st["E"]) == -1.0
    np.testing.assert_array_equal(bins["lnw_total"], [-0.5, 0.0, 0.0])
    np.testing.assert_array_equal(bins["lnw_count"], [0, 0, 0])
    np.testing.assert_array_equal(bins["energy_total"], [-1.0, 0.0, 1.0])
    np.testing.assert_array_equal(bins["energy_count"], [1, 1, 1])
    assert _bc(st, "lnw") == dict(min_total=0.0, max_total=1.0, e_max_total=-1.0, e_max_count=-1.0,
                                  min_count=0, max_count=2, total_count=3, exists=1)
    assert _bc(st, "energy") == dict(min_total=-1.0, max_total=1.0, e_max_total=-np.inf, e_max_count=-np.inf,
                                     min_count=1, max_count=1, total_count=3, exists=1)
    assert (float(st["too_lo"]), float(st["too_hi"])) == (-1.0, 1.0)
    assert int(st["num_states"]) == 2 and int(st["tL"]) == 3 and float(st["tF"]) == 2.0
    assert float(st["latest_parameter"]) == 2.0
    assert int(st["moves"]) == 3 and int(st["accepted_moves"]) == 2


def test_first_datum_known_answer():
    # binning.rs:340-364: accumulating a datum of 7 gives total, mean and maximum 7
    b = _script_walker([np.nan], e0=7.0)
    b.run(1)
    st, bins = b.state(), b.bins()
    assert bins["energy_total"][0] == 7.0 and bins["energy_count"][0] == 1
    assert bins["energy_total"][0] / bins["energy_count"][0] == 7.0
    assert _bc(st, "energy")["max_total"] == 7.0
    assert bins["lnw_count"].sum() > 0


def _exact_lng():
    d = json.load(open(os.path.join(GOLDEN, "ising4_dos.json")))
    return {float(e): np.log(float(int(c))) for e, c in zip(d["energies"], d["counts"])}


@pytest.mark.parametrize("name,kw,tol", [
    ("sad", dict(method=O.SAD, method_param=1.0, energy_bin=1.0), 0.08),
    ("samc", dict(method=O.SAMC, method_param=1000.0, energy_bin=1.0), 0.3),
    ("wl", dict(method=O.WL, energy_bin=4.0, has_min_e=1, min_e=-24.0, has_max_e=1, max_e=24.0), 0.4),
    ("inv_t_wl", dict(method=O.INV_T_WL, energy_bin=4.0, has_min_e=1, min_e=-24.0, has_max_e=1, max_e=24.0), 0.06),
    ("wl_min_gamma", dict(method=O.WL, method_param=1e-4, has_min_gamma=1, energy_bin=4.0, has_min_e=1,
                          min_e=-24.0, has_max_e=1, max_e=24.0), 0.4),
])
def test_converges_to_exact_ising_dos(name, kw, tol):
    lng = _exact_lng()
    for seed in (1, 2):
        b = O.BWalker(O.make_params(system=O.ISING, n=4, ising_init=1, seed=seed, **kw))
        assert b.run(3000000) == 0
        st, bins = b.state(), b.bins()
        n = int(st["len"])
        lo_edge = float(st["min"]) + np.arange(n) * float(st["width"])
        diffs = []
        for i in range(n):
            e = lo_edge[i]
            if e not in lng:
                continue
            if name == "sad" and not (st["too_lo"] <= e <= st["too_hi"]):
                continue
            diffs.append(bins["lnw_total"][i] - lng[e])
        assert len(diffs) >= 8, (name, seed, len(diffs))
        assert max(diffs) - min(diffs) < tol, (name, seed, np.round(np.array(diffs) - np.mean(diffs), 3))
        if name == "inv_t_wl":
            assert int(st["method"]) == O.SAMC and float(st["t0"]) == n  # switched to 1/t with t0 = num_states
        if name == "wl_min_gamma":
            assert float(st["gamma"]) == 0.0


def test_energy_extra_counts_every_move_lj_and_wca():
    p = O.make_params(system=O.LJ, n=13, radius=2.5, method=O.SAD, method_param=0.05, energy_bin=0.05,
                      move_param=0.1, has_max_e=1, max_e=0.0, order_g=4, seed=5)
    b = O.BWalker(p)
    b.run(50000)
    st, bins = b.state(), b.bins()
    assert int(st["status"]) == 0 and float(st["E"]) < 0.0
    assert int(st["energy"]["total_count"]) == 50000 == int(bins["energy_count"].sum())
    assert int(st["lnw"]["total_count"]) == 50000
    assert bins["t_found_count"].sum() >= 1
    assert st["energy"]["max_count"] == bins["energy_count"].max()
    assert st["energy"]["min_count"] == bins["energy_count"].min()
    n = 32
    L = (n / 0.5) ** (1 / 3)
    pw = O.make_params(system=O.WCA, n=n, box=L, method=O.SAD, method_param=0.1, energy_bin=0.1, move_param=0.1,
                       order_g=8, collect_pressure=1, has_max_e=1, max_e=2.5 * n, seed=6)
    bw = O.BWalker(pw)
    moves = 3 * n * n + 7
    bw.run(moves)
    st, bins = bw.state(), bw.bins()
    assert int(st["pressure"]["total_count"]) == 3 == int(bins["pressure_count"].sum())
    assert np.all(bins["pressure_total"][bins["pressure_count"] > 0] > 0.0)


# ---- the linear interpolation (binning/linear.rs)

def test_linear_known_answers():
    # linear.rs:170-186: one increment of weight 1 at E = 1.9 with bins of width 1
    p = O.make_params(system=O.SCRIPTED, n=1, method=O.SAMC, method_param=1000.0, energy_bin=1.0, binning=1)
    b = O.BWalker(p, script=np.array([np.nan]), e0=1.9)
    b.run(1)  # SAMC's gamma is 1 at t = 1 <= t0
    st, bins = b.state(), b.bins()
    assert float(st["min"]) == 1.0 and int(st["len"]) == 2  # one more bin kept above the energy
    assert b.lnw_at(1.95) > b.lnw_at(1.85)
    assert b.lnw_at(1.0) > 0.0 and b.lnw_at(0.51) > 0.0 and b.lnw_at(2.49) > 0.0
    assert b.lnw_at(3.01) == 0.0 and b.lnw_at(-0.01) == 0.0
    # the weight splits 0.1 / 0.9 over the bins at 1 and 2 (fractional counts too)
    np.testing.assert_allclose(bins["lnw_total"], [0.1, 0.9], rtol=1e-12)
    np.testing.assert_allclose(bins["lnw_count"].view(np.float64), [0.1, 0.9], rtol=1e-12)
    assert abs(b.lnw_at(1.95) - (0.9 * 0.95 + 0.1 * 0.05)) < 1e-12
    assert abs(b.lnw_at(0.51) - 0.1 * 0.51) < 1e-12  # below the first bin: one-sided weight 1 + fidx


@pytest.mark.parametrize("name,kw,tol", [
    ("sad", dict(method=O.SAD, method_param=1.0), 0.35),
    ("samc", dict(method=O.SAMC, method_param=1000.0), 0.2),
])
def test_linear_converges_to_exact_ising_dos(name, kw, tol):
    lng = _exact_lng()
    for seed in (1, 2):
        b = O.BWalker(O.make_params(system=O.ISING, n=4, ising_init=1, seed=seed, energy_bin=4.0, binning=1, **kw))
        assert b.run(3000000) == 0
        st = b.state()
        levels = [e for e in lng if name != "sad" or st["too_lo"] <= e <= st["too_hi"]]
        diffs = np.array([b.lnw_at(e) - lng[e] for e in levels])
        assert len(levels) >= 8
        assert diffs.max() - diffs.min() < tol, (name, seed, np.round(diffs - diffs.mean(), 3))


def test_linear_energy_extra_weights_sum_to_moves():
    p = O.make_params(system=O.LJ, n=13, radius=2.5, method=O.SAD, method_param=0.05, energy_bin=0.05,
                      move_param=0.1, has_max_e=1, max_e=0.0, order_g=4, seed=5, binning=1)
    b = O.BWalker(p)
    b.run(40000)
    st, bins = b.state(), b.bins()
    assert int(st["status"]) == 0
    assert int(st["energy"]["total_count"]) == 40000 == int(st["lnw"]["total_count"])
    w = bins["energy_count"].view(np.float64)
    assert abs(w.sum() - 40000) < 1e-6  # each move adds weights 1 - offset and offset


# ---- the high_resolution side histogram (energy_binning.rs:125, 328-330)

@pytest.mark.parametrize("method,kw", [(O.SAMC, dict(method_param=300.0)),
                                       (O.WL, dict(has_min_e=1, min_e=-32.0, has_max_e=1, max_e=32.0))],
                         ids=["samc", "wl"])
def test_high_resolution_at_own_width_repeats_ln_w_counts(method, kw):
    # SAMC and WL never reset ln w counts, and the side histogram is
    # incremented with the same energies: at the walker's own width it holds
    # the same window, counts and count statistics (a wrong index, edge or
    # window growth in either histogram breaks the equality)
    p = O.make_params(system=O.ISING, n=4, ising_init=1, method=method, energy_bin=1.0, high_resolution_de=1.0,
                      seed=3, **kw)
    b = O.BWalker(p)
    assert b.run(20000) == 0
    h, counts = b.hires()
    s, bins = b.state(), b.bins()
    np.testing.assert_array_equal(counts, bins["lnw_count"])
    assert (h["min"], h["len"], h["min_e"], h["max_e"]) == (s["min"], s["len"], s["min_e"], s["max_e"])
    for k in ("min_count", "max_count", "total_count", "e_max_count"):
        assert h[k] == s["lnw"][k], k
    assert h["total_count"] == 20000


def test_high_resolution_fine_bins_nest_in_coarse_ones():
    # LJ13 under SAMC with ln w bins of 1 and side bins of 1/4: every side
    # bin lies inside one ln w bin (all edges are exact multiples of 1/4), so
    # the side counts summed per ln w bin are the ln w counts; the statistics
    # are those of the counts themselves
    p = O.make_params(system=O.LJ, n=13, radius=2.5, order_g=4, method=O.SAMC, method_param=500.0, energy_bin=1.0,
                      move_param=0.1, has_max_e=1, max_e=0.0, high_resolution_de=0.25, seed=9)
    b = O.BWalker(p)
    assert b.run(30000) == 0
    h, fine = b.hires()
    s, coarse = b.state(), b.bins()["lnw_count"]
    assert h["len"] > 4 * 3
    edges = h["min"] + 0.25 * np.arange(h["len"])
    ci = np.floor((edges - s["min"]) / 1.0).astype(int)
    assert ci.min() >= 0 and ci.max() < len(coarse)
    summed = np.zeros(len(coarse), np.uint64)
    np.add.at(summed, ci, fine)
    np.testing.assert_array_equal(summed, coarse)
    assert h["total_count"] == int(fine.sum()) == 30000
    assert h["min_count"] == int(fine.min()) and h["max_count"] == int(fine.max())
    # e_max_count is the energy that first took a bin past the old maximum:
    # its bin holds the maximum
    k = int(np.floor((h["e_max_count"] - h["min"]) / 0.25))
    assert fine[k] == h["max_count"]
    assert h["min_e"] == s["min_e"] and h["max_e"] == s["max_e"]
    assert h["min"] <= h["min_e"] and h["max_e"] < h["min"] + 0.25 * h["len"]


def test_no_high_resolution_by_default():
    b = O.BWalker(O.make_params(system=O.ISING, n=4, method=O.SAD, method_param=1.0, energy_bin=1.0, seed=1))
    b.run(100)
    assert b.hires() is None
