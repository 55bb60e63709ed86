"""Pins of the oracle's Monte Carlo methods (energy.rs) against exact
results on the 4x4 and 8x8 Ising models, whose densities of states are
known exactly (tests/golden/ising*_dos.json).

- Canonical Metropolis (energy.rs:504-510) samples exp(-E/T) g(E).
- SAD, SAMC and 1/t-WL converge to ln g(E) (energy.rs:514-761).
- SAD gamma (energy.rs:22-38) against its closed-form limits.
- Bookkeeping invariants of EnergyMC on long runs: histogram total, energy
  sums against the recorded trajectory, t_found ordering, SAD num_states
  against the visited bins inside [too_lo, too_hi].
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def exact_lng(L):
    d = json.load(open(os.path.join(GOLDEN, f"ising{L}_dos.json")))
    e = np.array(d["energies"], float)
    c = np.array([float(int(x)) for x in d["counts"]])
    return e, np.log(c)


def walker_lng(w):
    """(bin centres, lnw, histogram) of one oracle walker."""
    s = w.scalars()
    b = w.bins()
    centres = s["bins_min"] + (np.arange(s["len"]) + 0.5) * s["bins_width"]
    return centres, b["lnw"], b["histogram"]


def test_canonical_samples_boltzmann():
    e, lng = exact_lng(4)
    T = 2.5
    w = O.Walker(O.make_params(system=O.ISING, n=4, method=O.CANONICAL, method_param=T, seed=3, ising_init=1))
    w.run(2_000_000)
    centres, _, hist = walker_lng(w)
    p_exact = np.exp(lng - e / T)
    p_exact /= p_exact.sum()
    counts = dict(zip(np.round(centres).astype(int), hist.astype(float)))
    obs = np.array([counts.get(int(x), 0.0) for x in e])
    n = obs.sum()
    mask = p_exact * n > 5000
    # 2e6 correlated samples: allow a few percent on every well-populated bin
    assert np.all(np.abs(obs[mask] / n - p_exact[mask]) / p_exact[mask] < 0.05)


@pytest.mark.parametrize("method,param,kw", [
    (O.SAD, 1.0, {}),
    (O.SAMC, 2000.0, {}),
    (O.INV_T_WL, 0.0, dict(has_min_e=1, min_e=-32.0, has_max_e=1, max_e=32.0)),
])
def test_methods_converge_to_exact_lng_4x4(method, param, kw):
    e, lng = exact_lng(4)
    ests = []
    for seed in range(4):
        w = O.Walker(O.make_params(system=O.ISING, n=4, method=method, method_param=param, seed=seed,
                                   ising_init=1, **kw))
        w.run(3_000_000)
        centres, lnw, hist = walker_lng(w)
        m = dict(zip(np.round(centres).astype(int), lnw))
        ests.append(np.array([m.get(int(x), np.nan) for x in e]))
    est = np.nanmean(np.array(ests), axis=0)
    # compare shapes over the energies every walker visited, relative to E = 0
    ok = ~np.isnan(est)
    if method == O.SAD:
        s = w.scalars()
        ok &= (e >= s["too_lo"]) & (e <= s["too_hi"])
    ref = int(np.where(e == 0)[0][0])
    diff = (est - est[ref]) - (lng - lng[ref])
    assert ok.sum() >= 8
    assert np.max(np.abs(diff[ok])) < 0.15


def test_sad_gamma_limits():
    g = O.sad_gamma
    # lambda * tF * num_states == 0 gives 0 (energy.rs:23-24)
    assert g(0.0, 10.0, 5.0, 3.0) == 0.0 and g(1.0, 10.0, 0.0, 3.0) == 0.0
    # t = tF = num_states: (lambda + 1) / (lambda + 1) = 1
    for lam in (0.5, 3.0, 1e4):
        assert abs(g(lam, 7.0, 7.0, 7.0) - 1.0) < 1e-15
    # long times: gamma -> num_states / t (the 1/t regime)
    for t in (1e6, 1e8):
        assert abs(g(1e-3, t, 100.0, 50.0) * t / 50.0 - 1.0) < 1e-3
    # large lambda: gamma -> 1
    assert abs(g(1e12, 100.0, 10.0, 5.0) - 1.0) < 1e-9


@pytest.mark.parametrize("system", ["ising", "lj"])
def test_energy_mc_bookkeeping_invariants(system):
    if system == "ising":
        p = O.make_params(system=O.ISING, n=8, method=O.SAD, method_param=1.0, seed=5, ising_init=1)
        moves = 200_000
    else:
        p = O.make_params(system=O.LJ, n=13, radius=2.5, method=O.SAD, method_param=0.05, move_param=0.1,
                          energy_bin=0.01, has_max_e=1, max_e=0.0, order_g=4, seed=5)
        moves = 50_000
    w = O.Walker(p)
    e0 = w.scalars()["E"]
    w.trace(moves)
    w.run(moves)
    energies, acc = w.trace_copy(moves)
    s = w.scalars()
    b = w.bins()
    assert b["histogram"].sum() == moves + 1
    assert s["accepted"] == acc.sum()
    np.testing.assert_allclose(b["energy_total"].sum(), e0 + energies.sum(), rtol=1e-12)
    # every visited bin's t_found is the first move that landed there
    idx = np.floor((energies - s["bins_min"]) / s["bins_width"]).astype(int)
    first = {}
    for k, i in enumerate(idx, start=1):
        first.setdefault(i, k)
    for i, t in first.items():
        if b["t_found"][i] != 0:
            assert b["t_found"][i] == t
    ilo = int((s["too_lo"] - s["bins_min"]) / s["bins_width"])
    ihi = int((s["too_hi"] - s["bins_min"]) / s["bins_width"])
    assert s["tF"] == int(b["t_found"][ilo:ihi + 1].max())
    if system == "ising":
        # with integer energies on bin centres, SAD's num_states is exactly the
        # number of visited bins of [too_lo, too_hi]
        assert s["num_states"] == int((b["histogram"][ilo:ihi + 1] > 0).sum())
