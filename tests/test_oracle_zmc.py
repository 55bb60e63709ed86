"""Pins of the oracle's energy-ceiling replicas (src/mc/energy_replicas.rs).

- Sampling below a ceiling is uniform over configurations: every replica's
  fraction of samples above its cutoff equals the exact ratio
  sum_{cutoff < E < max} g(E) / sum_{E < max} g(E) of the 4x4 Ising density
  of states (transfer-matrix fixture); a mistaken boundary (>= for >, or
  <= for <) moves whole levels across and fails it.
- Bookkeeping: the two initial replicas count every sample (a randomization
  per round for the unbounded one, N^2 moves per round for the other); the
  ceilings decrease down the ladder and every cutoff lies below its ceiling;
  each new replica's ceiling is its predecessor's cutoff.
- LJ13: new replicas are made, each cut near the median of its band.
- WCA (8 atoms, box 3): the initial ceilings sit at the median and lower
  quartile of numpy's uniform periodic configurations; every later cutoff
  halves the uniform weight under its ceiling; the unbounded replica's
  pressure samples arrive exactly when the move count is a multiple of N^2
  and its energy is above the cutoff, each equal to the configuration's
  virial sum of -r dU/dr / (3V).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dos():
    d = json.load(open(os.path.join(GOLDEN, "ising4_dos.json")))
    return np.array(d["energies"], float), np.array([float(int(c)) for c in d["counts"]])


def test_ising_fractions_match_exact_volume_ratios():
    E, g = _dos()
    z = O.Zmc(O.make_params(system=O.ISING, n=4, ising_init=1), seed=5, min_T=0.2)
    rounds = 60000
    z.run(rounds)
    reps = z.replicas()
    assert len(reps) > 5
    r0, r1 = reps[0], reps[1]
    assert np.isinf(r0["max_energy"])
    assert r0["above_count"] + r0["below_count"] == rounds
    assert r1["above_count"] + r1["below_count"] == rounds * 16
    for k, r in enumerate(reps):
        n = r["above_count"] + r["below_count"]
        if n < 20000:
            continue
        inside = E < r["max_energy"]
        exact = g[inside & (E > r["cutoff_energy"])].sum() / g[inside].sum()
        frac = r["above_count"] / n
        tol = 4 * np.sqrt(max(exact * (1 - exact), 1e-4) / n) * (1 if k == 0 else 8)  # moves are correlated
        assert abs(frac - exact) <= tol + 1e-12, (k, r["max_energy"], r["cutoff_energy"], frac, exact)
    for a, b in zip(reps[1:], reps[2:]):
        assert b["max_energy"] == a["cutoff_energy"] and b["cutoff_energy"] < b["max_energy"]


def test_lj13_makes_replicas_cut_near_their_medians():
    p = O.make_params(system=O.LJ, n=13, radius=2.5, order_g=4)
    z = O.Zmc(p, seed=3, min_T=0.05)
    z.run(30000)
    reps = z.replicas()
    assert len(reps) >= 8
    for a, b in zip(reps[1:], reps[2:]):
        assert b["max_energy"] == a["cutoff_energy"] and b["cutoff_energy"] < b["max_energy"]
    fr = [r["above_count"] / (r["above_count"] + r["below_count"]) for r in reps[2:-2]
          if r["above_count"] + r["below_count"] > 5000]
    assert fr and all(0.2 < f < 0.8 for f in fr), fr
    for k in range(1, len(reps)):
        c = z.coords(k).reshape(3, -1)
        assert np.all(np.sum(c * c, axis=0) <= 2.5 ** 2 + 1e-12)
        assert abs(O.lj_energy_ld(z.coords(k)) - reps[k]["E"]) <= 1e-9 * max(1.0, abs(reps[k]["E"]))


# ---- WCA (wca.rs): randomization, sampling below ceilings, the pressure extra

WCA_RC2 = 2.0 ** (1.0 / 3.0)


def _wca_uniform_energies(seed, n, L, count):
    """Energies of `count` uniform configurations in the periodic box, from the
    WCA pair potential 4(r^-12 - r^-6) + 1 inside r < 2^(1/6) (numpy)."""
    g = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    out = []
    for _ in range(count // 32768):
        x = g.uniform(0.0, L, (32768, n, 3))
        d = x[:, :, None, :] - x[:, None, :, :]
        d -= L * np.round(d / L)
        r2 = (d ** 2).sum(-1)[:, iu[0], iu[1]]
        s6 = r2 ** -3.0
        out.append(np.where(r2 < WCA_RC2, 4.0 * (s6 * s6 - s6) + 1.0, 0.0).sum(1))
    return np.concatenate(out)


def _wca_virial(c, n, L):
    """Excess pressure of a configuration: sum over pairs of -r dU/dr / (3V)."""
    x = np.asarray(c).reshape(3, n).T
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)[np.triu_indices(n, 1)]
    s6 = r2 ** -3.0
    return np.where(r2 < WCA_RC2, 48.0 * s6 * s6 - 24.0 * s6, 0.0).sum() / (3.0 * L ** 3)


def test_wca_initial_ceilings_are_quantiles_of_uniform_configurations():
    # from_params' ceilings are the median and lower quartile of 2^15
    # randomizations: against numpy's uniform configurations, the CDF there
    # is 1/2 and 1/4 (a randomization in the wrong region, or without the
    # periodic images, moves both by far more than the sampling error)
    n, L = 8, 3.0
    E = np.sort(_wca_uniform_energies(1, n, L, 1 << 17))
    for seed in (3, 4):
        r = O.Zmc(O.make_params(system=O.WCA, n=n, box=L, order_g=4), seed=seed, min_T=0.05).replicas()
        assert np.isinf(r[0]["max_energy"]) and r[0]["cutoff_energy"] == r[1]["max_energy"]
        sig = np.sqrt(0.25 / (1 << 15) + 0.25 / len(E))
        for q, e in ((0.5, r[1]["max_energy"]), (0.25, r[1]["cutoff_energy"])):
            cdf = np.searchsorted(E, e) / len(E)
            assert abs(cdf - q) < 5 * sig, (seed, q, cdf)


def test_wca_replicas_cut_at_medians_of_their_uniform_bands():
    # a bounded replica samples uniformly below its ceiling: each new cutoff
    # is the median of the configurations under the ceiling, so numpy's
    # uniform configurations with E < max put half their weight above it
    n, L = 8, 3.0
    z = O.Zmc(O.make_params(system=O.WCA, n=n, box=L, order_g=4), seed=7, min_T=0.05)
    z.run(4000)
    reps = z.replicas()
    assert len(reps) >= 5
    E = _wca_uniform_energies(2, n, L, 1 << 18)
    checked = 0
    for r in reps[1:]:
        inside = E < r["max_energy"]
        if inside.sum() < 4000:
            continue
        frac = (E[inside] > r["cutoff_energy"]).mean()
        assert abs(frac - 0.5) < 0.08, (r["max_energy"], r["cutoff_energy"], frac)
        checked += 1
    assert checked >= 3


def test_wca_pressure_collected_every_n2_moves_above_the_cutoff():
    # the unbounded replica randomizes once a round and, while above its
    # cutoff (so no exchange can take its system), collects the virial
    # pressure when the run's move count is a multiple of N^2
    n, L = 8, 3.0
    z = O.Zmc(O.make_params(system=O.WCA, n=n, box=L, order_g=4), seed=5, min_T=0.05)
    hits = 0
    for _ in range(400):
        before = z.replicas()[0]
        mv = z.moves()
        z.run(1)
        after = z.replicas()[0]
        due = mv % (n * n) == 0 and after["E"] > after["cutoff_energy"]
        assert after["pressure_count"] - before["pressure_count"] == int(due)
        if due:
            p = _wca_virial(z.coords(0), n, L)
            want = before["pressure_total"] + p  # totals reach 1e12: compare them, not their difference
            assert after["pressure_total"] == pytest.approx(want, rel=1e-12, abs=1e-9 * abs(p))
            hits += 1
    assert hits >= 10
