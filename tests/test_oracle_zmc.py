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
"""
import json
import os

import numpy as np

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
