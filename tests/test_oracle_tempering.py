"""Pins of the oracle's parallel tempering (src/mc/tempering.rs):

- the ladder generator's 2^64-step jump against the jump matrix M^(2^64)
  of the xoroshiro128 linear engine (computed independently in
  test_oracle_rng.py's GF(2) algebra);
- canonical averages <E>(T) and Cv(T) of every replica of a 4x4 Ising
  ladder against the exact density of states;
- swap bookkeeping: every replica's sample count is its moves plus its
  swap attempts, and the energies of the ladder are a permutation of
  configurations that exist (each replica's E equals its system's).
"""
import json
import os

import numpy as np

import oracle as O
from test_oracle_rng import _matrix, _pow

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _apply(Mat, s):
    """Apply a GF(2) 128x128 matrix (rows as ints) to state (s0, s1)."""
    v = s[0] | (s[1] << 64)
    out = 0
    for i, row in enumerate(Mat):
        if bin(row & v).count("1") & 1:
            out |= 1 << i
    return [out & (2 ** 64 - 1), out >> 64]


def test_jump_is_two_to_the_64_steps():
    J = _pow(_matrix(), 1 << 64)
    for state in ([1, 0], [0x0123456789ABCDEF, 0xFEDCBA9876543210], O.RNG(99).state):
        assert O.rng_jump(list(state)) == _apply(J, list(state))


def test_ising_ladder_canonical_averages_are_exact():
    d = json.load(open(os.path.join(GOLDEN, "ising4_dos.json")))
    E = np.array(d["energies"], float)
    g = np.array([float(int(c)) for c in d["counts"]])
    T = [1.5, 2.0, 2.5, 3.0, 4.0]
    p = O.make_params(system=O.ISING, n=4, ising_init=1, move_param=0.05)
    pt = O.Tempering(p, T, seed=11, spins=O.ising_spins(4, 1, 11))
    pt.run(40000)
    for r, t in enumerate(T):
        s = pt.replica(r)
        count = s["accepted"] + s["rejected"] + s["accepted_swaps"] + s["rejected_swaps"]
        assert count == 40000 * 16 + s["accepted_swaps"] + s["rejected_swaps"]
        w = g * np.exp(-(E - E.min()) / t)
        w /= w.sum()
        u_exact = (E * w).sum()
        cv_exact = ((E - u_exact) ** 2 * w).sum() / t ** 2
        u = s["total_energy"] / count
        cv = (s["total_energy_squared"] / count - u * u) / t ** 2
        assert abs(u - u_exact) < 0.01 * abs(u_exact) + 0.05, (t, u, u_exact)
        assert abs(cv - cv_exact) < 0.08 * cv_exact, (t, cv, cv_exact)  # a fluctuation: noisier
    # the swap bookkeeping is symmetric within each accepted pair
    total = sum(pt.replica(r)["accepted_swaps"] for r in range(len(T)))
    assert total % 2 == 0


def test_lj_ladder_keeps_configurations():
    p = O.make_params(system=O.LJ, n=13, radius=2.5, move_param=0.1, order_g=4, energy_bin=0.01)
    c0 = O.lj_init(13, 2.5, 3)
    pt = O.Tempering(p, [0.05, 0.1, 0.2, 0.4], seed=3, coords=c0)
    pt.run(300)
    for r in range(4):
        s = pt.replica(r)
        c = pt.coords(r)
        assert abs(O.lj_energy_ld(c) - s["E"]) <= 1e-9 * abs(s["E"])
        assert np.all(np.sum(c.reshape(3, 13) ** 2, axis=0) <= 2.5 ** 2 + 1e-12)
