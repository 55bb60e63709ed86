"""S(E) and Cv(T) extraction (plotting/parse-binning.py:150-171,
plotting/final_heat_capacity.py:81-89) pinned by exact results:

- Kaufman's closed-form partition function of the periodic L x L Ising
  model (B. Kaufman, Phys. Rev. 76, 1232 (1949); the form of P. D. Beale,
  Phys. Rev. Lett. 76, 78 (1996)) against the transfer-matrix density of
  states in tests/golden -- an independent check of the fixture;
- `heat_capacity` on the exact g(E) against Cv = beta^2 d^2 ln Z / d beta^2
  from Kaufman's Z by finite differences;
- the extended SAD entropy of oracle walkers (8x8 Ising, SAD min_T = 1)
  against the exact Cv(T) above min_T.
"""
import json
import os

import numpy as np
import pytest

import analysis as A
import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def exact_dos(L):
    d = json.load(open(os.path.join(GOLDEN, f"ising{L}_dos.json")))
    return np.array(d["energies"], float), np.log(np.array([float(int(c)) for c in d["counts"]]))


def kaufman_lnz(L, beta):
    """ln Z of the L x L periodic Ising model (J = 1) at inverse temperature beta."""
    K = beta
    k = np.arange(2 * L)
    c = np.cosh(2 * K) / np.tanh(2 * K) - np.cos(np.pi * k / L)
    gam = np.arccosh(c)
    gam[0] = 2 * K + np.log(np.tanh(K))  # signed, as in the closed form
    odd, even = gam[1::2], gam[0::2]
    terms = [np.prod(2 * np.cosh(L * odd / 2)), np.prod(2 * np.sinh(L * odd / 2)),
             np.prod(2 * np.cosh(L * even / 2)), np.prod(2 * np.sinh(L * even / 2))]
    return np.log(0.5) + (L * L / 2) * np.log(2 * np.sinh(2 * K)) + np.log(sum(terms))


def dos_lnz(E, lng, beta):
    a = lng - beta * E
    m = a.max()
    return m + np.log(np.exp(a - m).sum())


@pytest.mark.parametrize("L", [4, 6, 8])
def test_kaufman_partition_function_equals_fixture(L):
    E, lng = exact_dos(L)
    for T in (0.8, 1.5, 2.0, 2.269, 2.5, 3.5, 8.0):
        assert abs(kaufman_lnz(L, 1 / T) - dos_lnz(E, lng, 1 / T)) < 1e-11 * L * L


def test_heat_capacity_of_exact_dos_equals_kaufman():
    E, lng = exact_dos(8)
    T = np.array([1.0, 1.8, 2.2, 2.3, 2.6, 3.0, 4.0])
    got = A.heat_capacity(T, E, lng)
    h = 1e-4
    want = []
    for t in T:
        b = 1 / t
        d2 = (kaufman_lnz(8, b + h) - 2 * kaufman_lnz(8, b) + kaufman_lnz(8, b - h)) / h ** 2
        want.append(b * b * d2)
    np.testing.assert_allclose(got, want, rtol=2e-6)
    # and the peak sits near the finite-size critical temperature
    assert 2.2 < T[np.argmax(got)] < 2.7


def test_heat_capacity_limits():
    # two levels 0 and 1 (Schottky anomaly): Cv = x^2 e^x / (1 + e^x)^2 with x = 1/T
    T = np.array([0.2, 0.5, 1.0, 3.0])
    x = 1 / T
    np.testing.assert_allclose(A.heat_capacity(T, [0.0, 1.0], [0.0, 0.0]), x * x * np.exp(x) / (1 + np.exp(x)) ** 2,
                               rtol=1e-12)
    # a single level has no fluctuations
    assert np.all(A.heat_capacity(T, [3.0], [0.0]) == 0)


def test_sad_walkers_give_exact_heat_capacity():
    E, lng = exact_dos(8)
    T = np.linspace(1.5, 5.0, 20)
    want = A.heat_capacity(T, E, lng)
    walkers = []
    for seed in range(4):
        w = O.Walker(O.make_params(system=O.ISING, n=8, method=O.SAD, method_param=1.0, seed=seed, ising_init=1))
        w.run(2_000_000)
        st = w.scalars()
        st = {"len": st["len"], "bins_min": st["bins_min"], "bins_width": st["bins_width"], "method": 0,
              "too_lo": st["too_lo"], "too_hi": st["too_hi"], "min_T": 1.0, "wl_gamma": 0.0, "status": 0}
        walkers.append((st, w.bins()))
    Eg, S, count = A.mean_entropy(walkers)
    got = A.heat_capacity(T, Eg, S)
    assert count.max() == 4 and np.all(count[(Eg <= 0) & (Eg > -120)] == 4)
    assert np.max(np.abs(got / want - 1)) < 0.03


def test_walker_entropy_extension_below_too_lo():
    """Below too_lo the entropy continues with slope 1/min_T (plus the
    histogram correction); with a flat histogram the slope is exact."""
    st = {"len": 10, "bins_min": 0.0, "bins_width": 1.0, "method": 0, "too_lo": 4.5, "too_hi": 7.5,
          "min_T": 0.5, "wl_gamma": 0.0}
    b = {"lnw": np.array([0, 0, 0, 0, 3.0, 5.0, 6.0, 6.5, 0, 0.0]), "histogram": np.full(10, 7)}
    E, S = A.walker_entropy(st, b)
    np.testing.assert_allclose(np.diff(S[:5]), 1 / 0.5)
    np.testing.assert_allclose(S[7:], S[7])  # flat above too_hi
    assert S.max() == 0.0
