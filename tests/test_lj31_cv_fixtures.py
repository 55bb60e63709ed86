"""The LJ31 literature heat capacities the reference holds, and the
transform its analysis applies (final_heat_capacity.py:33-53): pinned by the
harmonic limit of a 31-atom cluster's configurational heat capacity,
(3N - 6)/2 = 43.5 k_B, and by the five sources agreeing with one another."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lj31_cv")
# file, whether the file holds the total heat capacity per atom
SOURCES = [("LJ31_Cv_Reference", True), ("LJ31_Cv_Reference_alt", False), ("LJ31_Cv_Reference_3", True),
           ("LJ31_Cv_Reference_4", True), ("tRem_Ref", False)]


def literature(name):
    """(T, configurational Cv) of one source, sorted by T."""
    per_atom = dict(SOURCES)[name]
    T, c = np.loadtxt(os.path.join(HERE, name + ".csv"), delimiter=",", unpack=True)
    cv = (c - 1.5) * 31 if per_atom else c
    o = np.argsort(T)
    return T[o], cv[o]


@pytest.mark.parametrize("name", [s for s, _ in SOURCES])
def test_low_temperature_limit_is_harmonic(name):
    T, cv = literature(name)
    assert abs(cv[0] - 43.5) < 0.05 * 43.5, (name, T[0], cv[0])


def test_sources_agree_below_the_melting_peak():
    grid = np.linspace(0.1, 0.28, 19)
    curves = np.array([np.interp(grid, *literature(s)) for s, _ in SOURCES])
    spread = (curves.max(0) - curves.min(0)) / curves.mean(0)
    assert spread.max() < 0.08, spread
    # the melting peak near T = 0.33
    for s, _ in SOURCES:
        T, cv = literature(s)
        assert 0.31 < T[np.argmax(cv)] < 0.35 and cv.max() > 110, s
