"""Pins of the oracle's systems (Ising, LJ, WCA) against the reference's own
invariant tests and against mathematics the reference does not state.

- ising.rs:125-147 energy_works: energy() == compute_energy() exactly after
  every flip, N in {2, 3, 10, 15} (137 and 150 at fewer flips).
- Brute-force and transfer-matrix densities of states (fixtures in
  tests/golden, produced by tests/golden/make_ising_dos.py).
- LJ pair potential minimum at r = 2^(1/6) with V = -1; the LJ13
  icosahedral global minimum -44.326801 (Wales & Doye) reached by relaxing
  the oracle's energy from an icosahedron.
- lj.rs:380-427 energy_is_right pattern: incremental energy within the
  reference's error bound of a recompute, and per-move energies within
  1e-9 relative of a long-double recompute.
- WCA: minimum-image energy equals the reference's cell-list energy
  (optcell.rs) to rounding; E >= 0 (wca.rs:234-236).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ Ising
@pytest.mark.parametrize("n,flips", [(2, 10000), (3, 10000), (10, 10000), (15, 10000), (137, 300), (150, 300)])
def test_ising_energy_works(n, flips):
    # the reference seeds the move RNG with 10137 as well (ising.rs:133)
    w = O.Walker(O.make_params(system=O.ISING, n=n, method=O.SAD, method_param=1.0, seed=10137))
    assert w.energy() == O.ising_energy(n, w.spins())
    for _ in range(flips):
        w.plan_move(0.0)
        w.confirm()
        assert w.energy() == O.ising_energy(n, w.spins())


def test_ising_reference_initial_spins_rule():
    # ising.rs:45-49: low bit of successive draws of seed_from_u64(10137)
    r = O.RNG(10137)
    want = np.array([1 if r.next_u64() & 1 else -1 for _ in range(64)], np.int8)
    np.testing.assert_array_equal(O.ising_spins(8, 0, 12345), want)


def _brute_dos(L):
    n = L * L
    s = np.where((np.arange(1 << n)[:, None] >> np.arange(n)) & 1, 1, -1).astype(np.int8)
    E = np.array([O.ising_energy(L, row) for row in s])
    vals, cnt = np.unique(E, return_counts=True)
    return dict(zip(vals.astype(int).tolist(), cnt.tolist()))


def test_ising_dos_fixtures_match_brute_force_and_sum_rules():
    bf = _brute_dos(4)  # 65536 configurations through the oracle's energy
    d4 = json.load(open(os.path.join(GOLDEN, "ising4_dos.json")))
    assert dict(zip(d4["energies"], map(int, d4["counts"]))) == bf
    for L in (6, 8):
        d = json.load(open(os.path.join(GOLDEN, f"ising{L}_dos.json")))
        counts = dict(zip(d["energies"], map(int, d["counts"])))
        assert sum(counts.values()) == 2 ** (L * L)
        assert counts[-2 * L * L] == 2            # the two checkerboards
        assert counts[-2 * L * L + 8] == 2 * L * L  # one flipped spin breaks four bonds
        assert all(counts[e] == counts[-e] for e in counts)


# --------------------------------------------------------------------- LJ
def test_lj_pair_minimum():
    r2 = 2.0 ** (1.0 / 3.0)
    assert abs(O.lib().or_lj_pair(r2) + 1.0) < 1e-15
    h = 1e-4
    left, right = O.lib().or_lj_pair(r2 * (1 - h)), O.lib().or_lj_pair(r2 * (1 + h))
    assert left > -1.0 and right > -1.0


def _icosahedron():
    g = (1 + 5 ** 0.5) / 2
    v = []
    for a in (-1, 1):
        for b in (-g, g):
            v += [(0, a, b), (a, b, 0), (b, 0, a)]
    v = np.array(v, float)
    v /= np.linalg.norm(v[0])
    return np.vstack([np.zeros(3), v])


def test_lj13_global_minimum():
    from scipy.optimize import minimize
    x0 = _icosahedron() * 1.1

    def f(flat):
        c = flat.reshape(13, 3).T.copy().ravel()
        return O.lj_energy(c)

    res = minimize(f, x0.ravel(), method="L-BFGS-B", options=dict(ftol=1e-15, gtol=1e-10, maxiter=5000))
    assert abs(res.fun - (-44.326801)) < 2e-6


@pytest.mark.parametrize("n,radius", [(13, 2.5), (31, 3.0), (38, 3.2)])
def test_lj_init_rule(n, radius):
    for seed in range(5):
        c = O.lj_init(n, radius, seed).reshape(3, n)
        assert (np.sum(c * c, 0) < radius * radius).all()
        d = c[:, :, None] - c[:, None, :]
        r2 = (d * d).sum(0) + np.eye(n) * 10
        assert r2.min() >= 0.8 * 0.8


@pytest.mark.parametrize("g", [0, 4, 8])
def test_lj_incremental_energy_per_move(g):
    """Per-move energies within 1e-9 relative of a long-double recompute
    (the bound the GPU path is held to), and the reference's own bound
    |E| 1e-14 N^2 after many moves (lj.rs:249-261)."""
    p = O.make_params(system=O.LJ, n=31, radius=3.0, method=O.SAD, method_param=0.05, move_param=0.1,
                      energy_bin=0.01, has_max_e=1, max_e=0.0, order_g=g, seed=4)
    w = O.Walker(p)
    worst = 0.0
    for _ in range(1500):
        w.run(1)
        e_ref = O.lj_energy_ld(w.coords())
        worst = max(worst, abs(w.scalars()["E"] - e_ref) / abs(e_ref))
    assert worst < 1e-9
    w.run(50000)
    s = w.scalars()
    assert abs(s["E"] - O.lj_energy_ld(w.coords())) <= abs(s["E"]) * 1e-14 * 31 * 31


def test_lj_fp32_pairs_error_reported():
    """FP32 pair terms drift between FP64 refreshes; the drift stays well
    inside 1e-4 relative and every refresh restores FP64 accuracy."""
    p = O.make_params(system=O.LJ, n=31, radius=3.0, method=O.SAD, method_param=0.05, move_param=0.1,
                      energy_bin=0.01, has_max_e=1, max_e=0.0, precision=1, refresh=1024, order_g=8, seed=2)
    w = O.Walker(p)
    w.run(20000)
    w.run(1024 - (20000 % 1024))          # land exactly on a refresh
    assert abs(w.scalars()["E"] - O.lj_energy_ld(w.coords())) <= 1e-12 * abs(w.scalars()["E"])
    w.run(1000)
    assert abs(w.scalars()["E"] - O.lj_energy_ld(w.coords())) <= 1e-4 * abs(w.scalars()["E"])


def test_lj_fp32_pair_term_is_binary32():
    """The precision variant's pair term is computed in binary32 (numpy's
    float32 arithmetic, no extended precision or contraction), and lies
    within its rounding bound of the binary64 term."""
    rng = np.random.default_rng(7)
    r2 = rng.uniform(0.6, 9.0, 20000).astype(np.float32)
    got = np.array([O.lib().or_lj_pair_f32(float(x)) for x in r2], np.float32)
    s = np.float32(1.0) / r2
    s2 = s * s
    s4 = s2 * s2
    s6 = s2 * s4
    s3 = s * s2
    want = np.float32(4.0) * (s6 - s3)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    s64 = 1.0 / r2.astype(np.float64)
    exact = 4.0 * (s64 ** 6 - s64 ** 3)
    bound = 4.0 * (s64 ** 6 + s64 ** 3) * 2.0 ** -20
    assert np.all(np.abs(got.astype(np.float64) - exact) <= bound)
    assert np.count_nonzero(got.astype(np.float64) != exact) > len(r2) // 2


def test_lj_fp32_refresh_past_the_window_edge_grows_the_bins():
    """LJ31 walker 83142 (4 lanes) drifts to an FP32-summed energy inside the
    window's top bin at move 2047 whose FP64 recompute lies above that bin;
    the refresh at move 2048 grows the bins to the refreshed energy, as a
    proposed energy would, and the walker goes on."""
    p = O.make_params(system=O.LJ, n=31, radius=3.0, method=O.SAD, method_param=0.05, move_param=0.1,
                      energy_bin=0.01, has_max_e=1, max_e=0.0, precision=1, refresh=1024, order_g=4, seed=83142)
    w = O.Walker(p)
    assert w.run(2047) == 0
    s = w.scalars()
    top = s["bins_min"] + s["bins_width"] * len(w.bins()["lnw"])
    e64 = O.lj_energy_ld(w.coords())
    assert s["E"] < top <= e64 < 0.0
    n0 = len(w.bins()["lnw"])
    assert w.run(1) == 0
    s = w.scalars()
    assert s["status"] == 0 and len(w.bins()["lnw"]) == n0 + 1
    assert s["bins_min"] + s["bins_width"] * (n0 + 1) > e64
    assert w.run(3000) == 0 and w.scalars()["status"] == 0


@pytest.mark.parametrize("seed", [4, 5])
def test_lj_fp32_refresh_below_the_window_grows_the_front(seed):
    """A walker at move 1000 given the positions of its own state at move
    200000, which lie below every bin it has: the refresh at move 1024
    recomputes the energy from those positions and the bins grow at the
    front to reach it; the walker goes on."""
    p = O.make_params(system=O.LJ, n=31, radius=3.0, method=O.SAD, method_param=0.05, move_param=0.1,
                      energy_bin=0.01, has_max_e=1, max_e=0.0, precision=1, refresh=1024, order_g=8, seed=seed)
    deep = O.Walker(p)
    deep.run(200000)
    w = O.Walker(p)
    w.run(1000)
    s0 = w.scalars()
    assert O.lj_energy_ld(deep.coords()) < s0["bins_min"]
    n0 = len(w.bins()["lnw"])
    w.load_state(s0, w.bins(), deep.coords())
    assert w.run(24) == 0
    s = w.scalars()
    assert s["status"] == 0 and s["bins_min"] <= s["E"] < s0["bins_min"]
    # refreshed at the start of move 1024; that move's FP32 difference follows
    assert abs(s["E"] - O.lj_energy_ld(w.coords())) <= 1e-6 * abs(s["E"])
    assert len(w.bins()["lnw"]) >= n0 + round((s0["bins_min"] - s["bins_min"]) / 0.01) > n0
    assert w.run(2000) == 0 and w.scalars()["status"] == 0


def test_lj_orders_agree():
    for seed in range(3):
        c = O.lj_init(38, 3.2, seed)
        e0 = O.lj_energy(c, 0)
        for g in (2, 4, 8, 16, 32, 64):
            assert abs(O.lj_energy(c, g) - e0) <= 1e-13 * max(1.0, abs(e0))
        assert abs(O.lj_energy_ld(c) - e0) <= 1e-12 * max(1.0, abs(e0))


# -------------------------------------------------------------------- WCA
RC2 = 2.0 ** (1.0 / 3.0)


@pytest.mark.parametrize("n", [32, 64, 128, 256])
def test_wca_celllist_equals_minimum_image(n):
    for eta in (0.3, 0.4, 0.5):
        box = np.cbrt(n / (6 * eta / np.pi))
        p = O.make_params(system=O.WCA, n=n, box=box, rc2=RC2, method=O.CANONICAL, method_param=2.0,
                          move_param=0.3, energy_bin=0.05, seed=3)
        w = O.Walker(p)
        w.run(3000)  # disorder the lattice
        c = w.coords()
        e_cell = O.wca_energy_celllist(c, box, RC2)
        e_mi = O.wca_energy(c, box, RC2, 0)
        e_ld = O.wca_energy_ld(c, box, RC2)
        assert e_cell >= 0.0
        assert abs(e_cell - e_mi) <= 1e-12 * max(1.0, e_ld)
        assert abs(e_ld - e_mi) <= 1e-12 * max(1.0, e_ld)
        for g in (8, 16, 64):
            assert abs(O.wca_energy(c, box, RC2, g) - e_ld) <= 1e-12 * max(1.0, e_ld)


def test_wca_energy_is_right():
    """wca.rs:647-694: incremental energy within the reference's bound."""
    for g in (0, 16):
        box = np.cbrt(64 / (6 * 0.45 / np.pi))
        p = O.make_params(system=O.WCA, n=64, box=box, rc2=RC2, method=O.SAD, method_param=0.1, move_param=0.05,
                          energy_bin=0.05, has_max_e=1, max_e=160.0, order_g=g, seed=8, collect_pressure=1)
        w = O.Walker(p)
        for _ in range(25):
            w.run(200)
            s = w.scalars()
            e_ld = O.wca_energy_ld(w.coords(), box, RC2)
            assert abs(s["E"] - e_ld) <= max(abs(e_ld) * 1e-13 * 64 * 64, 1e-12)
        b = w.bins()
        assert s["has_pressure"] == 1 and b["pressure_count"].sum() == s["moves"] // (64 * 64)
