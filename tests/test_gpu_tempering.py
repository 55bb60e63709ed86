"""GPU parallel tempering (src/mc/tempering.rs) against the oracle's ladder,
bit for bit: every replica's generator, energy, error estimate, sums of E
and E^2, move and swap counters, configuration, and the ladder generator,
after several launches.  Then the device ladders' canonical averages on a
4x4 Ising lattice against the exact density of states, and the argument
checks of the entry points.
"""
import json
import os

import numpy as np
import pytest

import sadmc_amd as S

pytestmark = pytest.mark.gpu

FIELDS = ["E", "error", "T", "total_energy", "total_energy_squared", "accepted", "rejected", "ignored",
          "accepted_swaps", "rejected_swaps"]


def _oracle_ladder(sp, temps, seed, geom, coords, spins, steps):
    import oracle as O
    op = O.make_params(system=sp.system, n=sp.n, move_param=sp.move_param, rc2=sp.rc2,
                       radius=geom if sp.system == S.LJ else 0.0, box=geom if sp.system == S.WCA else 0.0,
                       order_g=sp.lanes if sp.system != S.ISING else 0, ising_init=sp.ising_init, seed=int(seed))
    return O.Tempering(op, temps, seed=int(seed), coords=coords, spins=spins, canonical_steps=steps)


def _compare(ctx, pt, sp, temps, seeds, geom, rounds, chunks, steps=1, init_from_seed=False):
    import oracle as O
    ctx.pt_run(pt, rounds - sum(chunks))
    for c in chunks:
        ctx.pt_run(pt, c)
    reps = ctx.pt_replicas(pt)
    lrng = ctx.pt_ladder_rng(pt)
    for l, seed in enumerate(seeds):
        if sp.system == S.ISING:
            sp0 = None if init_from_seed else S.init_spins(sp.n, sp.ising_init, int(seed))
            orc = _oracle_ladder(sp, temps, seed, 0.0, None, sp0, steps)
        else:
            c0 = S.init_coords(sp.system, sp.n, geom, int(seed))
            orc = _oracle_ladder(sp, temps, seed, geom, c0, None, steps)
        orc.run(rounds)
        assert [int(x) for x in lrng[l]] == orc.rng(), f"ladder {l} generator"
        for r in range(len(temps)):
            o = orc.replica(r)
            g = reps[l, r]
            label = f"ladder {l} seed {seed} replica {r}"
            assert [int(x) for x in g["rng"]] == list(o["rng"]), label
            for f in FIELDS:
                assert g[f] == o[f], f"{label} {f}: gpu {g[f]!r} oracle {o[f]!r}"
            if sp.system == S.ISING:
                np.testing.assert_array_equal(ctx.pt_spins(pt, l, r), orc.spins(r), err_msg=label)
            else:
                np.testing.assert_array_equal(ctx.pt_coords(pt, l, r), orc.coords(r), err_msg=label)
        assert sum(int(reps[l, r]["accepted_swaps"]) for r in range(len(temps))) > 0


def test_ising_ladders_bit_exact(gpu_ctx_factory):
    print("case pt ising 4x4 R=5 L=3", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=4, ising_init=1)
    temps = [1.5, 2.0, 2.5, 3.0, 4.0]
    seeds = [11, 12, 13]
    pt = ctx.pt_add(sp, temps, seeds)
    _compare(ctx, pt, sp, temps, seeds, None, 3000, [1000, 1], init_from_seed=True)


def test_ising_8x8_sixteen_replicas(gpu_ctx_factory):
    print("case pt ising 8x8 R=16", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=8, ising_init=1)
    temps = list(np.linspace(1.6, 3.5, 16))
    seeds = [5, 6]
    spins = np.concatenate([S.init_spins(8, 1, s) for s in seeds])
    pt = ctx.pt_add(sp, temps, seeds, spins=spins, canonical_steps=2)
    _compare(ctx, pt, sp, temps, seeds, None, 300, [100], steps=2)


def test_lj13_ladders_bit_exact(gpu_ctx_factory):
    print("case pt lj13 R=8 G=4 L=2", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.LJ, n=13, move_param=0.1, lanes=4)
    temps = list(0.05 * 1.4 ** np.arange(8))
    seeds = [3, 4]
    coords = np.concatenate([S.init_coords(S.LJ, 13, 2.5, s) for s in seeds])
    pt = ctx.pt_add(sp, temps, seeds, geom=2.5, coords=coords)
    _compare(ctx, pt, sp, temps, seeds, 2.5, 600, [200, 7])


def test_lj31_wide_ladder(gpu_ctx_factory):
    # 64 replicas x 8 lanes: one 512-lane workgroup per ladder
    print("case pt lj31 R=64 G=8", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.LJ, n=31, move_param=0.08, lanes=8)
    temps = list(0.05 * 1.05 ** np.arange(64))
    seeds = [21]
    coords = S.init_coords(S.LJ, 31, 3.0, 21)
    pt = ctx.pt_add(sp, temps, seeds, geom=3.0, coords=coords)
    _compare(ctx, pt, sp, temps, seeds, 3.0, 40, [10])


def test_wca_ladders_bit_exact(gpu_ctx_factory):
    print("case pt wca32 R=4 G=8", flush=True)
    ctx = gpu_ctx_factory()
    L = (32 / 0.6) ** (1.0 / 3.0)
    sp = S.make_params(system=S.WCA, n=32, move_param=0.1, lanes=8)
    temps = [0.5, 0.8, 1.2, 2.0]
    seeds = [7, 8]
    coords = np.concatenate([S.init_coords(S.WCA, 32, L, s) for s in seeds])
    pt = ctx.pt_add(sp, temps, seeds, geom=L, coords=coords, canonical_steps=2)
    _compare(ctx, pt, sp, temps, seeds, L, 150, [50], steps=2)


def test_ising_ladders_match_exact_canonical_averages(gpu_ctx_factory):
    """<E>(T) and Cv(T) of 64 device ladders against the exact 4x4 density of states."""
    print("case pt ising 4x4 statistics L=64", flush=True)
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ising4_dos.json")))
    E = np.array(d["energies"], float)
    g = np.array([float(int(c)) for c in d["counts"]])
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=4, ising_init=1)
    temps = [1.5, 2.0, 2.5, 3.0, 4.0]
    rounds = 20000
    pt = ctx.pt_add(sp, temps, np.arange(64, dtype=np.uint64) + 100)
    ctx.pt_run(pt, rounds)
    assert ctx.last_run_ms() > 0.0  # the launch is timed by device events
    reps = ctx.pt_replicas(pt)
    for r, t in enumerate(temps):
        col = reps[:, r]
        count = (col["accepted"] + col["rejected"] + col["accepted_swaps"] + col["rejected_swaps"]).astype(float)
        assert np.all(count == rounds * 16 + col["accepted_swaps"] + col["rejected_swaps"])
        u = col["total_energy"].sum() / count.sum()
        cv = (col["total_energy_squared"].sum() / count.sum() - u * u) / t ** 2
        w = g * np.exp(-(E - E.min()) / t)
        w /= w.sum()
        u_exact = (E * w).sum()
        cv_exact = ((E - u_exact) ** 2 * w).sum() / t ** 2
        assert abs(u - u_exact) < 0.003 * abs(u_exact) + 0.01, (t, u, u_exact)
        assert abs(cv - cv_exact) < 0.02 * cv_exact, (t, cv, cv_exact)


def test_tempering_argument_errors(gpu_ctx_factory):
    print("case pt argument errors", flush=True)
    ctx = gpu_ctx_factory()
    lj = S.make_params(system=S.LJ, n=13, move_param=0.1, lanes=4)
    with pytest.raises(S.SadmcError, match="temperatures"):
        ctx.pt_add(lj, [0.1, 0.0], [1], geom=2.5)
    with pytest.raises(S.SadmcError, match="512"):
        ctx.pt_add(lj, np.full(129, 0.1), [1], geom=2.5)
    with pytest.raises(S.SadmcError, match="FP64"):
        ctx.pt_add(S.make_params(system=S.LJ, n=13, move_param=0.1, lanes=4, refresh=1024), [0.1], [1], geom=2.5)
    with pytest.raises(S.SadmcError, match="geometry"):
        ctx.pt_add(lj, [0.1], [1])
    pt = ctx.pt_add(S.make_params(system=S.ISING, n=4), [2.0, 3.0], [1])
    with pytest.raises(S.SadmcError, match="coordinates"):
        ctx.pt_coords(pt, 0, 0)
    with pytest.raises(S.SadmcError, match="out of range"):
        ctx.pt_spins(pt, 0, 2)
