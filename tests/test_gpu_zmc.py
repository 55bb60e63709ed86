"""GPU energy-ceiling replicas (src/mc/energy_replicas.rs) against the
oracle, bit for bit: every replica record (ceilings, cutoffs, counters,
sums, lowest_max_energy, unique visitors, generator, translation scale,
its system's energy and error), every replica's configuration, the run's
own generator, move count and median estimator, through replica splits and
several launches.  Then the argument and capacity checks.
"""
import numpy as np
import pytest

import sadmc_amd as S

pytestmark = pytest.mark.gpu

FIELDS = ["max_energy", "cutoff_energy", "rejected_count", "accepted_count", "above_count", "below_count",
          "upwelling_count", "above_total", "below_total", "above_total_squared", "below_total_squared",
          "lowest_max_energy", "unique_visitors", "collecting_data", "translation_scale", "E", "error",
          "pressure_total", "pressure_count"]


def _oracle(sp, seed, geom, min_T, isbnb, coords=None, spins=None):
    import oracle as O
    op = O.make_params(system=sp.system, n=sp.n, move_param=sp.move_param,
                       radius=(geom or 0.0) if sp.system == S.LJ else 0.0,
                       box=(geom or 0.0) if sp.system == S.WCA else 0.0, rc2=sp.rc2,
                       order_g=sp.lanes if sp.system != S.ISING else 0, ising_init=sp.ising_init)
    return O.Zmc(op, seed=int(seed), min_T=min_T, isbnb=isbnb, coords=coords, spins=spins)


def _compare(ctx, zid, sp, seeds, geom, rounds, chunks, min_T, isbnb, min_replicas):
    for c in chunks:
        ctx.zmc_run(zid, c)
    ctx.zmc_run(zid, rounds - sum(chunks))
    assert ctx.last_run_ms() > 0.0  # the last launch, timed by device events
    for s, seed in enumerate(seeds):
        orc = _oracle(sp, seed, geom, min_T, isbnb)
        orc.run(rounds)
        want = orc.replicas()
        got = ctx.zmc_replicas(zid, s)
        label = f"run {s} seed {seed}"
        assert len(got) == len(want), (label, len(got), len(want))
        assert len(got) >= min_replicas, (label, len(got))
        for k, (g, o) in enumerate(zip(got, want)):
            assert [int(x) for x in g["rng"]] == o["rng"], f"{label} replica {k} rng"
            for f in FIELDS:
                assert np.asarray(g[f]).tobytes() == np.asarray(o[f], dtype=g[f].dtype).tobytes(), \
                    f"{label} replica {k} {f}: gpu {g[f]!r} oracle {o[f]!r}"
            if sp.system == S.ISING:
                np.testing.assert_array_equal(ctx.zmc_spins(zid, s, k), orc.spins(k), err_msg=f"{label} {k}")
            else:
                np.testing.assert_array_equal(ctx.zmc_coords(zid, s, k), orc.coords(k), err_msg=f"{label} {k}")
        rng, moves, med = ctx.zmc_state(zid, s)
        assert [int(x) for x in rng] == orc.rng(), label
        assert moves == orc.moves(), label
        np.testing.assert_array_equal(med, orc.median(), err_msg=label)


def test_ising_runs_bit_exact(gpu_ctx_factory):
    print("case zmc ising 4x4 S=3", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=4, ising_init=1)
    seeds = [5, 6, 7]
    zid = ctx.zmc_add(sp, seeds, min_T=0.2, isbnb=64, max_replicas=48)
    _compare(ctx, zid, sp, seeds, None, 12000, [3000, 1], 0.2, 64, 10)


def test_lj13_runs_bit_exact(gpu_ctx_factory):
    print("case zmc lj13 S=2 G=4", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.LJ, n=13, lanes=4, move_param=0.1)
    seeds = [3, 4]
    zid = ctx.zmc_add(sp, seeds, geom=2.5, min_T=0.05, isbnb=64, max_replicas=32)
    _compare(ctx, zid, sp, seeds, 2.5, 12000, [5000], 0.05, 64, 6)


def test_lj31_eight_lanes(gpu_ctx_factory):
    print("case zmc lj31 S=1 G=8", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.LJ, n=31, lanes=8, move_param=0.1)
    zid = ctx.zmc_add(sp, [9], geom=3.0, min_T=0.05, isbnb=16, max_replicas=48)
    _compare(ctx, zid, sp, [9], 3.0, 3000, [], 0.05, 16, 4)


def test_lj13_more_slot_lanes_than_a_workgroup(gpu_ctx_factory):
    """48 slots of 16 lanes (768 lanes' worth) on one wave: its four lane
    groups take the slots in turn, so only LDS bounds the replica count."""
    print("case zmc lj13 S=2 G=16 R=48", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.LJ, n=13, lanes=16, move_param=0.1)
    seeds = [3, 4]
    zid = ctx.zmc_add(sp, seeds, geom=2.5, min_T=0.05, isbnb=64, max_replicas=48)
    _compare(ctx, zid, sp, seeds, 2.5, 12000, [5000], 0.05, 64, 6)


def test_wca8_runs_bit_exact(gpu_ctx_factory):
    # WCA's randomization (atoms added back through the error estimate) and
    # the pressure collected above each cutoff every N^2 moves
    print("case zmc wca8 S=2 G=4", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.WCA, n=8, lanes=4, move_param=0.1)
    seeds = [7, 8]
    zid = ctx.zmc_add(sp, seeds, geom=3.0, min_T=0.05, isbnb=64, max_replicas=32)
    _compare(ctx, zid, sp, seeds, 3.0, 6000, [2000], 0.05, 64, 5)


def test_wca32_eight_lanes(gpu_ctx_factory):
    print("case zmc wca32 S=1 G=8", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.WCA, n=32, lanes=8, move_param=0.1)
    zid = ctx.zmc_add(sp, [11], geom=4.5, min_T=0.05, isbnb=16, max_replicas=40)
    _compare(ctx, zid, sp, [11], 4.5, 1500, [], 0.05, 16, 4)


def test_zmc_capacity_and_arguments(gpu_ctx_factory):
    print("case zmc capacity and arguments", flush=True)
    ctx = gpu_ctx_factory()
    with pytest.raises(S.SadmcError, match="cutoff wide"):
        ctx.zmc_add(S.make_params(system=S.WCA, n=32), [1], geom=1.0)
    # a run's systems must fit the 64 KiB of LDS its wave uses
    with pytest.raises(S.SadmcError, match="64 KiB"):
        ctx.zmc_add(S.make_params(system=S.LJ, n=38, lanes=16), [1], geom=3.2, max_replicas=64)
    zid = ctx.zmc_add(S.make_params(system=S.ISING, n=4, ising_init=1), [5], max_replicas=4)
    with pytest.raises(S.SadmcError, match="split beyond 4"):
        ctx.zmc_run(zid, 20000)
    assert len(ctx.zmc_replicas(zid, 0)) == 4
