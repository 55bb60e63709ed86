"""GPU parity over the configurations' own horizons.

The short parity cases (test_gpu_parity.py) cover SAD's start-up phase.
These run whole configurations for 1e7 moves per walker -- the late phase of
SAD with its rare tF updates, windows of thousands of bins and, for the LJ31
precision configuration, about ten thousand 1024-move FP64 refreshes -- and
compare walkers sampled across the full launch with the oracle, bit for bit
(every EnergyMC field, every bin, the generator and the configuration).
The oracle runs the sampled walkers on parallel host threads.
"""
import time

import numpy as np
import pytest

import sadmc_amd as S
from conftest import compare_walker, oracle_params, run_oracles

pytestmark = pytest.mark.gpu

MOVES = 10_000_000


def _run_long(ctx_factory, sp, W, geom, sample, label, launches=20):
    seeds = np.arange(W, dtype=np.uint64)
    ctx = ctx_factory()
    sid = ctx.add_set(sp, seeds, geom=geom)
    ctx.init()
    t0 = time.perf_counter()
    for k in range(launches):
        ctx.run(MOVES // launches)
        print(f"{label}: launch {k + 1}/{launches} done at {time.perf_counter() - t0:.1f} s", flush=True)
    st, _ = ctx.download_raw(sid, with_bins=False)
    assert np.all(st["status"] == 0) and np.all(st["moves"] == MOVES)
    g = lambda w: 0.0 if geom is None else float(geom)
    orcs = run_oracles([oracle_params(sp, seeds[w], g(w)) for w in sample], MOVES)
    print(f"{label}: oracle walkers done at {time.perf_counter() - t0:.1f} s", flush=True)
    for w, orc in zip(sample, orcs):
        lab = f"{label} walker {w}"
        compare_walker(ctx.state(sid, w), ctx.bins(sid, w), orc, sp.method, label=lab)
        if sp.system == S.ISING:
            np.testing.assert_array_equal(ctx.spins(sid, w), orc.spins(), err_msg=lab)
        else:
            np.testing.assert_array_equal(ctx.coords(sid, w), orc.coords(), err_msg=lab)
    lens = st["len"][list(sample)]
    print(f"{label}: {len(sample)} walkers bit-identical after {MOVES} moves; bins {lens.min()}..{lens.max()}")
    return st


def test_lj31_precision_configuration_1e7_moves(gpu_ctx_factory):
    """LJ31 precision configuration (the benchmark workload): 16384 walkers,
    sphere 3.0, cube moves 0.1, SAD min_T 0.05, bins 0.01, max energy 0, FP32
    pair terms with FP64 sums and an FP64 refresh every 1024 moves, 8 lanes;
    16 walkers spread over the launch."""
    print("case long lj31 fp32 16384 walkers", flush=True)
    sp = S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024)
    W = 16384
    sample = np.linspace(0, W - 1, 16).astype(int)
    st = _run_long(gpu_ctx_factory, sp, W, 3.0, sample, "lj31 fp32")
    assert np.all(st["len"][sample] > 3000)  # the late phase: thousands of bins


def test_ising_8x8_configuration_1e7_moves(gpu_ctx_factory):
    """Ising 8x8 configuration: 4096 walkers, spins i.i.d. from each seed,
    bins of width 4, SAD min_T 1.0; 8 walkers spread over the launch."""
    print("case long ising 8x8 4096 walkers", flush=True)
    sp = S.make_params(system=S.ISING, n=8, ising_init=1, method=S.SAD, method_param=1.0)
    W = 4096
    _run_long(gpu_ctx_factory, sp, W, None, np.linspace(0, W - 1, 8).astype(int), "ising 8x8", launches=4)


def test_lj13_fp64_configuration_1e7_moves(gpu_ctx_factory):
    """LJ13 configuration: 4096 walkers, sphere 2.5, cube moves 0.1, SAD
    min_T 0.05, all FP64 with the reference's error-triggered refresh, bins
    0.01, max energy 0, 8 lanes; 8 walkers spread over the launch."""
    print("case long lj13 fp64 4096 walkers", flush=True)
    sp = S.make_params(system=S.LJ, n=13, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    W = 4096
    _run_long(gpu_ctx_factory, sp, W, 2.5, np.linspace(0, W - 1, 8).astype(int), "lj13 fp64", launches=10)


def test_batched_bench_composition_1e6_moves(gpu_ctx_factory):
    """The mixed batch's eight size classes (LJ13/19/31/38, WCA32/64/128/256
    at the autotuned lanes, FP64 refresh every 1024 moves) as one
    graph-captured launch, replayed for 1e6 moves: about a thousand refreshes
    and SAD past its start-up; the first and last walker of every class
    against the oracle, bit for bit."""
    from conftest import run_oracles
    from test_gpu_batched import _bench_sets

    print("case long batched composition", flush=True)
    moves, launches = 1_000_000, 4
    sets = _bench_sets()
    ctx = gpu_ctx_factory()
    ids = [ctx.add_set(sp, seeds, geom=geom) for sp, seeds, geom in sets]
    ctx.init()
    t0 = time.perf_counter()
    for k in range(launches):
        ctx.run_batched(moves // launches, graph=True)
        ctx.sync()
        print(f"batched: replay {k + 1}/{launches} done at {time.perf_counter() - t0:.1f} s", flush=True)
    picks = [(sid, sp, seeds, geom, w) for sid, (sp, seeds, geom) in zip(ids, sets) for w in (0, len(seeds) - 1)]
    orcs = run_oracles([oracle_params(sp, seeds[w], float(geom[w])) for _, sp, seeds, geom, w in picks], moves)
    print(f"batched: oracle walkers done at {time.perf_counter() - t0:.1f} s", flush=True)
    for (sid, sp, seeds, geom, w), orc in zip(picks, orcs):
        label = f"{'LJ' if sp.system == S.LJ else 'WCA'}{sp.n} G{sp.lanes} walker {w}"
        st = ctx.state(sid, w)
        assert st["status"] == 0 and st["moves"] == moves, label
        compare_walker(st, ctx.bins(sid, w), orc, sp.method, label=label)
        np.testing.assert_array_equal(ctx.coords(sid, w), orc.coords(), err_msg=label)


def test_wca64_fp64_1e7_moves(gpu_ctx_factory):
    """A WCA fluid of the mixed batch's classes on its own: N = 64 at packing
    fraction 0.4 (periodic box), cube moves 0.05, SAD min_T 0.1, bins 0.05,
    max energy 2.5 N, FP64 refresh every 1024 moves, 8 lanes; 64 walkers for
    1e7 moves, 4 of them against the oracle."""
    print("case long wca64 fp64 64 walkers", flush=True)
    sp = S.make_params(system=S.WCA, n=64, lanes=8, method=S.SAD, method_param=0.1, move_param=0.05,
                       energy_bin=0.05, has_max_e=1, max_e=160.0, refresh=1024)
    W = 64
    box = float(np.cbrt(64 / (6.0 * 0.4 / np.pi)))
    _run_long(gpu_ctx_factory, sp, W, box, np.array([0, 21, 42, 63]), "wca64 fp64", launches=10)


@pytest.mark.parametrize("kind", ["histogram", "linear"])
def test_binning_walker_lj31_2e6_moves(gpu_ctx_factory, kind):
    """The `binning` binary's walker on the LJ31 configuration (SAD min_T 0.05,
    bins 0.01, max energy 0, 8 lanes, FP64) for 2e6 moves in four launches:
    past the histogram's growth into the phase where the BinCounts trackers
    carry the statistics; 8 walkers against the oracle, every scalar,
    statistic and bin."""
    from test_gpu_binning import _run

    print(f"case long binning lj31 {kind}", flush=True)
    sp = S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0,
                       binning=S.BIN_HISTOGRAM if kind == "histogram" else S.BIN_LINEAR)
    _run(gpu_ctx_factory, sp, np.arange(8, dtype=np.uint64) + 40, 3.0, [500_000] * 4, f"binning {kind}")
