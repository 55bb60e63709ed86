"""GPU edge cases: the widest lane groups and largest systems of the
workloads, the bin window running out (reported, never silent), argument and
state errors of the C ABI, and the FP32-pair energy drift bound."""
import numpy as np
import pytest

import oracle as O
import sadmc_amd as S
from conftest import compare_walker, oracle_params

pytestmark = pytest.mark.gpu


def test_wca256_with_64_lanes_bit_exact(gpu_ctx_factory):
    """N = 256 over a full wavefront (the cross-row shuffles of the group sum)."""
    print("case wca256 g64", flush=True)
    n, eta = 256, 0.45
    box = float(np.cbrt(n / (6.0 * eta / np.pi)))
    sp = S.make_params(system=S.WCA, n=n, lanes=64, method=S.SAD, method_param=0.1, move_param=0.05,
                       energy_bin=0.05, has_max_e=1, max_e=2.5 * n)
    ctx = gpu_ctx_factory()
    seeds = np.array([11, 12], dtype=np.uint64)
    sid = ctx.add_set(sp, seeds, geom=box)
    ctx.init()
    ctx.run(1500)
    for w, seed in enumerate(seeds):
        orc = O.Walker(oracle_params(sp, seed, box))
        orc.run(1500)
        compare_walker(ctx.state(sid, w), ctx.bins(sid, w), orc, sp.method, label=f"wca256 walker {w}")
        np.testing.assert_array_equal(ctx.coords(sid, w), orc.coords())


def test_lj38_with_64_lanes_bit_exact(gpu_ctx_factory):
    print("case lj38 g64", flush=True)
    sp = S.make_params(system=S.LJ, n=38, lanes=64, method=S.SAD, method_param=0.03, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    ctx = gpu_ctx_factory()
    seeds = np.array([5, 6, 7], dtype=np.uint64)
    sid = ctx.add_set(sp, seeds, geom=3.2)
    ctx.init()
    ctx.run(3000)
    for w, seed in enumerate(seeds):
        orc = O.Walker(oracle_params(sp, seed, 3.2))
        orc.run(3000)
        compare_walker(ctx.state(sid, w), ctx.bins(sid, w), orc, sp.method, label=f"lj38 g64 walker {w}")


def test_bin_window_overflow_is_reported(gpu_ctx_factory):
    """A window too narrow for the walk stops the walker with a capacity
    status, which sync reports; the walker's state up to that move is the
    oracle's."""
    print("case capacity", flush=True)
    sp = S.make_params(system=S.ISING, n=8, method=S.SAD, method_param=1.0, ising_init=1,
                       e_lo_bound=-40.0, e_hi_bound=40.0)
    ctx = gpu_ctx_factory()
    sid = ctx.add_set(sp, np.arange(4, dtype=np.uint64))
    ctx.init()
    with pytest.raises(S.SadmcError) as e:
        ctx.run(200000)
    assert e.value.code == -3
    st, _ = ctx.download_raw(sid, with_bins=False)
    assert np.all(st["status"] != 0)
    for w in range(4):
        # the move that left the window counted but did not complete
        stopped = int(st[w]["moves"])
        orc = O.Walker(oracle_params(sp, w))
        orc.run(stopped - 1)
        o = orc.scalars()
        assert st[w]["E"] == o["E"]
        assert st[w]["bins_min"] == o["bins_min"] and st[w]["len"] == o["len"]
        np.testing.assert_array_equal(ctx.bins(sid, w)["lnw"], orc.bins()["lnw"])
        np.testing.assert_array_equal(ctx.bins(sid, w)["histogram"], orc.bins()["histogram"])


def test_abi_errors(gpu_ctx_factory):
    print("case abi errors", flush=True)
    ctx = gpu_ctx_factory()
    with pytest.raises(S.SadmcError) as e:
        ctx.init()  # no sets
    assert e.value.code == -6
    sp = S.make_params(system=S.LJ, n=13, method=S.SAD, method_param=0.05, energy_bin=0.01, has_max_e=1, max_e=0.0)
    with pytest.raises(S.SadmcError) as e:
        ctx.add_set(sp, np.zeros(0, dtype=np.uint64), geom=2.5)
    assert e.value.code == -1
    with pytest.raises(S.SadmcError) as e:
        ctx.add_set(S.make_params(system=S.LJ, n=13, lanes=3, method=S.SAD, method_param=0.05), np.arange(2),
                    geom=2.5)
    assert e.value.code == -1
    with pytest.raises(S.SadmcError) as e:
        ctx.add_set(S.make_params(system=S.LJ, n=200, lanes=8, method=S.SAD, method_param=0.05), np.arange(2),
                    geom=6.0)
    assert e.value.code == -1  # more than 16 atoms per lane
    with pytest.raises(S.SadmcError) as e:
        ctx.run(10)  # before init
    assert e.value.code == -6
    sid = ctx.add_set(sp, np.arange(2, dtype=np.uint64), geom=2.5)
    ctx.init()
    with pytest.raises(S.SadmcError):
        ctx.add_set(sp, np.arange(2, dtype=np.uint64), geom=2.5)  # after init
    with pytest.raises(S.SadmcError):
        ctx.state(sid, 5)  # no such walker


def test_fp32_pairs_drift_stays_bounded(gpu_ctx_factory):
    """FP32 pair terms with FP64 sums and the 1024-move FP64 refresh: the
    running energy stays within 1e-4 relative of an FP64 recompute."""
    print("case fp32 drift", flush=True)
    sp = S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024)
    ctx = gpu_ctx_factory()
    W = 512
    sid = ctx.add_set(sp, np.arange(W, dtype=np.uint64), geom=3.0)
    ctx.init()
    worst = 0.0
    for _ in range(4):
        ctx.run(50_000 + 333)  # ends between refreshes
        st, _ = ctx.download_raw(sid, with_bins=False)
        for w in range(0, W, 8):
            e_ref = O.lj_energy_ld(ctx.coords(sid, w))
            worst = max(worst, abs(st[w]["E"] - e_ref) / abs(e_ref))
    print("max relative drift", worst)
    assert worst < 1e-4


def test_fp32_refresh_past_the_window_edge(gpu_ctx_factory):
    """LJ31 walker 83142 at 4 lanes: the FP64 refresh at move 2048 puts the
    energy one bin above the window it had grown to; the bins grow to it and
    the walker goes on, bit-identical to the oracle (with its neighbours)."""
    print("case fp32 refresh past the window edge", flush=True)
    sp = S.make_params(system=S.LJ, n=31, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024)
    ctx = gpu_ctx_factory()
    seeds = np.array([83141, 83142, 83143], dtype=np.uint64)
    sid = ctx.add_set(sp, seeds, geom=3.0)
    ctx.init()
    ctx.run(2047)
    before = int(ctx.state(sid, 1)["len"])
    ctx.run(3000 - 2047)
    assert int(ctx.state(sid, 1)["len"]) > before
    for w, seed in enumerate(seeds):
        orc = O.Walker(oracle_params(sp, seed, 3.0))
        orc.run(3000)
        compare_walker(ctx.state(sid, w), ctx.bins(sid, w), orc, sp.method, label=f"lj31 fp32 g4 walker {seed}")
        np.testing.assert_array_equal(ctx.coords(sid, w), orc.coords())


def test_fp32_refresh_below_the_window_grows_the_front(gpu_ctx_factory):
    """Walkers at move 1000 given the positions of their own states at move
    200000, below every bin they have: the refresh at move 1024 grows the
    bins at the front (hundreds of bins) on the device as in the oracle,
    bit for bit through move 3000."""
    print("case fp32 refresh below the window", flush=True)
    sp = S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024)
    seeds = np.array([4, 5], dtype=np.uint64)
    deep = gpu_ctx_factory()
    sd = deep.add_set(sp, seeds, geom=3.0)
    deep.init()
    deep.run(200000)
    ctx = gpu_ctx_factory()
    sid = ctx.add_set(sp, seeds, geom=3.0)
    ctx.init()
    ctx.run(1000)
    st, bins = ctx.download_raw(sid)
    coords = np.concatenate([deep.coords(sd, w) for w in range(len(seeds))])
    ctx.upload_raw(sid, st, bins, coords=coords)
    ctx.run(24)
    for w in range(len(seeds)):
        assert float(ctx.state(sid, w)["bins_min"]) < float(st[w]["bins_min"])
    ctx.run(3000 - 1024)
    for w, seed in enumerate(seeds):
        d = O.Walker(oracle_params(sp, seed, 3.0))
        d.run(200000)
        orc = O.Walker(oracle_params(sp, seed, 3.0))
        orc.run(1000)
        orc.load_state(orc.scalars(), orc.bins(), d.coords())
        orc.run(2000)
        compare_walker(ctx.state(sid, w), ctx.bins(sid, w), orc, sp.method, label=f"lj31 fp32 transplant {seed}")
        np.testing.assert_array_equal(ctx.coords(sid, w), orc.coords())
