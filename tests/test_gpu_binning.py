"""GPU parity of the binning-path walker (src/mc/energy_binning.rs over
src/mc/binning/histogram.rs) against the oracle, bit for bit: every scalar
of the walker, BinCounts' running statistics of ln w and of the "energy",
"t_found", "hist" and "pressure" extras, every logical bin's totals and
counts, the generator and the configuration, after several launches.
"""
import numpy as np
import pytest

import sadmc_amd as S
from conftest import oracle_params

pytestmark = pytest.mark.gpu

BIN_FIELDS = ["lnw_total", "lnw_count", "energy_total", "energy_count", "t_found_total", "t_found_count",
              "hist_count", "pressure_total", "pressure_count"]


def _compare(ctx, s, w, orc, label):
    import oracle as O
    g, o = ctx.bin_state(s, w), orc.state()
    for name in O.BSTATE_DTYPE.names:
        if O.BSTATE_DTYPE.fields[name][0].names:
            for sub in O.BSTATE_DTYPE.fields[name][0].names:
                if sub == "reserved":
                    continue
                assert g[name][sub].tobytes() == o[name][sub].tobytes(), \
                    f"{label} {name}.{sub}: gpu {g[name][sub]!r} oracle {o[name][sub]!r}"
        else:
            assert g[name].tobytes() == o[name].tobytes(), f"{label} {name}: gpu {g[name]!r} oracle {o[name]!r}"
    gb, ob = ctx.bin_bins(s, w), orc.bins()
    for f in BIN_FIELDS:
        assert gb[f].tobytes() == np.ascontiguousarray(ob[f]).tobytes(), f"{label} bins {f}"


def _run(ctx_factory, sp, seeds, geom, chunks, label):
    import oracle as O
    ctx = ctx_factory()
    s = ctx.bin_add_set(sp, seeds, geom=None if sp.system == S.ISING else geom)
    for c in chunks:
        ctx.bin_run(s, c)
    if chunks and chunks[-1] > 0:
        assert ctx.last_run_ms() > 0.0  # the last launch, timed by device events
    for w, seed in enumerate(seeds):
        g = 0.0 if sp.system == S.ISING else float(np.broadcast_to(geom, (len(seeds),))[w])
        orc = O.BWalker(oracle_params(sp, seed, g))
        assert orc.run(sum(chunks)) == 0
        _compare(ctx, s, w, orc, f"{label} walker {w} seed {seed}")
        if sp.system == S.ISING:
            np.testing.assert_array_equal(ctx.bin_spins(s, w), orc.spins())
        else:
            np.testing.assert_array_equal(ctx.bin_coords(s, w), orc.coords())
    return ctx, s


ISING_CASES = [
    ("sad", dict(method=S.SAD, method_param=1.0, energy_bin=1.0)),
    ("samc", dict(method=S.SAMC, method_param=1000.0, energy_bin=1.0)),
    ("wl", dict(method=S.WL, energy_bin=4.0, has_min_e=1, min_e=-24.0, has_max_e=1, max_e=24.0)),
    ("inv_t_wl", dict(method=S.INV_T_WL, energy_bin=4.0, has_min_e=1, min_e=-24.0, has_max_e=1, max_e=24.0)),
    ("wl_min_gamma", dict(method=S.WL, method_param=5e-3, has_min_gamma=1, energy_bin=4.0, has_min_e=1,
                          min_e=-24.0, has_max_e=1, max_e=24.0)),
]


@pytest.mark.parametrize("name,kw", ISING_CASES, ids=[c[0] for c in ISING_CASES])
def test_ising_4x4_bit_exact(gpu_ctx_factory, name, kw):
    print(f"case binning ising 4x4 {name}", flush=True)
    sp = S.make_params(system=S.ISING, n=4, ising_init=1, **kw)
    _run(gpu_ctx_factory, sp, np.arange(16, dtype=np.uint64) * 5 + 1, None, [7000, 1, 12999], name)


def test_ising_8x8_sad_default_width(gpu_ctx_factory):
    print("case binning ising 8x8 sad width 1", flush=True)
    sp = S.make_params(system=S.ISING, n=8, ising_init=1, method=S.SAD, method_param=2.0)
    _run(gpu_ctx_factory, sp, np.arange(8, dtype=np.uint64) + 40, None, [30000], "ising8")


@pytest.mark.parametrize("plan", [S.TRANSLATION_SCALE, S.ACCEPTANCE_RATE])
def test_lj13_sad_bit_exact(gpu_ctx_factory, plan):
    print(f"case binning lj13 sad move_plan {plan}", flush=True)
    sp = S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, energy_bin=0.05,
                       move_plan=plan, move_param=0.1 if plan == S.TRANSLATION_SCALE else 0.3, has_max_e=1,
                       max_e=0.0)
    _run(gpu_ctx_factory, sp, np.arange(8, dtype=np.uint64) + 3, 2.5, [15000, 15000], "lj13")


def test_lj31_samc_and_wl(gpu_ctx_factory):
    for name, kw in (("samc", dict(method=S.SAMC, method_param=500.0)),
                     ("wl", dict(method=S.WL, has_min_e=1, min_e=-100.0))):
        print(f"case binning lj31 {name}", flush=True)
        sp = S.make_params(system=S.LJ, n=31, lanes=8, energy_bin=0.1, move_param=0.1, has_max_e=1, max_e=0.0,
                           **kw)
        _run(gpu_ctx_factory, sp, np.arange(4, dtype=np.uint64) + 11, 3.0, [10000], f"lj31 {name}")


def test_wca32_sad_with_pressure(gpu_ctx_factory):
    print("case binning wca32 sad pressure", flush=True)
    n = 32
    L = (n / 0.6) ** (1.0 / 3.0)
    sp = S.make_params(system=S.WCA, n=n, lanes=8, method=S.SAD, method_param=0.1, energy_bin=0.1, move_param=0.1,
                       has_max_e=1, max_e=2.5 * n, collect_pressure=1)
    ctx, s = _run(gpu_ctx_factory, sp, np.arange(4, dtype=np.uint64) + 7, L, [2100, 2000], "wca32")
    assert int(ctx.bin_state(s, 0)["pressure"]["total_count"]) == 4100 // (n * n)


def test_binning_argument_errors(gpu_ctx_factory):
    print("case binning argument errors", flush=True)
    ctx = gpu_ctx_factory()
    with pytest.raises(S.SadmcError, match="SAD, SAMC, WL"):
        ctx.bin_add_set(S.make_params(system=S.ISING, n=4, method=S.CANONICAL, method_param=2.0), [1])
    with pytest.raises(S.SadmcError, match="FP64"):
        ctx.bin_add_set(S.make_params(system=S.LJ, n=13, method=S.SAD, method_param=0.1, precision=S.FP32_PAIRS,
                                      refresh=1024), [1], geom=2.5)
    s = ctx.bin_add_set(S.make_params(system=S.ISING, n=4, method=S.SAD, method_param=1.0), [1])
    with pytest.raises(S.SadmcError, match="coordinates"):
        ctx.bin_coords(s, 0)


LINEAR_CASES = [
    ("ising_sad", S.ISING, dict(n=4, ising_init=1, method=S.SAD, method_param=1.0, energy_bin=4.0), None, [9000, 6000]),
    ("ising_samc", S.ISING, dict(n=4, ising_init=1, method=S.SAMC, method_param=500.0, energy_bin=1.0), None,
     [15000]),
    ("ising_wl", S.ISING, dict(n=4, ising_init=1, method=S.WL, energy_bin=4.0, has_min_e=1, min_e=-24.0,
                               has_max_e=1, max_e=24.0), None, [15000]),
    ("lj13_sad", S.LJ, dict(n=13, lanes=4, method=S.SAD, method_param=0.05, energy_bin=0.05, move_param=0.1,
                            has_max_e=1, max_e=0.0), 2.5, [10000, 10000]),
    ("lj13_sad_acceptance", S.LJ, dict(n=13, lanes=4, method=S.SAD, method_param=0.05, energy_bin=0.05,
                                       move_plan=S.ACCEPTANCE_RATE, move_param=0.3, has_max_e=1, max_e=0.0),
     2.5, [20000]),
    ("lj31_wl", S.LJ, dict(n=31, lanes=8, method=S.INV_T_WL, energy_bin=0.5, move_param=0.1, has_min_e=1,
                           min_e=-60.0, has_max_e=1, max_e=0.0), 3.0, [10000]),
]


@pytest.mark.parametrize("name,system,kw,geom,chunks", LINEAR_CASES, ids=[c[0] for c in LINEAR_CASES])
def test_linear_binning_bit_exact(gpu_ctx_factory, name, system, kw, geom, chunks):
    print(f"case linear binning {name}", flush=True)
    sp = S.make_params(system=system, binning=S.BIN_LINEAR, **kw)
    W = 16 if system == S.ISING else 4
    _run(gpu_ctx_factory, sp, np.arange(W, dtype=np.uint64) * 3 + 2, geom, chunks, f"linear {name}")


def test_linear_wca32_with_pressure(gpu_ctx_factory):
    print("case linear binning wca32 sad pressure", flush=True)
    n = 32
    L = (n / 0.6) ** (1.0 / 3.0)
    sp = S.make_params(system=S.WCA, n=n, lanes=8, method=S.SAD, method_param=0.1, energy_bin=0.1, move_param=0.1,
                       has_max_e=1, max_e=2.5 * n, collect_pressure=1, binning=S.BIN_LINEAR)
    _run(gpu_ctx_factory, sp, np.arange(4, dtype=np.uint64) + 9, L, [3100], "linear wca32")


@pytest.mark.parametrize("binning", [S.BIN_HISTOGRAM, S.BIN_LINEAR])
def test_window_overflow_is_reported(gpu_ctx_factory, binning):
    # a 9-bin window (e_lo_bound / e_hi_bound) cannot hold a 4x4 Ising walk:
    # every walker stops with a capacity status and its bins stay inside the window
    print(f"case binning window overflow binning={binning}", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=4, ising_init=1, method=S.SAD, method_param=1.0, energy_bin=1.0,
                       e_lo_bound=-10.0, e_hi_bound=-9.0, binning=binning)
    s = ctx.bin_add_set(sp, np.arange(8, dtype=np.uint64) + 1)
    with pytest.raises(S.SadmcError, match="status 4"):
        ctx.bin_run(s, 5000)
    for w in range(8):
        st = ctx.bin_state(s, w)
        assert int(st["status"]) == S.WS_CAPACITY
        assert 0 <= int(st["base"]) and int(st["base"]) + int(st["len"]) <= int(st["cap"]) == 9
        assert float(st["E"]) % 4 == 0 and -32 <= float(st["E"]) <= 32


# ---- the high_resolution side histogram (energy_binning.rs:125, 328-330)

def _compare_hires(ctx, s, w, orc, label):
    gs, gc = ctx.bin_hires(s, w)
    os_, oc = orc.hires()
    for k in ("min", "width", "min_e", "max_e", "e_max_count"):
        assert np.float64(gs[k]).tobytes() == np.float64(os_[k]).tobytes(), f"{label} {k}: {gs[k]!r} {os_[k]!r}"
    for k in ("len", "total_count", "min_count", "max_count"):
        assert int(gs[k]) == os_[k], f"{label} {k}: {gs[k]!r} {os_[k]!r}"
    assert os_["min_total"] == 0.0
    np.testing.assert_array_equal(gc, oc, err_msg=label)


HIRES_CASES = [
    ("ising_sad", dict(system=S.ISING, n=4, ising_init=1, method=S.SAD, method_param=1.0, energy_bin=1.0,
                       high_resolution_de=2.0), None, [6000, 4000]),
    ("lj13_samc", dict(system=S.LJ, n=13, lanes=4, method=S.SAMC, method_param=500.0, energy_bin=0.1,
                       move_param=0.1, has_max_e=1, max_e=0.0, high_resolution_de=0.01), 2.5, [8000, 4000]),
    ("lj13_wl_linear", dict(system=S.LJ, n=13, lanes=4, method=S.WL, energy_bin=0.1, move_param=0.1,
                            has_min_e=1, min_e=-40.0, has_max_e=1, max_e=0.0, binning=S.BIN_LINEAR,
                            high_resolution_de=0.025), 2.5, [10000]),
    ("wca32_sad_pressure", dict(system=S.WCA, n=32, lanes=8, method=S.SAD, method_param=0.1, energy_bin=0.1,
                                move_param=0.1, has_max_e=1, max_e=80.0, collect_pressure=1,
                                high_resolution_de=0.05), (32 / 0.6) ** (1.0 / 3.0), [2100, 2000]),
]


@pytest.mark.parametrize("name,kw,geom,chunks", HIRES_CASES, ids=[c[0] for c in HIRES_CASES])
def test_high_resolution_histogram_bit_exact(gpu_ctx_factory, name, kw, geom, chunks):
    import oracle as O
    print(f"case binning high_resolution {name}", flush=True)
    sp = S.make_params(**kw)
    seeds = np.arange(4, dtype=np.uint64) + 21
    ctx, s = _run(gpu_ctx_factory, sp, seeds, geom, chunks, name)
    for w, seed in enumerate(seeds):
        orc = O.BWalker(oracle_params(sp, seed, 0.0 if geom is None else geom))
        assert orc.run(sum(chunks)) == 0
        _compare_hires(ctx, s, w, orc, f"{name} walker {w}")
        st, counts = ctx.bin_hires(s, w)
        assert int(st["total_count"]) == sum(chunks) == int(counts.sum())


def test_high_resolution_window_overflow_is_reported(gpu_ctx_factory):
    # the side histogram's window (bins of 1 over [-32, -20] plus margins)
    # overflows while the walker's own (bins of 16) does not
    print("case binning high_resolution overflow", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=4, ising_init=1, method=S.SAD, method_param=1.0, energy_bin=16.0,
                       e_lo_bound=-32.0, e_hi_bound=-20.0, high_resolution_de=1.0)
    s = ctx.bin_add_set(sp, np.arange(8, dtype=np.uint64) + 1)
    with pytest.raises(S.SadmcError, match="status 4"):
        ctx.bin_run(s, 20000)
    stopped = 0
    for w in range(8):
        st, counts = ctx.bin_hires(s, w)
        assert 0 <= int(st["base"]) and int(st["base"]) + int(st["len"]) <= 20
        stopped += int(ctx.bin_state(s, w)["status"]) == S.WS_CAPACITY
    assert stopped >= 1
    with pytest.raises(S.SadmcError, match="no high_resolution"):
        ctx.bin_hires(ctx.bin_add_set(S.make_params(system=S.ISING, n=4, method=S.SAD, method_param=1.0), [1]), 0)
