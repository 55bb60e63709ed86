"""GPU parity: the HIP walker engine against the C oracle, bit for bit.

Every case builds the same walkers on the GPU (through the C ABI) and in the
oracle (with the oracle summing in the same lane-group order), runs the same
number of moves, and requires every EnergyMC field, every bin array, the RNG
state and the configuration to be identical.  Sizes are chosen so that the
oracle finishes in a second or two.
"""
import numpy as np
import pytest

import sadmc_amd as S
from conftest import compare_walker, oracle_params

pytestmark = pytest.mark.gpu


def _run_and_compare(ctx_factory, sp, seeds, geom, moves, chunks=1, label=""):
    import oracle as O
    ctx = ctx_factory()
    sid = ctx.add_set(sp, seeds, geom=None if sp.system == S.ISING else geom)
    ctx.init()
    for _ in range(chunks):
        ctx.run(moves // chunks)
    for w, seed in enumerate(seeds):
        g = 0.0 if sp.system == S.ISING else float(np.broadcast_to(geom, (len(seeds),))[w])
        orc = O.Walker(oracle_params(sp, seed, g))
        orc.run(moves)
        assert orc.scalars()["status"] == 0
        st = ctx.state(sid, w)
        compare_walker(st, ctx.bins(sid, w), orc, sp.method, label=f"{label} walker {w} seed {seed}")
        if sp.system == S.ISING:
            np.testing.assert_array_equal(ctx.spins(sid, w), orc.spins())
        else:
            np.testing.assert_array_equal(ctx.coords(sid, w), orc.coords())
    return ctx, sid


METHODS = [
    ("sad", dict(method=S.SAD, method_param=1.0)),
    ("samc", dict(method=S.SAMC, method_param=1000.0)),
    ("wl", dict(method=S.WL, has_min_e=1, min_e=-128.0, has_max_e=1, max_e=128.0)),
    ("wl_min_gamma", dict(method=S.WL, method_param=1e-3, has_min_gamma=1)),
    ("inv_t_wl", dict(method=S.INV_T_WL, has_min_e=1, min_e=-128.0, has_max_e=1, max_e=128.0)),
    ("canonical", dict(method=S.CANONICAL, method_param=2.5)),
]


@pytest.mark.parametrize("name,kw", METHODS, ids=[m[0] for m in METHODS])
def test_ising_8x8_bit_exact(gpu_ctx_factory, name, kw):
    print(f"case ising 8x8 {name}", flush=True)
    sp = S.make_params(system=S.ISING, n=8, ising_init=1, **kw)
    _run_and_compare(gpu_ctx_factory, sp, np.arange(48, dtype=np.uint64) * 7 + 3, None, 30000, label=name)


def test_ising_reference_spins_and_small_lattices(gpu_ctx_factory):
    for n in (2, 4, 6, 8):
        print(f"case ising {n}x{n} reference spins", flush=True)
        sp = S.make_params(system=S.ISING, n=n, ising_init=0, method=S.SAD, method_param=1.0)
        _run_and_compare(gpu_ctx_factory, sp, np.arange(8, dtype=np.uint64), None, 5000, label=f"N={n}")


def test_ising_odd_lattice_edge_bin_fails_like_reference(gpu_ctx_factory):
    """Odd N: an initial E = 2 (mod 4) lies on the upper edge of its one
    bin (energy.rs:852 rounds half away from zero), so the first move that
    does not grow the bins indexes past them -- the reference panics there.
    Both engines must stop the same walkers at the same move."""
    import oracle as O
    n = 5
    sp = S.make_params(system=S.ISING, n=n, ising_init=1, method=S.SAD, method_param=1.0)
    seeds = np.arange(32, dtype=np.uint64)
    ctx = gpu_ctx_factory()
    sid = ctx.add_set(sp, seeds)
    ctx.init()
    with pytest.raises(S.SadmcError) as ei:
        ctx.run(200)
    assert ei.value.code == -8
    st, _ = ctx.download_raw(sid, with_bins=False)
    failed = 0
    for w, seed in enumerate(seeds):
        orc = O.Walker(oracle_params(sp, seed))
        orc.run(200)
        s = orc.scalars()
        assert int(st[w]["status"]) == s["status"], f"walker {w}"
        if s["status"]:
            failed += 1
            assert int(st[w]["moves"]) == s["moves"]
    assert failed > 0


def test_ising_chunked_equals_single_launch(gpu_ctx_factory):
    print("case ising chunked", flush=True)
    sp = S.make_params(system=S.ISING, n=8, ising_init=1, method=S.SAD, method_param=1.0)
    _run_and_compare(gpu_ctx_factory, sp, np.arange(16, dtype=np.uint64), None, 24000, chunks=6, label="chunks")


LJ_CASES = [
    ("lj13_g4", 13, 2.5, 4, {}),
    ("lj13_g16", 13, 2.5, 16, {}),
    ("lj31_g8", 31, 3.0, 8, {}),
    ("lj31_g8_fp32", 31, 3.0, 8, dict(precision=S.FP32_PAIRS, refresh=1024)),
    ("lj38_g16", 38, 3.2, 16, {}),
    ("lj31_g8_accrate", 31, 3.0, 8, dict(move_plan=S.ACCEPTANCE_RATE, move_param=0.5)),
]


@pytest.mark.parametrize("name,n,radius,lanes,kw", LJ_CASES, ids=[c[0] for c in LJ_CASES])
def test_lj_sad_bit_exact(gpu_ctx_factory, name, n, radius, lanes, kw):
    print(f"case {name}", flush=True)
    args = dict(system=S.LJ, n=n, lanes=lanes, method=S.SAD, method_param=0.05, move_param=0.1,
                energy_bin=0.01, has_max_e=1, max_e=0.0)
    args.update(kw)
    sp = S.make_params(**args)
    _run_and_compare(gpu_ctx_factory, sp, np.arange(6, dtype=np.uint64) + 100, radius, 8000, chunks=2, label=name)


def test_lj31_fp32_refresh_into_another_bin(gpu_ctx_factory):
    """The fixed FP64 refresh of the precision configuration can move the
    tracked energy across a bin edge; the move that follows must then use
    the bin of the refreshed energy, as move_once indexes the bins by the
    energy itself (energy.rs:904-974).  The oracle shows these seeds'
    refreshes at moves 1024-9216 crossing an edge."""
    print("case lj31 fp32 refresh across a bin edge", flush=True)
    sp = S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024)
    seeds = np.array([97, 255, 406, 1275, 1309, 1492, 1721, 1804, 2289, 2119], dtype=np.uint64)
    _run_and_compare(gpu_ctx_factory, sp, seeds, 3.0, 12000, chunks=3, label="refresh crossing")


@pytest.mark.parametrize("n,lanes", [(32, 8), (64, 16), (128, 32)])
def test_wca_sad_bit_exact(gpu_ctx_factory, n, lanes):
    print(f"case wca n={n} lanes={lanes}", flush=True)
    etas = np.array([0.3, 0.4, 0.5, 0.45])
    box = np.cbrt(n / (6 * etas / np.pi))
    sp = S.make_params(system=S.WCA, n=n, lanes=lanes, method=S.SAD, method_param=0.1, move_param=0.05,
                       energy_bin=0.05, has_max_e=1, max_e=2.5 * n, collect_pressure=1)
    _run_and_compare(gpu_ctx_factory, sp, np.arange(4, dtype=np.uint64) + 7, box, 4000, label=f"wca{n}")


def test_plan_confirm_pair_matches_oracle(gpu_ctx_factory):
    """MovableSystem::plan_move / ConfirmSystem::confirm through the C ABI."""
    import oracle as O
    print("case plan/confirm", flush=True)
    sp = S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    ctx = gpu_ctx_factory()
    sid = ctx.add_set(sp, np.array([5], np.uint64), geom=2.5)
    ctx.init()
    st = ctx.state(sid, 0)
    c0 = ctx.coords(sid, 0)
    # Replay the same plan/confirm sequence on the oracle's pair-term sums.
    E = st["E"]
    coords = c0.copy()
    rng = O.RNG(state=tuple(int(v) for v in st["rng"]))
    n = 13
    for k in range(200):
        e = ctx.plan_move(sid, 0, 0.3)
        w = rng.uniform_index(n)
        d = [rng.uniform_f64(-0.3, 0.3) for _ in range(3)]
        f = coords.reshape(3, n)[:, w].copy()
        t = f + np.array(d)
        if t @ t > 2.5 ** 2 and t @ t > f @ f:
            assert e is None
            continue
        acc = np.zeros(4)
        for lane in range(4):
            for j in range(lane, n, 4):
                if j == w:
                    continue
                rj = coords.reshape(3, n)[:, j]
                acc[lane] += O.lib().or_lj_pair(float(((rj[0] - t[0]) ** 2 + (rj[1] - t[1]) ** 2) + (rj[2] - t[2]) ** 2)) \
                    - O.lib().or_lj_pair(float(((rj[0] - f[0]) ** 2 + (rj[1] - f[1]) ** 2) + (rj[2] - f[2]) ** 2))
        tot = ((acc[0] + acc[1]) + (acc[2] + acc[3]))
        assert e == E + tot
        if k % 3 == 0:
            ctx.confirm(sid, 0)
            coords.reshape(3, n)[:, w] = t
            E = ctx.state(sid, 0)["E"]
    np.testing.assert_array_equal(ctx.coords(sid, 0), coords)


def test_resume_is_bit_identical(gpu_ctx_factory):
    """tests/resume-sad.rs pattern: stop, dump, reload into a fresh context,
    continue -- identical to an uninterrupted run."""
    print("case resume", flush=True)
    sp = S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    seeds = np.arange(8, dtype=np.uint64)
    a = gpu_ctx_factory()
    sa = a.add_set(sp, seeds, geom=2.5)
    a.init()
    a.run(3000)
    st, bins = a.download_raw(sa)
    coords = np.concatenate([a.coords(sa, w) for w in range(len(seeds))])
    a.run(2000)
    b = gpu_ctx_factory()
    sb = b.add_set(sp, seeds, geom=2.5)
    b.init()
    b.upload_raw(sb, st, bins, coords=coords)
    b.run(2000)
    st_a, bins_a = a.download_raw(sa)
    st_b, bins_b = b.download_raw(sb)
    assert st_a.tobytes() == st_b.tobytes()
    assert bins_a.tobytes() == bins_b.tobytes()


CONT_METHODS = [
    ("lj13_samc", S.LJ, 13, 2.5, 4, dict(method=S.SAMC, method_param=500.0, energy_bin=0.01, has_max_e=1,
                                         max_e=0.0)),
    ("lj13_wl", S.LJ, 13, 2.5, 4, dict(method=S.WL, energy_bin=0.05, has_min_e=1, min_e=-44.0, has_max_e=1,
                                       max_e=0.0)),
    ("lj13_inv_t_wl", S.LJ, 13, 2.5, 4, dict(method=S.INV_T_WL, energy_bin=0.05, has_min_e=1, min_e=-44.0,
                                             has_max_e=1, max_e=0.0)),
    ("lj19_canonical", S.LJ, 19, 2.7, 8, dict(method=S.CANONICAL, method_param=0.3, energy_bin=0.01,
                                              has_max_e=1, max_e=0.0)),
    ("lj13_fp32_canonical", S.LJ, 13, 2.5, 4, dict(method=S.CANONICAL, method_param=0.2, energy_bin=0.01,
                                                   has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024)),
    ("wca32_canonical", S.WCA, 32, 3.6, 8, dict(method=S.CANONICAL, method_param=1.0, energy_bin=0.05,
                                                has_max_e=1, max_e=80.0)),
    ("wca64_samc", S.WCA, 64, 4.3, 16, dict(method=S.SAMC, method_param=500.0, energy_bin=0.05, has_max_e=1,
                                            max_e=160.0)),
]


@pytest.mark.parametrize("name,system,n,geom,lanes,kw", CONT_METHODS, ids=[c[0] for c in CONT_METHODS])
def test_continuous_methods_bit_exact(gpu_ctx_factory, name, system, n, geom, lanes, kw):
    """SAMC, WL, 1/t-WL and canonical sampling of the LJ and WCA systems."""
    print(f"case {name}", flush=True)
    sp = S.make_params(system=system, n=n, lanes=lanes, move_param=0.1 if system == S.LJ else 0.05, **kw)
    _run_and_compare(gpu_ctx_factory, sp, np.arange(4, dtype=np.uint64) + 300, geom, 6000, chunks=2, label=name)
