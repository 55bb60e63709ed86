"""GPU: the boundary entry points a MonteCarlo shim needs besides the run:
advancing one set only (the drain shared by one set's walkers), several
sets advanced side by side on their own streams, and the System trait's
compute_energy / verify_energy on the device."""
import numpy as np
import pytest

import oracle as O
import sadmc_amd as S
from conftest import compare_walker, oracle_params

pytestmark = pytest.mark.gpu


def test_run_set_advances_only_its_set(gpu_ctx_factory):
    print("case run_set", flush=True)
    ctx = gpu_ctx_factory()
    lj = S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    ising = S.make_params(system=S.ISING, n=8, method=S.SAD, method_param=1.0, ising_init=1)
    a = ctx.add_set(lj, np.arange(8, dtype=np.uint64), geom=2.5)
    b = ctx.add_set(ising, np.arange(8, dtype=np.uint64))
    ctx.init()
    ctx.run_set(a, 3000)
    ctx.run_set(b, 500)
    ctx.run_set(a, 1000)
    sa, _ = ctx.download_raw(a, with_bins=False)
    sb, _ = ctx.download_raw(b, with_bins=False)
    assert np.all(sa["moves"] == 4000) and np.all(sb["moves"] == 500)
    for w in (0, 7):
        orc = O.Walker(oracle_params(lj, w, 2.5))
        orc.run(4000)
        compare_walker(ctx.state(a, w), ctx.bins(a, w), orc, lj.method, label=f"lj walker {w}")
        orc = O.Walker(oracle_params(ising, w))
        orc.run(500)
        compare_walker(ctx.state(b, w), ctx.bins(b, w), orc, ising.method, label=f"ising walker {w}")


def test_sets_run_side_by_side_equal_sets_alone(gpu_ctx_factory):
    """Five sets in one context (more than the four side streams, so two
    share one) advanced together, against each set in a context of its own:
    every walker record and bin byte-identical, after launches interleaved
    with downloads and a single-set run."""
    print("case concurrent sets", flush=True)
    specs = [
        (S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0), 2.5, 96),
        (S.make_params(system=S.ISING, n=8, method=S.SAD, method_param=1.0, ising_init=1), None, 200),
        (S.make_params(system=S.WCA, n=32, lanes=8, method=S.SAD, method_param=0.1, move_param=0.05,
                       energy_bin=0.05, has_max_e=1, max_e=80.0), 3.6, 64),
        (S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, precision=S.FP32_PAIRS, refresh=1024), 3.0, 128),
        (S.make_params(system=S.LJ, n=19, lanes=4, method=S.WL, move_param=0.1, energy_bin=0.05,
                       has_min_e=1, min_e=-75.0, has_max_e=1, max_e=0.0), 2.7, 64),
    ]

    def drive(ctx, ids):
        ctx.init()
        ctx.run(3000)
        first = [ctx.download_raw(i, with_bins=False)[0] for i in ids]
        ctx.run(2000)
        ctx.run_set(ids[0], 500)
        ctx.run(1000)
        return first, [ctx.download_raw(i) for i in ids]

    ctx = gpu_ctx_factory()
    ids = [ctx.add_set(sp, np.arange(W, dtype=np.uint64) + 10 * k, geom=g)
           for k, (sp, g, W) in enumerate(specs)]
    first, together = drive(ctx, ids)
    for k, (sp, g, W) in enumerate(specs):
        alone = gpu_ctx_factory()
        i = alone.add_set(sp, np.arange(W, dtype=np.uint64) + 10 * k, geom=g)
        alone.init()
        alone.run(3000)
        f = alone.download_raw(i, with_bins=False)[0]
        alone.run(2000)
        if k == 0:
            alone.run_set(i, 500)
        alone.run(1000)
        st, bins = alone.download_raw(i)
        assert f.tobytes() == first[k].tobytes(), f"set {k} after the first launch"
        assert st.tobytes() == together[k][0].tobytes(), f"set {k} walker records"
        assert bins.tobytes() == together[k][1].tobytes(), f"set {k} bins"
        assert np.all(st["moves"] == 6000 + (500 if k == 0 else 0))


@pytest.mark.parametrize("system,n,lanes,geom", [(S.LJ, 31, 8, 3.0), (S.WCA, 64, 16, 4.2), (S.ISING, 8, 1, 0.0)])
def test_compute_and_verify_energy(gpu_ctx_factory, system, n, lanes, geom):
    print(f"case compute/verify energy system {system} n {n}", flush=True)
    kw = dict(energy_bin=0.01, has_max_e=1, max_e=0.0) if system == S.LJ else (
        dict(energy_bin=0.05, has_max_e=1, max_e=2.5 * n) if system == S.WCA else dict(ising_init=1))
    sp = S.make_params(system=system, n=n, lanes=lanes if system != S.ISING else 0, method=S.SAD,
                       method_param=0.1 if system != S.ISING else 1.0, move_param=0.05, **kw)
    ctx = gpu_ctx_factory()
    sid = ctx.add_set(sp, np.arange(4, dtype=np.uint64), geom=None if system == S.ISING else geom)
    ctx.init()
    ctx.run(5000)
    for w in range(4):
        orc = O.Walker(oracle_params(sp, w, geom))
        orc.run(5000)
        e = ctx.compute_energy(sid, w)
        # the oracle's recompute in the same lane-group order, bit for bit
        assert e == orc.compute_energy(), (w, e, orc.compute_energy())
        ok, eg = ctx.verify_energy(sid, w)
        assert ok and eg == e
        if system == S.ISING:
            assert e == ctx.state(sid, w)["E"]
        elif system == S.LJ:
            assert abs(e - O.lj_energy_ld(ctx.coords(sid, w))) <= 1e-12 * abs(e)
