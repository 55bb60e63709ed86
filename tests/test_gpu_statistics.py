"""GPU statistics: converged ln g(E) and Cv(T) of the 8x8 Ising model from
the device's SAD walkers (min_T = 1, i.i.d. initial spins, integer energy
bins) against the exact density of states (tests/golden/ising8_dos.json,
itself checked against Kaufman's closed form in test_analysis.py)."""
import json
import os

import numpy as np
import pytest

import analysis as A
import sadmc_amd as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_ising_sad_walkers_converge_to_exact_cv(gpu_ctx_factory):
    print("case ising 8x8 SAD statistics", flush=True)
    d = json.load(open(os.path.join(GOLDEN, "ising8_dos.json")))
    E = np.array(d["energies"], float)
    lng = np.log(np.array([float(int(c)) for c in d["counts"]]))
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.ISING, n=8, method=S.SAD, method_param=1.0, ising_init=1)
    W = 1024
    sid = ctx.add_set(sp, np.arange(W, dtype=np.uint64))
    ctx.init()
    ctx.run(1_000_000)
    states, bins = ctx.download_raw(sid)
    assert np.all(states["status"] == 0)
    walkers = [(states[w], S.logical_bins(states[w], bins[w])) for w in range(W)]
    Eg, Sg, count = A.mean_entropy(walkers)
    # the device reduction gives the same estimate
    Ed, Sd, cd = ctx.mean_entropy(sid)
    np.testing.assert_array_equal(cd, count)
    np.testing.assert_allclose(Ed, Eg, atol=1e-9)
    np.testing.assert_allclose(Sd, Sg, atol=1e-9)
    # ln g on the energies every walker reached, relative to E = 0
    m = dict(zip(np.round(Eg).astype(int), Sg))
    eg = list(np.round(Eg).astype(int))
    # the SAD range of min_T = 1 is the lower half of the spectrum
    both = [k for k, e in enumerate(E) if e <= 0 and int(e) in m and count[eg.index(int(e))] >= W // 2]
    est = np.array([m[int(E[k])] for k in both])
    ref = lng[both]
    z = both.index(int(np.where(E == 0)[0][0]))
    assert len(both) >= 28
    assert np.max(np.abs((est - est[z]) - (ref - ref[z]))) < 0.05
    T = np.linspace(1.2, 5.0, 40)
    cv = A.heat_capacity(T, Eg, Sg)
    want = A.heat_capacity(T, E, lng)
    assert np.max(np.abs(cv / want - 1)) < 0.02


def test_device_entropy_reduction_matches_host_lj(gpu_ctx_factory):
    print("case LJ13 entropy reduction", flush=True)
    ctx = gpu_ctx_factory()
    sp = S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    W = 256
    sid = ctx.add_set(sp, np.arange(W, dtype=np.uint64), geom=2.5)
    ctx.init()
    ctx.run(100000)
    states, bins = ctx.download_raw(sid)
    walkers = [(states[w], S.logical_bins(states[w], bins[w])) for w in range(W)]
    for window in (None, (-10.0, -5.0)):
        Eg, Sg, count = A.mean_entropy(walkers, window)
        Ed, Sd, cd = ctx.mean_entropy(sid, window)
        np.testing.assert_array_equal(cd, count)
        np.testing.assert_allclose(Ed, Eg, atol=1e-9)
        np.testing.assert_allclose(Sd, Sg, atol=1e-8)
