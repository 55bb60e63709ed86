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
    # ln g on the energies every walker reached, relative to E = 0
    m = dict(zip(np.round(Eg).astype(int), Sg))
    both = [k for k, e in enumerate(E) if int(e) in m and count[list(np.round(Eg).astype(int)).index(int(e))] == W]
    est = np.array([m[int(E[k])] for k in both])
    ref = lng[both]
    z = both.index(int(np.where(E == 0)[0][0]))
    assert len(both) >= 30
    assert np.max(np.abs((est - est[z]) - (ref - ref[z]))) < 0.05
    T = np.linspace(1.2, 5.0, 40)
    cv = A.heat_capacity(T, Eg, Sg)
    want = A.heat_capacity(T, E, lng)
    assert np.max(np.abs(cv / want - 1)) < 0.02
