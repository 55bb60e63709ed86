"""Checkpoint documents (the serde layout of EnergyMC, src/mc/energy.rs:
145-245): YAML round trip, the keys plotting/parse-binning.py:102-171 reads,
and the document -> raw device records mapping, on the CPU; resume on the
device from a document is bit-identical to an uninterrupted run (the
tests/resume-sad.rs pattern) on the GPU."""
import numpy as np
import pytest

import checkpoint as CK
import sadmc_amd as S


def synthetic_walker(n_bins=40, method=S.SAD):
    rng = np.random.default_rng(3)
    st = np.zeros((), S.WALKER_STATE_DTYPE)
    st["rng"] = [2 ** 64 - 5, 12345]
    st["E"], st["error"], st["moves"], st["accepted_moves"] = -40.123456789, 3e-13, 100000, 61234
    st["acceptance_rate"], st["translation_scale"] = 0.61, 0.1
    st["bins_min"], st["bins_width"], st["len"] = -44.015, 0.01, n_bins
    st["max_S"], st["max_S_index"], st["method"] = 12.5, 7, method
    st["too_lo"], st["too_hi"], st["min_T"] = -43.9, -43.7, 0.05
    st["tL"], st["tF"], st["num_states"], st["highest_hist"], st["latest_parameter"] = 99000, 98000, 20, 800, 4.0
    st["geom"] = 2.5
    bins = {"lnw": rng.normal(size=n_bins), "histogram": rng.integers(0, 1000, n_bins).astype(np.uint64),
            "t_found": np.sort(rng.integers(0, 100000, n_bins)).astype(np.uint64),
            "energy_total": rng.normal(size=n_bins), "energy_squared_total": rng.uniform(size=n_bins),
            "visited": rng.integers(0, 2, n_bins).astype(np.uint8),
            "round_trips": rng.integers(0, 9, n_bins).astype(np.uint64),
            "wl_hist": np.zeros(n_bins, np.uint64), "pressure_count": np.zeros(n_bins, np.uint64),
            "pressure_total": np.zeros(n_bins)}
    coords = rng.normal(size=3 * 13)
    return st, bins, coords


def test_document_yaml_round_trip_and_parse_binning_keys():
    p = S.make_params(system=S.LJ, n=13, method=S.SAD, method_param=0.05, move_param=0.1, energy_bin=0.01,
                      has_max_e=1, max_e=0.0)
    st, bins, coords = synthetic_walker()
    doc = CK.to_document(p, st, bins, coords=coords)
    back = CK.loads(CK.dumps(doc))
    assert back == doc
    # what parse-binning.py reads from a SAD file
    b = back["bins"]
    assert set(["min", "width", "lnw", "histogram", "extra", "energy_total", "energy_squared_total"]) <= set(b)
    assert len(b["lnw"]) == len(b["histogram"]) == len(b["energy_total"]) == 40
    assert set(["too_lo", "too_hi", "min_T"]) <= set(back["method"]["Sad"])
    assert isinstance(back["moves"], int)
    # values survive exactly
    np.testing.assert_array_equal(np.array(b["lnw"]), bins["lnw"])
    assert back["rng"] == {"s0": 2 ** 64 - 5, "s1": 12345}
    pos = np.array(back["system"]["Lj"]["positions"]).T.reshape(-1)
    np.testing.assert_array_equal(pos, coords)


def test_document_to_device_records():
    p = S.make_params(system=S.LJ, n=13, method=S.SAD, method_param=0.05, move_param=0.1, energy_bin=0.01,
                      has_max_e=1, max_e=0.0)
    st, bins, coords = synthetic_walker()
    doc = CK.loads(CK.dumps(CK.to_document(p, st, bins, coords=coords)))
    cap = 4500
    e_lo = -45.326801
    rec, rows, c, sp = CK.from_document(doc, p, cap, e_lo, 2.5)
    base = int(rec["base"])
    assert 0 <= base and base + 40 <= cap
    for k in ("E", "error", "moves", "accepted_moves", "acceptance_rate", "translation_scale", "bins_min",
              "bins_width", "len", "max_S", "max_S_index", "method", "too_lo", "too_hi", "min_T", "tL", "tF",
              "num_states", "highest_hist", "latest_parameter"):
        assert rec[k] == st[k], k
    assert list(rec["rng"]) == list(st["rng"])
    sl = slice(base, base + 40)
    np.testing.assert_array_equal(rows["lnw"][sl], bins["lnw"])
    np.testing.assert_array_equal(rows["t_found"][sl], bins["t_found"])
    np.testing.assert_array_equal(rows["visit_epoch"][sl], bins["visited"])
    assert rows["histogram"][:base].sum() == 0 and rows["histogram"][base + 40:].sum() == 0
    # the SAD running maximum of t_found over [too_lo, too_hi]
    lo = int(np.floor((st["too_lo"] - st["bins_min"]) / 0.01))
    hi = int(np.floor((st["too_hi"] - st["bins_min"]) / 0.01))
    assert rec["tmax"] == bins["t_found"][lo:hi + 1].max()
    np.testing.assert_array_equal(c, coords)
    assert sp is None


@pytest.mark.gpu
def test_resume_from_document_is_bit_identical(gpu_ctx_factory):
    print("case resume from yaml", flush=True)
    sp = S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0)
    seeds = np.arange(8, dtype=np.uint64) + 40
    a = gpu_ctx_factory()
    sa = a.add_set(sp, seeds, geom=2.5)
    a.init()
    a.run(3000)
    text = [CK.dumps(d) for d in CK.save_set(a, sa)]
    a.run(2000)
    b = gpu_ctx_factory()
    sb = b.add_set(sp, seeds, geom=2.5)
    b.init()
    CK.resume_set(b, sb, [CK.loads(t) for t in text])
    b.run(2000)
    da, db = CK.save_set(a, sa), CK.save_set(b, sb)
    for w in range(len(seeds)):
        assert da[w] == db[w], f"walker {w}"
    # and the Ising spins path
    si = S.make_params(system=S.ISING, n=8, method=S.SAD, method_param=1.0, ising_init=1)
    c = gpu_ctx_factory()
    sc = c.add_set(si, np.arange(4, dtype=np.uint64))
    c.init()
    c.run(20000)
    text = [CK.dumps(d) for d in CK.save_set(c, sc)]
    c.run(10000)
    d = gpu_ctx_factory()
    sd = d.add_set(si, np.arange(4, dtype=np.uint64))
    d.init()
    CK.resume_set(d, sd, [CK.loads(t) for t in text])
    d.run(10000)
    assert CK.save_set(c, sc) == CK.save_set(d, sd)
