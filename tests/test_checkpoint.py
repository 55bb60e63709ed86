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
    rec, rows, c, sp, pr = CK.from_document(doc, p, cap, e_lo, 2.5)
    assert pr is None
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


# RFC 8949 Appendix A: published encodings of data items (preferred
# serialisation: the shortest float that holds the value exactly).
RFC8949_VECTORS = [
    (0, "00"), (1, "01"), (10, "0a"), (23, "17"), (24, "1818"), (25, "1819"), (100, "1864"), (1000, "1903e8"),
    (1000000, "1a000f4240"), (1000000000000, "1b000000e8d4a51000"), (18446744073709551615, "1bffffffffffffffff"),
    (-1, "20"), (-10, "29"), (-100, "3863"), (-1000, "3903e7"),
    (0.0, "f90000"), (-0.0, "f98000"), (1.0, "f93c00"), (1.1, "fb3ff199999999999a"), (1.5, "f93e00"),
    (65504.0, "f97bff"), (100000.0, "fa47c35000"), (3.4028234663852886e+38, "fa7f7fffff"),
    (1.0e+300, "fb7e37e43c8800759c"), (5.960464477539063e-8, "f90001"), (0.00006103515625, "f90400"),
    (-4.0, "f9c400"), (-4.1, "fbc010666666666666"), (float("inf"), "f97c00"), (float("-inf"), "f9fc00"),
    (False, "f4"), (True, "f5"), (None, "f6"),
    ("", "60"), ("a", "6161"), ("IETF", "6449455446"), ("\"\\", "62225c"), ("ü", "62c3bc"),
    ("水", "63e6b0b4"),
    ([], "80"), ([1, 2, 3], "83010203"), ([1, [2, 3], [4, 5]], "8301820203820405"),
    (list(range(1, 26)), "98190102030405060708090a0b0c0d0e0f101112131415161718181819"),
    ({}, "a0"), ({"a": 1, "b": [2, 3]}, "a26161016162820203"), (["a", {"b": "c"}], "826161a161626163"),
]


@pytest.mark.parametrize("value,hexstr", RFC8949_VECTORS, ids=[h for _, h in RFC8949_VECTORS])
def test_cbor_rfc8949_vectors(value, hexstr):
    assert CK.cbor_dumps(value).hex() == hexstr
    back = CK.cbor_loads(bytes.fromhex(hexstr))
    assert back == value and type(back) is type(value)
    if isinstance(value, float):
        assert np.float64(back).tobytes() == np.float64(value).tobytes()


def test_cbor_decodes_indefinite_lengths_and_nan():
    # RFC 8949 Appendix A: indefinite-length array and map
    assert CK.cbor_loads(bytes.fromhex("9f018202039f0405ffff")) == [1, [2, 3], [4, 5]]
    assert CK.cbor_loads(bytes.fromhex("bf61610161629f0203ffff")) == {"a": 1, "b": [2, 3]}
    assert np.isnan(CK.cbor_loads(bytes.fromhex("f97e00")))
    with pytest.raises(ValueError):
        CK.cbor_loads(bytes.fromhex("8301"))  # truncated


@pytest.mark.parametrize("ext", ["yaml", "json", "cbor"])
def test_document_files_by_extension(tmp_path, ext):
    """save/load pick the format from the extension (src/mc/mod.rs:72-79,
    113-119); every value of a document survives exactly."""
    p = S.make_params(system=S.WCA, n=32, method=S.SAD, method_param=0.1, move_param=0.05, energy_bin=0.05,
                      has_max_e=1, max_e=80.0, collect_pressure=1)
    st, bins, _ = synthetic_walker()
    st["has_pressure"] = 1
    rng = np.random.default_rng(9)
    bins["pressure_count"] = rng.integers(0, 50, 40).astype(np.uint64)
    bins["pressure_total"] = rng.normal(size=40) * 1e3
    coords = rng.uniform(0, 3.5, 3 * 32)
    doc = CK.to_document(p, st, bins, coords=coords, save_as=f"w.{ext}")
    path = tmp_path / f"w.{ext}"
    CK.save(doc, path)
    assert CK.load(path) == doc
    with pytest.raises(ValueError):
        CK.save(doc, tmp_path / "w.toml")
    # the pressure extra reaches the physical rows of the device layout
    st2, rows, c, sp, pr = CK.from_document(CK.load(path), p, 4000, 0.0, 3.5)
    sl = slice(int(st2["base"]), int(st2["base"]) + 40)
    np.testing.assert_array_equal(pr["count"][sl], bins["pressure_count"])
    np.testing.assert_array_equal(pr["total"][sl], bins["pressure_total"])
    assert pr["count"].sum() == bins["pressure_count"].sum()
    assert int(st2["has_pressure"]) == 1


@pytest.mark.gpu
def test_resume_wca_with_pressure_is_bit_identical(gpu_ctx_factory, tmp_path):
    """A WCA walker set collecting the pressure, checkpointed to CBOR
    mid-run and resumed on a fresh context, continues bit for bit,
    pressure extra included."""
    print("case resume wca pressure from cbor", flush=True)
    sp = S.make_params(system=S.WCA, n=32, lanes=8, method=S.SAD, method_param=0.1, move_param=0.05,
                       energy_bin=0.05, has_max_e=1, max_e=80.0, collect_pressure=1)
    seeds = np.arange(6, dtype=np.uint64) + 7
    box = np.cbrt(32 / (6.0 * np.array([0.3, 0.35, 0.4, 0.45, 0.5, 0.42]) / np.pi))
    a = gpu_ctx_factory()
    sa = a.add_set(sp, seeds, geom=box)
    a.init()
    a.run(6000)
    for w, d in enumerate(CK.save_set(a, sa)):
        assert d["bins"]["extra"]["pressure"]["count"], "pressure collected"
        CK.save(d, tmp_path / f"w{w}.cbor")
    a.run(4000)
    b = gpu_ctx_factory()
    sb = b.add_set(sp, seeds, geom=box)
    b.init()
    CK.resume_set(b, sb, [CK.load(tmp_path / f"w{w}.cbor") for w in range(len(seeds))])
    b.run(4000)
    da, db = CK.save_set(a, sa), CK.save_set(b, sb)
    for w in range(len(seeds)):
        assert da[w] == db[w], f"walker {w}"
