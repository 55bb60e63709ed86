"""Checkpoint documents in the layout the reference serialises `EnergyMC`
with (serde, externally tagged enums; src/mc/energy.rs:145-245,
src/system/{lj,ising,wca}.rs), as YAML, JSON or CBOR chosen by the file's
extension as the reference does (src/mc/mod.rs:72-79, 104-119).

A document carries what `plotting/parse-binning.py:102-171` reads (`moves`,
`bins.{min, width, histogram, lnw, energy_total, energy_squared_total,
extra}`, `method.Sad.{too_lo, too_hi, min_T}`) and everything a walker needs
to resume on the device bit for bit: the method state, the counters, the
acceptance rate and translation scale, the generator state
(`rng.{s0, s1}`), `t_found`, `round_trips`, `have_visited_since_maxentropy`,
`max_S`, WCA's `pressure` extra, and the system (energy, its error estimate,
and the positions or spins).  `time_L` is written as 0: the reference
initialises it to 0 (energy.rs:862) and never changes it.  Not written: the
reference's plugin state (`report`, `movies`, `save`, `manager`) and, for
WCA, its cell lists (`optcell.rs`), which a resume on the device rebuilds
from the positions; `save_as` is the file name.  Units are plain numbers in
sigma and epsilon.

`to_document` builds one walker's document from the device views
(`Context.state`, `Context.bins`, `Context.coords`/`spins`);
`from_document` turns it back into the raw records `Context.upload_raw`
takes (and the pressure rows `Context.upload_pressure` takes).
`Context`-level helpers: `save_set` and `resume_set`; files: `save`, `load`.

CBOR follows RFC 8949 with definite lengths, maps with text keys in field
order, and each float in the shortest of half, single or double precision
that holds it exactly (the preferred serialisation, which serde_cbor's
float writer also uses as recalled); byte equality with serde_cbor's output
is parity unpinned (no Rust build here), while every value round-trips
exactly.
"""
import json
import os
import struct

import numpy as np
import yaml

import sadmc_amd as S


def _f(x):
    return float(x)


def _method_doc(st, params):
    m = int(st["method"])
    if m == S.SAD:
        return {"Sad": {"min_T": _f(st["min_T"]), "too_lo": _f(st["too_lo"]), "too_hi": _f(st["too_hi"]),
                        "tL": int(st["tL"]), "tF": int(st["tF"]), "num_states": int(st["num_states"]),
                        "highest_hist": int(st["highest_hist"]), "version": "Sad",
                        "latest_parameter": _f(st["latest_parameter"])}}
    if m == S.SAMC:
        return {"Samc": {"t0": _f(st["t0"])}}
    if m in (S.WL, S.INV_T_WL):
        return {"WL": {"gamma": _f(st["wl_gamma"]), "lowest_hist": int(st["wl_lowest_hist"]),
                       "highest_hist": int(st["wl_highest_hist"]), "total_hist": int(st["wl_total_hist"]),
                       "num_states": _f(st["wl_num_states"]), "hist": None, "min_energy": _f(st["wl_min_energy"]),
                       "inv_t": m == S.INV_T_WL,
                       "min_gamma": _f(params.method_param) if params.has_min_gamma else None}}
    return {"Canonical": {"temperature": _f(params.method_param)}}


def to_document(params, st, bins, coords=None, spins=None, save_as="resume.yaml"):
    """One walker's checkpoint document (a dict ready for yaml.safe_dump)."""
    n = int(st["len"])
    method = _method_doc(st, params)
    if "WL" in method:
        method["WL"]["hist"] = [int(x) for x in bins["wl_hist"][:int(st["wl_len"])]]
    extra = {}
    if int(st["has_pressure"]):
        extra["pressure"] = {"total": [_f(x) for x in bins["pressure_total"]],
                             "count": [int(x) for x in bins["pressure_count"]]}
    if params.system == S.ISING:
        system = {"Ising": {"E": _f(st["E"]), "N": int(params.n), "S": [int(x) for x in spins],
                            "possible_change": None}}
    else:
        xyz = np.asarray(coords, dtype=np.float64).reshape(3, -1).T
        pos = [[_f(a), _f(b), _f(c)] for a, b, c in xyz]
        if params.system == S.LJ:
            R = _f(st["geom"])
            system = {"Lj": {"E": _f(st["E"]), "error": _f(st["error"]), "possible_change": "None",
                             "positions": pos, "max_radius_squared": R * R, "max_radius": R}}
        else:
            system = {"Wca": {"E": _f(st["E"]), "error": _f(st["error"]), "possible_change": "None",
                              "cell": {"box_diagonal": [_f(st["geom"])] * 3, "positions": pos}}}
    move_plan = ({"TranslationScale": _f(params.move_param)} if params.move_plan == S.TRANSLATION_SCALE
                 else {"AcceptanceRate": _f(params.move_param)})
    return {
        "system": system,
        "method": method,
        "moves": int(st["moves"]),
        "time_L": 0,
        "accepted_moves": int(st["accepted_moves"]),
        "min_allowed_energy": _f(params.min_e) if params.has_min_e else None,
        "max_allowed_energy": _f(params.max_e) if params.has_max_e else None,
        "move_plan": move_plan,
        "translation_scale": _f(st["translation_scale"]),
        "acceptance_rate": _f(st["acceptance_rate"]),
        "rng": {"s0": int(st["rng"][0]), "s1": int(st["rng"][1])},
        "save_as": save_as,
        "bins": {"min": _f(st["bins_min"]), "width": _f(st["bins_width"]),
                 "histogram": [int(x) for x in bins["histogram"][:n]],
                 "t_found": [int(x) for x in bins["t_found"][:n]],
                 "lnw": [_f(x) for x in bins["lnw"][:n]],
                 "energy_total": [_f(x) for x in bins["energy_total"][:n]],
                 "energy_squared_total": [_f(x) for x in bins["energy_squared_total"][:n]],
                 "extra": extra},
        "have_visited_since_maxentropy": [bool(x) for x in bins["visited"][:n]],
        "round_trips": [int(x) for x in bins["round_trips"][:n]],
        "max_S": _f(st["max_S"]),
        "max_S_index": int(st["max_S_index"]),
    }


def dumps(doc):
    return yaml.safe_dump(doc, sort_keys=False)


def loads(text):
    return yaml.safe_load(text)


def from_document(doc, params, cap, e_lo, geom):
    """(walker-state record, physical bin row of length cap, coords or None,
    spins or None, pressure rows or None) from a checkpoint document: the
    first four for Context.upload_raw, the pressure extra's physical rows
    ({"count", "total"}, length cap) for Context.upload_pressure."""
    st = np.zeros((), S.WALKER_STATE_DTYPE)
    b = doc["bins"]
    n = len(b["lnw"])
    width = float(b["width"])
    # any window position that holds the logical bins will do; keep the
    # front margin the initial placement uses
    base = int(np.floor((float(b["min"]) - e_lo) / width)) + 2
    base = max(0, min(base, cap - n))
    if base + n > cap:
        raise ValueError("the checkpoint's bins do not fit the set's bin window")
    method = doc["method"]
    sysdoc = next(iter(doc["system"].values()))
    st["rng"] = [int(doc["rng"]["s0"]), int(doc["rng"]["s1"])]
    st["E"] = float(sysdoc["E"])
    st["error"] = float(sysdoc.get("error", 0.0))
    st["moves"] = int(doc["moves"])
    st["accepted_moves"] = int(doc["accepted_moves"])
    st["acceptance_rate"] = float(doc["acceptance_rate"])
    st["translation_scale"] = float(doc["translation_scale"])
    st["bins_min"] = float(b["min"])
    st["bins_width"] = width
    st["base"] = base
    st["len"] = n
    st["max_S"] = float(doc["max_S"])
    st["max_S_index"] = int(doc["max_S_index"])
    # have_visited_since_maxentropy as epochs: visited bins carry epoch 1
    st["epoch"] = 1
    st["bulk_visited"] = 0
    st["geom"] = geom
    st["cap"] = cap
    if "Sad" in method:
        m = method["Sad"]
        st["method"] = S.SAD
        st["too_lo"], st["too_hi"] = float(m["too_lo"]), float(m["too_hi"])
        st["latest_parameter"] = float(m["latest_parameter"])
        st["min_T"] = float(m["min_T"])
        st["tL"], st["tF"] = int(m["tL"]), int(m["tF"])
        st["num_states"], st["highest_hist"] = int(m["num_states"]), int(m["highest_hist"])
        lo = int(np.floor((st["too_lo"] - st["bins_min"]) / width))
        hi = int(np.floor((st["too_hi"] - st["bins_min"]) / width))
        tf = np.asarray(b["t_found"], dtype=np.uint64)
        st["tmax"] = int(tf[max(lo, 0):min(hi, n - 1) + 1].max()) if n else 0
    elif "Samc" in method:
        st["method"] = S.SAMC
        st["t0"] = float(method["Samc"]["t0"])
    elif "WL" in method:
        m = method["WL"]
        st["method"] = S.INV_T_WL if m["inv_t"] else S.WL
        st["wl_gamma"] = float(m["gamma"])
        st["wl_num_states"] = float(m["num_states"])
        st["wl_min_energy"] = float(m["min_energy"])
        st["wl_lowest_hist"], st["wl_highest_hist"] = int(m["lowest_hist"]), int(m["highest_hist"])
        st["wl_total_hist"] = int(m["total_hist"])
        st["wl_len"] = len(m["hist"])
        st["t0"] = float(params.method_param)
    else:
        st["method"] = S.CANONICAL
    rows = np.zeros(cap, S.BIN_DTYPE)
    sl = slice(base, base + n)
    rows["lnw"][sl] = b["lnw"]
    rows["histogram"][sl] = b["histogram"]
    rows["t_found"][sl] = b["t_found"]
    rows["energy_total"][sl] = b["energy_total"]
    rows["energy_squared_total"][sl] = b["energy_squared_total"]
    rows["round_trips"][sl] = doc["round_trips"]
    rows["visit_epoch"][sl] = [1 if v else 0 for v in doc["have_visited_since_maxentropy"]]
    if "WL" in method:
        h = method["WL"]["hist"]
        rows["wl_hist"][base:base + len(h)] = h
    extra = b.get("extra") or {}
    pressure = None
    if "pressure" in extra:
        st["has_pressure"] = 1
        pressure = {"count": np.zeros(cap, np.uint64), "total": np.zeros(cap)}
        pressure["count"][sl] = extra["pressure"]["count"]
        pressure["total"][sl] = extra["pressure"]["total"]
    coords = spins = None
    if "Ising" in doc["system"]:
        spins = np.asarray(sysdoc["S"], dtype=np.int8)
    else:
        pos = sysdoc["positions"] if "positions" in sysdoc else sysdoc["cell"]["positions"]
        coords = np.asarray(pos, dtype=np.float64).T.reshape(-1)
    return st, rows, coords, spins, pressure


def save_set(ctx, s):
    """Checkpoint documents of every walker of set s."""
    params = ctx.sets[s][0]
    W, _ = ctx.count(s)
    docs = []
    for w in range(W):
        st = ctx.state(s, w)
        bins = ctx.bins(s, w)
        if params.system == S.ISING:
            docs.append(to_document(params, st, bins, spins=ctx.spins(s, w)))
        else:
            docs.append(to_document(params, st, bins, coords=ctx.coords(s, w)))
    return docs


def resume_set(ctx, s, docs):
    """Overwrite every walker of set s (after Context.init) with documents."""
    params = ctx.sets[s][0]
    W, cap = ctx.count(s)
    if len(docs) != W:
        raise ValueError("one document per walker")
    e_lo = S.bin_window(params)[0]
    states = np.zeros(W, S.WALKER_STATE_DTYPE)
    rows = np.zeros((W, cap), S.BIN_DTYPE)
    coords, spins = [], []
    p_count = p_total = None
    if params.system == S.WCA and params.collect_pressure:
        p_count, p_total = np.zeros((W, cap), np.uint64), np.zeros((W, cap))
    cur = ctx.download_raw(s, with_bins=False)[0]
    for w, doc in enumerate(docs):
        st, r, c, sp, pr = from_document(doc, params, cap, e_lo, float(cur[w]["geom"]))
        states[w] = st
        rows[w] = r
        if c is not None:
            coords.append(c)
        if sp is not None:
            spins.append(sp)
        if pr is not None:
            if p_count is None:
                raise ValueError("a pressure extra in the document, but the set does not collect the pressure")
            p_count[w], p_total[w] = pr["count"], pr["total"]
    ctx.upload_raw(s, states, rows, coords=np.concatenate(coords) if coords else None,
                   spins=np.concatenate(spins) if spins else None)
    if p_count is not None:
        ctx.upload_pressure(s, p_count, p_total)


# ----------------------------------------------------------------- CBOR

def _cbor_head(major, n):
    if n < 24:
        return bytes([(major << 5) | n])
    for ai, fmt in ((24, ">B"), (25, ">H"), (26, ">I"), (27, ">Q")):
        if n < 1 << (8 * struct.calcsize(fmt)):
            return bytes([(major << 5) | ai]) + struct.pack(fmt, n)
    raise ValueError("integer out of CBOR's range")


def _cbor_float(x):
    if x != x:
        return b"\xf9\x7e\x00"  # the canonical NaN
    for head, fmt in ((b"\xf9", ">e"), (b"\xfa", ">f")):
        try:
            b = struct.pack(fmt, x)
        except OverflowError:
            continue
        if struct.unpack(fmt, b)[0] == x:
            return head + b
    return b"\xfb" + struct.pack(">d", x)


def _cbor_enc(v, out):
    if v is None:
        out.append(b"\xf6")
    elif v is True:
        out.append(b"\xf5")
    elif v is False:
        out.append(b"\xf4")
    elif isinstance(v, int):
        out.append(_cbor_head(0, v) if v >= 0 else _cbor_head(1, -1 - v))
    elif isinstance(v, float):
        out.append(_cbor_float(v))
    elif isinstance(v, str):
        b = v.encode("utf-8")
        out.append(_cbor_head(3, len(b)))
        out.append(b)
    elif isinstance(v, (list, tuple)):
        out.append(_cbor_head(4, len(v)))
        for x in v:
            _cbor_enc(x, out)
    elif isinstance(v, dict):
        out.append(_cbor_head(5, len(v)))
        for k, x in v.items():
            _cbor_enc(str(k), out)
            _cbor_enc(x, out)
    else:
        raise TypeError(f"no CBOR encoding for {type(v).__name__}")


def cbor_dumps(doc):
    out = []
    _cbor_enc(doc, out)
    return b"".join(out)


def cbor_loads(data):
    pos = 0

    def take(k):
        nonlocal pos
        if pos + k > len(data):
            raise ValueError("truncated CBOR")
        b = data[pos:pos + k]
        pos += k
        return b

    def arg(ai):
        if ai < 24:
            return ai
        if ai in (24, 25, 26, 27):
            return int.from_bytes(take(1 << (ai - 24)), "big")
        if ai == 31:
            return None  # indefinite length
        raise ValueError("reserved CBOR additional information")

    def item():
        ib = take(1)[0]
        major, ai = ib >> 5, ib & 31
        if major == 0:
            return arg(ai)
        if major == 1:
            return -1 - arg(ai)
        if major in (2, 3):
            n = arg(ai)
            if n is None:
                parts = []
                while data[pos] != 0xFF:
                    parts.append(item())
                take(1)
                raw = b"".join(p if isinstance(p, bytes) else p.encode("utf-8") for p in parts)
            else:
                raw = take(n)
            return raw if major == 2 else raw.decode("utf-8")
        if major == 4:
            n = arg(ai)
            if n is None:
                xs = []
                while data[pos] != 0xFF:
                    xs.append(item())
                take(1)
                return xs
            return [item() for _ in range(n)]
        if major == 5:
            n = arg(ai)
            d = {}
            if n is None:
                while data[pos] != 0xFF:
                    k = item()
                    d[k] = item()
                take(1)
                return d
            for _ in range(n):
                k = item()
                d[k] = item()
            return d
        if major == 6:
            arg(ai)
            return item()  # tags carry no meaning in these documents
        if ai == 20:
            return False
        if ai == 21:
            return True
        if ai in (22, 23):
            return None
        if ai == 25:
            return float(struct.unpack(">e", take(2))[0])
        if ai == 26:
            return float(struct.unpack(">f", take(4))[0])
        if ai == 27:
            return struct.unpack(">d", take(8))[0]
        raise ValueError(f"unsupported CBOR simple value {ai}")

    v = item()
    if pos != len(data):
        raise ValueError("trailing bytes after the CBOR item")
    return v


# ---------------------------------------------------------------- files

def save(doc, path):
    """Write one document, in the format the extension names (yaml, json or
    cbor, src/mc/mod.rs:113-119); another extension is an error, as the
    reference panics.  The file is replaced atomically (AtomicFile)."""
    ext = os.path.splitext(str(path))[1].lstrip(".")
    if ext == "yaml":
        data = dumps(doc).encode("utf-8")
    elif ext == "json":
        data = json.dumps(doc).encode("utf-8")
    elif ext == "cbor":
        data = cbor_dumps(doc)
    else:
        raise ValueError(f"I don't know how to create file {path!r}")
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        f.write(data)
    os.replace(tmp, path)


def load(path):
    """Read one document by the extension (src/mc/mod.rs:72-79)."""
    ext = os.path.splitext(str(path))[1].lstrip(".")
    with open(path, "rb") as f:
        data = f.read()
    if ext == "yaml":
        return loads(data.decode("utf-8"))
    if ext == "json":
        return json.loads(data.decode("utf-8"))
    if ext == "cbor":
        return cbor_loads(data)
    raise ValueError(f"I don't know how to read file {path!r}")
