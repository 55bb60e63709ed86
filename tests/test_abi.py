"""The C-ABI library on the CPU: it loads without a GPU, exports every
function include/sadmc.h declares, its struct layouts match the Python
mirrors, its host-side initial configurations equal the oracle's, and it
refuses to run without a gfx950 device instead of falling back to the CPU."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
import sadmc_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sadmc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sadmc_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = S.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(S.EXPORTS)
    assert lib.sadmc_abi_version() == 1


def test_struct_layouts_match_python_mirrors():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "sadmc.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(sadmc_walker_state), sizeof(sadmc_bin), sizeof(sadmc_set_params),
         offsetof(sadmc_walker_state, cap), offsetof(sadmc_set_params, move_kind));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(sadmc_replica), sizeof(sadmc_bwalker_state), sizeof(sadmc_bbin),
         sizeof(sadmc_bcounts), offsetof(sadmc_bwalker_state, pressure), offsetof(sadmc_bwalker_state, min_gamma));
  printf("%zu %zu %zu %zu\n", sizeof(sadmc_zreplica), offsetof(sadmc_zreplica, pressure_count),
         sizeof(sadmc_hires), offsetof(sadmc_set_params, high_resolution_de));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sizes.c")
        exe = os.path.join(d, "sizes")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    ws, b, sp, cap_off, mk_off = map(int, out[:5])
    rep, bws, bb, bc, pr_off, mg_off = map(int, out[5:11])
    zr, zpc_off, hr, hr_off = map(int, out[11:])
    assert hr == S.HIRES_DTYPE.itemsize == 80
    assert hr_off == S.SetParams.high_resolution_de.offset
    assert zr == S.ZREPLICA_DTYPE.itemsize
    assert zpc_off == S.ZREPLICA_DTYPE.fields["pressure_count"][1]
    assert rep == S.REPLICA_DTYPE.itemsize
    assert bws == S.BWALKER_STATE_DTYPE.itemsize and bb == S.BBIN_DTYPE.itemsize == 80
    assert bc == S.BCOUNTS_DTYPE.itemsize == 64
    assert pr_off == S.BWALKER_STATE_DTYPE.fields["pressure"][1]
    assert mg_off == S.BWALKER_STATE_DTYPE.fields["min_gamma"][1]
    # the binning walker's state starts with the oracle's record, field for field
    import oracle as Orc
    for name in Orc.BSTATE_DTYPE.names:
        assert S.BWALKER_STATE_DTYPE.fields[name][1] == Orc.BSTATE_DTYPE.fields[name][1], name
    assert ws == S.WALKER_STATE_DTYPE.itemsize
    assert b == S.BIN_DTYPE.itemsize == 64
    assert sp == __import__("ctypes").sizeof(S.SetParams)
    assert cap_off == S.WALKER_STATE_DTYPE.fields["cap"][1]
    assert mk_off == S.SetParams.move_kind.offset


@pytest.mark.parametrize("system,n,geom", [(S.LJ, 13, 2.5), (S.LJ, 31, 3.0), (S.LJ, 38, 3.2),
                                           (S.WCA, 32, 3.5), (S.WCA, 64, 4.2), (S.WCA, 128, 5.5),
                                           (S.WCA, 256, 6.8)])
def test_host_initial_configurations_equal_oracle(system, n, geom):
    for seed in (0, 1, 17, 2 ** 40 + 3):
        got = S.init_coords(system, n, geom, seed)
        want = O.lj_init(n, geom, seed) if system == S.LJ else O.wca_init(n, geom, seed)
        np.testing.assert_array_equal(got, want)


def test_host_initial_spins_equal_oracle():
    for rule in (0, 1):
        for seed in (0, 5, 99):
            np.testing.assert_array_equal(S.init_spins(8, rule, seed), O.ising_spins(8, rule, seed))


def test_bin_window_covers_the_configurations():
    lo, hi, cap = S.bin_window(S.make_params(system=S.ISING, n=8, energy_bin=0.0))
    assert lo == -128 and hi == 128 and cap >= 64
    lo, hi, cap = S.bin_window(S.make_params(system=S.LJ, n=31, energy_bin=0.01, has_max_e=1, max_e=0.0))
    assert lo < -133.586422 and hi == 0.0 and cap >= (hi - lo) / 0.01


def test_no_cpu_fallback_without_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU node is present")
    with pytest.raises(S.SadmcError) as e:
        S.Context(0)
    assert e.value.code == -7


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run(["python3", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_system_info_matches_the_reference_trait_constants():
    """System trait constants (src/system/mod.rs:54-98) per system:
    lowest_possible_energy (lj.rs:245-248, wca.rs:234-236; none for Ising),
    delta_energy (ising.rs:76-78), min_moves_to_randomize / dimensionality
    (ising.rs:86-91, lj.rs:280-285, wca.rs:268-273), max_size (ising.rs:120-122,
    lj.rs:375-377, wca.rs:354-356)."""
    i = S.system_info(S.make_params(system=S.ISING, n=8))
    assert i == dict(lowest_possible_energy=None, delta_energy=4.0, min_moves_to_randomize=64, dimensionality=64,
                     max_size=0.5)
    lj = S.system_info(S.make_params(system=S.LJ, n=31), 3.0)
    assert lj == dict(lowest_possible_energy=-465.0, delta_energy=None, min_moves_to_randomize=31,
                      dimensionality=93, max_size=3.0)
    w = S.system_info(S.make_params(system=S.WCA, n=64), 4.0)
    assert w["lowest_possible_energy"] == 0.0 and w["delta_energy"] is None
    assert (w["min_moves_to_randomize"], w["dimensionality"]) == (64, 192)
    assert w["max_size"] == np.sqrt(48.0)
    with pytest.raises(S.SadmcError):
        S.system_info(S.make_params(system=S.ISING, n=9))


def test_rust_binding_declares_every_function_with_its_arity():
    """The Rust extern block a maintainer would add names
    every function of include/sadmc.h with the same number of arguments."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = doc[doc.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    rust = {m.group(1): m.group(2) for m in re.finditer(r"pub fn (sadmc_\w+)\((.*?)\)\s*(?:->|;)", block, re.S)}
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    c = {m.group(1): m.group(2) for m in re.finditer(r"\b(sadmc_\w+)\s*\(([^;{]*?)\)\s*;", text, re.S)}
    assert set(rust) == set(c), (sorted(set(c) - set(rust)), sorted(set(rust) - set(c)))
    arity = lambda args: 0 if args.strip() in ("", "void") else args.count(",") + 1
    for name in c:
        assert arity(rust[name]) == arity(c[name]), name


def _c_struct_types(name):
    """Field types of a header struct in order, arrays expanded, nested
    structs by name."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = {m.group(2): m.group(1) for m in re.finditer(r"typedef struct(?: \w+)? \{(.*?)\}\s*(\w+);", text, re.S)}[name]
    out = []
    for decl in (d.strip() for d in body.split(";") if d.strip()):
        ty, rest = decl.split(None, 1)
        for v in (x.strip() for x in rest.split(",")):
            m = re.match(r"\w+(?:\[(\d+)\])?$", v)
            out += [ty] * int(m.group(1) or 1)
    return out


def _rust_struct_types(doc, name):
    body = re.search(r"pub struct " + name + r" \{(.*?)\n\}", doc, re.S).group(1)
    body = re.sub(r"//[^\n]*", "", body)
    out = []
    for ty in re.findall(r"pub \w+: ([^,]+?)\s*(?:,|$)", body):
        m = re.match(r"\[(\w+); (\d+)\]$", ty)
        out += [m.group(1)] * int(m.group(2)) if m else [ty]
    return out


def test_rust_structs_mirror_the_header():
    """Each #[repr(C)] struct of the Rust binding has its C struct's field
    types in the same order (so the same layout)."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    to_rust = {"int32_t": "i32", "int64_t": "i64", "uint64_t": "u64", "double": "f64", "sadmc_bcounts": "BCounts"}
    pairs = {"SetParams": "sadmc_set_params", "WalkerState": "sadmc_walker_state", "Bin": "sadmc_bin",
             "SystemInfo": "sadmc_system_info_t", "BCounts": "sadmc_bcounts", "BWalkerState": "sadmc_bwalker_state",
             "BBin": "sadmc_bbin", "HiRes": "sadmc_hires", "ZReplica": "sadmc_zreplica", "Replica": "sadmc_replica"}
    for rust, c in pairs.items():
        assert _rust_struct_types(doc, rust) == [to_rust[t] for t in _c_struct_types(c)], rust
