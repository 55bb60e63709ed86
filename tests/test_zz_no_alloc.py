"""GPU: the allocation-free entropy reduction (sadmc_mean_entropy_noalloc,
scratch in a fixed device array) against the allocating one
(sadmc_mean_entropy), both called through the C ABI: two calls in a row,
the second on a smaller set, each bit-identical to the allocating entry
point's result on the same walkers; and a set whose reduction needs more
scratch than the array holds is refused with SADMC_E_CAPACITY while the
allocating entry point still reduces it."""
import ctypes as C

import numpy as np
import pytest

import sadmc_amd as S

pytestmark = pytest.mark.gpu


def _fn(name):
    f = getattr(S.lib(), name)
    f.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_int64, C.POINTER(C.c_double),
                  C.POINTER(C.c_int64), C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _reduce(fn, ctx, sid, max_bins):
    e0, nb = C.c_double(), C.c_int64()
    s = np.full(max_bins, np.nan)
    cnt = np.full(max_bins, -1, np.int64)
    rc = fn(ctx.h, sid, float("nan"), float("nan"), max_bins, C.byref(e0), C.byref(nb),
            s.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p))
    return rc, e0.value, nb.value, s[:nb.value], cnt[:nb.value]


def test_noalloc_entropy_equals_allocating_entry_point(gpu_ctx_factory):
    print("case no-alloc entropy: LJ31 2048 walkers, then LJ13 96 walkers", flush=True)
    alloc, noalloc = _fn("sadmc_mean_entropy"), _fn("sadmc_mean_entropy_noalloc")
    ctx = gpu_ctx_factory()
    big = ctx.add_set(S.make_params(system=S.LJ, n=31, lanes=8, method=S.SAD, method_param=0.05, move_param=0.1,
                                    energy_bin=0.01, has_max_e=1, max_e=0.0), np.arange(2048, dtype=np.uint64),
                      geom=3.0)
    small = ctx.add_set(S.make_params(system=S.LJ, n=13, lanes=4, method=S.SAD, method_param=0.05,
                                      move_param=0.1, energy_bin=0.01, has_max_e=1, max_e=0.0),
                        np.arange(96, dtype=np.uint64) + 5000, geom=2.5)
    ctx.init()
    ctx.run(30000)
    sizes = []
    for sid in (big, small):  # the second call has fewer walkers and bins
        _, cap = ctx.count(sid)
        want = _reduce(alloc, ctx, sid, cap)
        got = _reduce(noalloc, ctx, sid, cap)
        assert want[0] == 0 and got[0] == 0, (want[0], got[0], S.lib().sadmc_last_error(ctx.h))
        assert got[1] == want[1] and got[2] == want[2]
        assert got[3].tobytes() == want[3].tobytes() and np.array_equal(got[4], want[4])
        assert (got[4] > 0).sum() > 10
        sizes.append(got[2])
    assert sizes[1] < sizes[0]
    # the same call again, after the other set's reduction left its contents
    again = _reduce(noalloc, ctx, big, ctx.count(big)[1])
    want = _reduce(alloc, ctx, big, ctx.count(big)[1])
    assert again[3].tobytes() == want[3].tobytes() and np.array_equal(again[4], want[4])


def test_noalloc_entropy_refuses_beyond_its_scratch(gpu_ctx_factory):
    """360000 walkers need 96 bytes of per-walker constants each, more than
    the 32 MiB array: refused, while the allocating entry point reduces the
    same set."""
    print("case no-alloc entropy refusal: 360000 Ising 2x2 walkers", flush=True)
    alloc, noalloc = _fn("sadmc_mean_entropy"), _fn("sadmc_mean_entropy_noalloc")
    ctx = gpu_ctx_factory()
    sid = ctx.add_set(S.make_params(system=S.ISING, n=2, method=S.SAD, method_param=1.0, ising_init=1),
                      np.arange(360000, dtype=np.uint64))
    ctx.init()
    ctx.run(200)
    _, cap = ctx.count(sid)
    rc = _reduce(noalloc, ctx, sid, cap)[0]
    assert rc == -3
    assert b"scratch" in S.lib().sadmc_last_error(ctx.h)
    assert _reduce(alloc, ctx, sid, cap)[0] == 0
