"""Pins of the oracle's random-number stream and sampling adapters.

- SplitMix64 against its published seed-0 output vector.
- The xoroshiro128 linear engine (rng.rs:48-57, a=24 b=16 c=37) against
  facts fixed by the mathematics: its GF(2) transition matrix has full
  period 2^128 - 1 (M^(2^128-1) = I, and M^((2^128-1)/q) != I for every
  prime q dividing 2^128 - 1), and the published jump polynomial of
  xoroshiro128+ applied to M equals M^(2^64).  A wrong shift or rotation
  constant fails these.
- The rand 0.7 integer adapters against their defining properties: exact
  uniformity of Uniform::new(0, n) over the accepted zone, and the
  power-of-two behaviour of the conservative gen_range zone.
"""
import numpy as np

import oracle as O

MASK64 = (1 << 64) - 1


def test_splitmix64_published_vector():
    import ctypes as C
    x = C.c_uint64(0)
    got = [O.lib().or_splitmix64_next(C.byref(x)) for _ in range(4)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_seed_from_u64_is_two_splitmix_outputs():
    import ctypes as C
    for seed in (0, 1, 10137, 2 ** 63 + 5):
        x = C.c_uint64(seed)
        a = O.lib().or_splitmix64_next(C.byref(x))
        b = O.lib().or_splitmix64_next(C.byref(x))
        assert O.RNG(seed).state == (a, b)


def _step_state(v):
    """One oracle step applied to a 128-bit state vector (bits 0-63 = s0)."""
    r = O.RNG(state=(v & MASK64, v >> 64))
    r.next_u64()
    s0, s1 = r.state
    return s0 | (s1 << 64)


def _matrix():
    # column k = image of basis vector e_k (the step is linear over GF(2))
    cols = [_step_state(1 << k) for k in range(128)]
    # store as rows: row i has bit k set if column k has bit i
    rows = [0] * 128
    for k, c in enumerate(cols):
        for i in range(128):
            if (c >> i) & 1:
                rows[i] |= 1 << k
    return rows


def _mul(A, B):
    # (A B) row i = XOR of rows of B selected by the bits of A's row i
    out = []
    for a in A:
        acc = 0
        k = 0
        while a:
            if a & 1:
                acc ^= B[k]
            a >>= 1
            k += 1
        out.append(acc)
    return out


def _pow(M, e):
    R = [1 << i for i in range(128)]
    P = M
    while e:
        if e & 1:
            R = _mul(R, P)
        P = _mul(P, P)
        e >>= 1
    return R


IDENT = [1 << i for i in range(128)]


def test_xoroshiro_engine_is_linear():
    rng = np.random.default_rng(1)
    for _ in range(8):
        a = int(rng.integers(0, 2 ** 63)) | (int(rng.integers(0, 2 ** 63)) << 64)
        b = int(rng.integers(0, 2 ** 63)) | (int(rng.integers(0, 2 ** 63)) << 64)
        assert _step_state(a ^ b) == _step_state(a) ^ _step_state(b)


def test_xoroshiro_engine_full_period():
    M = _matrix()
    order = (1 << 128) - 1
    assert _pow(M, order) == IDENT
    # 2^128 - 1 = F0 F1 ... F6 (Fermat numbers); F5 = 641 * 6700417,
    # F6 = 274177 * 67280421310721
    primes = (3, 5, 17, 257, 65537, 641, 6700417, 274177, 67280421310721)
    prod = 1
    for q in primes:
        prod *= q
    assert prod == order
    for q in primes:
        assert order % q == 0
        assert _pow(M, order // q) != IDENT


def test_xoroshiro128plus_jump_polynomial():
    M = _matrix()
    jump = [0xDF900294D8F554A5, 0x170865DF4B3201FC]  # xoroshiro128+ (24, 16, 37) jump, 2^64 steps
    J = [0] * 128
    P = IDENT
    for word in jump:
        for bit in range(64):
            if (word >> bit) & 1:
                J = [x ^ y for x, y in zip(J, P)]
            P = _mul(P, M)
    assert J == _pow(M, 1 << 64)


def test_output_is_sum_of_state():
    r = O.RNG(5)
    s0, s1 = r.state
    assert r.next_u64() == (s0 + s1) & MASK64


def test_gen_f64_53_bit_grid():
    r = O.RNG(3)
    for _ in range(1000):
        s0, s1 = r.state
        v = r.gen_f64()
        assert v == (((s0 + s1) & MASK64) >> 11) * 2.0 ** -53
        assert 0.0 <= v < 1.0


def test_uniform_index_exact_zone_uniformity():
    # For n dividing 2^64 there is no rejection and the result is the high
    # bits; for other n the accepted zone holds an exact multiple of n values.
    r = O.RNG(11)
    for _ in range(200):
        s0, s1 = r.state
        v = (s0 + s1) & MASK64
        got = r.uniform_index(16)
        assert got == v >> 60
    n = 31
    zone = MASK64 - ((2 ** 64) % n)
    assert (zone + 1) % n == 0
    counts = np.bincount([r.uniform_index(n) for _ in range(62000)], minlength=n)
    chi2 = ((counts - 2000) ** 2 / 2000).sum()
    assert chi2 < 80  # 30 dof; p ~ 1e-6


def test_gen_range_power_of_two_rejects_half():
    # zone = (8 << 60) - 1: accepted iff bit 60 of the draw is clear, and the
    # result is then the top three bits.
    r = O.RNG(7)
    for _ in range(300):
        probe = O.RNG(state=r.state)
        draws = []
        while True:
            v = probe.next_u64()
            draws.append(v)
            if not (v >> 60) & 1:
                break
        got = r.gen_range(0, 8)
        assert got == draws[-1] >> 61
        assert r.state == probe.state


def test_uniform_f64_range_and_grid():
    r = O.RNG(9)
    xs = np.array([r.uniform_f64(-0.1, 0.1) for _ in range(20000)])
    assert xs.min() >= -0.1 and xs.max() < 0.1
    assert abs(xs.mean()) < 0.003
    assert abs(xs.var() - 0.2 ** 2 / 12) < 2e-4
