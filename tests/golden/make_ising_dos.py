"""Exact density of states g(E) of the periodic L x L Ising model with the
reference's energy E = +sum_<ij> s_i s_j (src/system/ising.rs:59-75).

Method: row transfer matrix.  With rows r1, r2 as L-bit spin strings,
T[r1, r2] = z^((b(r1, r2) + h(r2) + 2L) / 2), where b is the vertical bond
energy between the rows and h the horizontal (periodic) bond energy inside
r2.  Then Z(z) = Tr T^L = sum_E g(E) z^((E + 2 L^2) / 2), a polynomial of
degree 2 L^2 with non-negative integer coefficients summing to 2^(L^2).
Z is evaluated at the integer points z = 1 .. 2 L^2 + 1 modulo several
primes (exact integer matrix products in float64 with entries < 2^22),
interpolated (Lagrange, exact modular arithmetic) and recombined by CRT.

Writes ising{L}_dos.json: {"L": L, "energies": [...], "counts": [...]}
(decimal strings for the counts).  Checks before writing: the counts sum to
2^(L^2), g(E) = g(-E), and for L <= 4 they equal brute-force enumeration.
"""
import json
import os
import sys

import numpy as np

PRIMES = [4194301, 4194287, 4194277, 4194271]  # < 2^22: 256 * p^2 < 2^53


def is_prime(n):
    if n < 2:
        return False
    i = 2
    while i * i <= n:
        if n % i == 0:
            return False
        i += 1
    return True


def row_tables(L):
    rows = np.arange(1 << L)
    spins = np.where((rows[:, None] >> np.arange(L)) & 1, 1, -1)        # [r, i]
    h = (spins * np.roll(spins, -1, axis=1)).sum(1)                       # horizontal bonds in a row
    b = spins @ spins.T                                                   # vertical bonds between rows
    return b, h


def z_poly_mod(L, p):
    b, h = row_tables(L)
    expo = (b + h[None, :] + 2 * L) // 2                                  # 0 .. 2L per row step
    deg = 2 * L * L
    vals = []
    for z in range(1, deg + 2):
        zp = np.array([pow(z, k, p) for k in range(2 * L + 1)], dtype=np.float64)
        T = zp[expo]
        M = np.eye(1 << L)
        P = T.copy()
        e = L
        while e:
            if e & 1:
                M = np.fmod(M @ P, p)
            P = np.fmod(P @ P, p)
            e >>= 1
        vals.append(int(np.trace(M)) % p)
    return vals


def interpolate_mod(vals, p):
    """Coefficients of the degree-(n-1) polynomial through (k, vals[k-1]),
    k = 1..n, modulo p."""
    n = len(vals)
    xs = list(range(1, n + 1))
    coef = [0] * n
    for i, (xi, yi) in enumerate(zip(xs, vals)):
        # basis polynomial prod_{j != i} (x - xj) / (xi - xj)
        num = [1]
        den = 1
        for j, xj in enumerate(xs):
            if j == i:
                continue
            num = [(a - xj * bb) % p for a, bb in zip([0] + num, num + [0])]
            den = den * (xi - xj) % p
        s = yi * pow(den, p - 2, p) % p
        for k in range(n):
            coef[k] = (coef[k] + s * num[k]) % p
    return coef


def crt(residues, primes):
    x, m = 0, 1
    for r, p in zip(residues, primes):
        t = (r - x) * pow(m, p - 2, p) % p
        x += m * t
        m *= p
    return x


def exact_dos(L):
    assert all(is_prime(p) for p in PRIMES)
    per_prime = [interpolate_mod(z_poly_mod(L, p), p) for p in PRIMES]
    counts = [crt([c[k] for c in per_prime], PRIMES) for k in range(2 * L * L + 1)]
    energies = [2 * k - 2 * L * L for k in range(2 * L * L + 1)]
    return energies, counts


def brute_dos(L):
    n = L * L
    s = np.where((np.arange(1 << n)[:, None] >> np.arange(n)) & 1, 1, -1).reshape(-1, L, L)
    E = (s * np.roll(s, -1, axis=2)).sum((1, 2)) + (s * np.roll(s, -1, axis=1)).sum((1, 2))
    vals, cnt = np.unique(E, return_counts=True)
    return dict(zip(vals.tolist(), cnt.tolist()))


def main(L=8, out_dir=None):
    energies, counts = exact_dos(L)
    assert sum(counts) == 2 ** (L * L)
    d = dict(zip(energies, counts))
    assert all(d[e] == d[-e] for e in energies)
    if L <= 4:
        bf = brute_dos(L)
        assert all(d.get(e, 0) == c for e, c in bf.items()) and sum(bf.values()) == sum(counts)
    keep = [(e, c) for e, c in zip(energies, counts) if c]
    out = {"L": L, "energy_sign": "E = +sum_<ij> s_i s_j (ising.rs:59-75)",
           "energies": [e for e, _ in keep], "counts": [str(c) for _, c in keep]}
    out_dir = out_dir or os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(out_dir, f"ising{L}_dos.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    return path


if __name__ == "__main__":
    print(main(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
