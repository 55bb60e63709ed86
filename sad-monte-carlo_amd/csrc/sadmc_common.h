// Numerics shared by the host side and the gfx950 kernels: the walker RNG
// stream, the rand 0.7 sampling adapters the reference calls, and the pair
// potentials.  Everything here is compiled with -ffp-contract=off so that
// each expression rounds exactly as written (Rust never contracts), which is
// what makes device results reproducible bit-for-bit on a CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SADMC_HD __host__ __device__ __forceinline__
#define SADMC_DF __device__ __forceinline__

namespace sadmc {

// Seed tweak for the initial-configuration stream (LJ/WCA positions, per-walker
// Ising spins), independent of the MC stream that seed_from_u64(seed) drives.
constexpr uint64_t kInitTag = 0xA5F0C3E1D2B49687ULL;

SADMC_HD double bits_f64(uint64_t b) { return __builtin_bit_cast(double, b); }
SADMC_HD uint64_t f64_bits(double d) { return __builtin_bit_cast(uint64_t, d); }
SADMC_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// SplitMix64 step (rand_xoshiro 0.4 SplitMix64::next_u64).
SADMC_HD uint64_t splitmix_next(uint64_t &x) {
  x += 0x9e3779b97f4a7c15ULL;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Xoroshiro128Plus::seed_from_u64 (called at src/mc/energy.rs:835): the state
// is the first two SplitMix64 outputs of `seed`.
SADMC_HD void rng_seed(uint64_t seed, uint64_t &s0, uint64_t &s1) {
  uint64_t x = seed;
  s0 = splitmix_next(x);
  s1 = splitmix_next(x);
}

// xoroshiro128+ next_u64, src/rng.rs:48-57.
SADMC_HD uint64_t rng_next(uint64_t &s0, uint64_t &s1) {
  uint64_t a = s0, b = s1;
  uint64_t r = a + b;
  b ^= a;
  s0 = rotl64(a, 24) ^ b ^ (b << 16);
  s1 = rotl64(b, 37);
  return r;
}

// The 2^64-step jump of xoroshiro128+ (the published jump polynomial), as
// rand_xoshiro's jump() (tempering.rs:158).
SADMC_HD void rng_jump(uint64_t &s0, uint64_t &s1) {
  const uint64_t J[2] = {0xDF900294D8F554A5ULL, 0x170865DF4B3201FCULL};
  uint64_t a = 0, b = 0;
  for (int w = 0; w < 2; w++)
    for (int k = 0; k < 64; k++) {
      if ((J[w] >> k) & 1ULL) {
        a ^= s0;
        b ^= s1;
      }
      rng_next(s0, s1);
    }
  s0 = a;
  s1 = b;
}

// rng.gen_range(low, high) (src/system/ising.rs:105-106): rand 0.7
// UniformInt::sample_single, conservative zone, widening multiply.
SADMC_HD uint64_t gen_range(uint64_t &s0, uint64_t &s1, uint64_t low, uint64_t high) {
  uint64_t range = high - low;
  uint64_t zone = (range << __builtin_clzll(range)) - 1;
  for (;;) {
    uint64_t v = rng_next(s0, s1);
    unsigned __int128 m = (unsigned __int128)v * range;
    if ((uint64_t)m <= zone) return low + (uint64_t)(m >> 64);
  }
}

// rng.sample(Uniform::new(0, n)) (src/system/lj.rs:368, wca.rs:345): exact
// zone MAX - (2^64 mod n).
SADMC_HD uint64_t uniform_index(uint64_t &s0, uint64_t &s1, uint64_t n) {
  uint64_t zone = ~0ULL - ((~0ULL - n + 1) % n);
  for (;;) {
    uint64_t v = rng_next(s0, s1);
    unsigned __int128 m = (unsigned __int128)v * n;
    if ((uint64_t)m <= zone) return (uint64_t)(m >> 64);
  }
}

// rng.gen::<f64>() (src/mc/energy.rs:465,489,498,508).
SADMC_HD double gen_f64(uint64_t &s0, uint64_t &s1) {
  uint64_t v = rng_next(s0, s1);
  return (1.0 / 9007199254740992.0) * (double)(v >> 11);
}

// Uniform::new(low, high) for f64, as rand 0.7's UniformFloat.
SADMC_HD double uniform_scale(double low, double high) {
  const double max_rand = bits_f64((~0ULL >> 12) | 0x3FF0000000000000ULL) - 1.0;
  double scale = high - low;
  while (scale * max_rand + low >= high) scale = bits_f64(f64_bits(scale) - 1);
  return scale;
}
SADMC_HD double uniform_sample(uint64_t &s0, uint64_t &s1, double low, double scale) {
  uint64_t v = rng_next(s0, s1) >> 12;
  double v12 = bits_f64(v | 0x3FF0000000000000ULL);
  return (v12 - 1.0) * scale + low;
}
SADMC_HD double uniform_f64(uint64_t &s0, uint64_t &s1, double low, double high) {
  return uniform_sample(s0, s1, low, uniform_scale(low, high));
}

// src/system/lj.rs:78-81 with sigma = epsilon = 1.  powi(s, 6) and powi(s, 3)
// expand as binary powering: s^6 = s^2 (s^2)^2, s^3 = s s^2.
SADMC_HD double lj_pair(double r2) {
  double s = 1.0 / r2;
  double s2 = s * s;
  double s4 = s2 * s2;
  double s6 = s2 * s4;
  double s3 = s * s2;
  return 4.0 * (s6 - s3);
}

// The same pair term in binary32 (precision variant).
SADMC_HD float lj_pair_f32(float r2) {
  float s = 1.0f / r2;
  float s2 = s * s;
  float s4 = s2 * s2;
  float s6 = s2 * s4;
  float s3 = s * s2;
  return 4.0f * (s6 - s3);
}

// 1/x by the hardware reciprocal and one FMA Newton step.  For every x in
// [2^-126, 2^126) this is the correctly rounded reciprocal, bit-identical to
// the IEEE quotient 1.0f / x (an exhaustive comparison over all 2^31
// non-negative binary32 patterns on gfx950, sadmc_check_recip_f32, finds no
// difference there); below 2^-126 (denormals, zero), from 2^126 up and at
// +inf it is not (the pair loops that use it handle those).
SADMC_DF float recip_f32_fast(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(y, __builtin_fmaf(-x, y, 1.0f), y);
}

// lj_pair_f32 with recip_f32_fast: the same value for r2 in [2^-126, 2^126).
SADMC_DF float lj_pair_f32_fast(float r2) {
  float s = recip_f32_fast(r2);
  float s2 = s * s;
  float s4 = s2 * s2;
  float s6 = s2 * s4;
  float s3 = s * s2;
  return 4.0f * (s6 - s3);
}

// src/system/wca.rs:66-76.
SADMC_HD double wca_pair(double r2, double rc2) {
  if (r2 < rc2) {
    double s = 1.0 / r2;
    double s2 = s * s;
    double s4 = s2 * s2;
    double s6 = s2 * s4;
    double s3 = s * s2;
    return 4.0 * (s6 - s3) + 1.0;
  }
  return 0.0;
}

// src/system/wca.rs:79-92.
SADMC_HD double wca_pair_pressure(double r2, double rc2) {
  if (r2 < rc2) {
    double s = 1.0 / r2;
    double s2 = s * s;
    double s4 = s2 * s2;
    double s6 = s2 * s4;
    double s3 = s * s2;
    return 12.0 * (2.0 * s6 - s3);
  }
  return 0.0;
}

SADMC_HD double norm2(double x, double y, double z) { return x * x + y * y + z * z; }

// Minimum-image component for two positions inside [0, L).
SADMC_HD double min_image(double d, double L, double halfL) {
  if (d > halfL) return d - L;
  if (d < -halfL) return d + L;
  return d;
}

// src/system/optcell.rs:277-309.
SADMC_HD double put_in_cell(double r, double L) {
  if (r < 0.0) {
    do {
      r += L;
    } while (r < 0.0);
  } else {
    while (r >= L) r -= L;
  }
  return r;
}

}  // namespace sadmc
