// Learned proposal for LJ walkers: device pieces shared by the walker
// engine, the probe kernel and the training kernels.
//
// Definition (no reference counterpart; the same in the CPU restatement):
// - features of atom i at position p: f_k = sum_{j != i} exp(-(r_ij - c_k)^2
//   / (2 w^2)) with c_k = 0.9 + k (1.6/7), w = 0.2, k = 0..7, in float, r_ij
//   the float square root of the float r^2 of the double differences;
// - network 8 -> 32 -> 32 -> 1 with ReLU, BF16 weights and FP32 arithmetic,
//   parameters in the order W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3
//   (1377); output s = ln sigma clamped to [ln 0.01, 0];
// - move: atom uniform, xi ~ N(0, I3) by Box-Muller from four uniform draws,
//   p' = p + sigma(p) xi; ln H = 3 (s - s') + |xi|^2 (1 - e^{2(s - s')}) / 2
//   is the log ratio q(p' -> p) / q(p -> p') of the two Gaussian proposals
//   (sigma' evaluated at p'), added to the acceptance exponent;
// - probe record (training data): the features of a random atom and, for the
//   16 scales ln sigma_m = ln 0.01 + m ln(100)/15, the mean over 2 Gaussian
//   draws of alpha |d|^2 (alpha = the walker's acceptance probability).
#pragma once

#include "sadmc_common.h"

namespace sadmc {

constexpr int kPropNF = 8, kPropH = 32, kPropNP = 1377, kPropNS = 16, kPropR = 2, kPropRec = 24;
constexpr int kPropLds = 1384;  // floats reserved for the weights in LDS
constexpr float kPropC0 = 0.9f, kPropDC = 1.6f / 7.0f, kPropA = 12.5f;
constexpr float kPropSMin = -4.605170186f, kPropSMax = 0.0f;
// parameter offsets
constexpr int kW1 = 0, kB1 = 256, kW2 = 288, kB2 = 1312, kW3 = 1344, kB3 = 1376;

SADMC_HD float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
// Round to nearest, ties to even (finite inputs).
SADMC_HD uint16_t f32_to_bf16(float f) {
  uint32_t b = __builtin_bit_cast(uint32_t, f);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Sum over the G lanes of a group (the same pattern as the FP64 group sum).
__device__ __forceinline__ float group_sum_f32(float v, int G) {
  if (G >= 2) v = v + dpp_f32<0xB1>(v);
  if (G >= 4) v = v + dpp_f32<0x4E>(v);
  if (G >= 8) v = v + dpp_f32<0x141>(v);
  if (G >= 16) v = v + dpp_f32<0x140>(v);
  if (G >= 32) v = v + __shfl_xor(v, 16, 64);
  if (G >= 64) v = v + __shfl_xor(v, 32, 64);
  return v;
}

// Contribution of one neighbour at float distance^2 r2 to the 8 features
// (hardware square root and exponential: the proposal only has to be the
// same function at both ends of a move, not correctly rounded).
__device__ __forceinline__ void feature_add(float r2, float *f) {
  const float r = __builtin_amdgcn_sqrtf(r2);
#pragma unroll
  for (int k = 0; k < kPropNF; k++) {
    const float d = r - (kPropC0 + (float)k * kPropDC);
    f[k] += __expf(-kPropA * d * d);
  }
}

// Forward pass for one walker whose G lanes all hold the features f; lane l
// computes hidden units l, l+G, ... and the group exchanges the first layer
// through `xc` (32 floats of LDS per walker).  W: the weights in LDS.
__device__ __forceinline__ float mlp_forward(const float *W, float *xc, const float *f, int lane, int G) {
  for (int u = lane; u < kPropH; u += G) {
    const float4 *w1 = (const float4 *)(W + kW1 + u * kPropNF);
    const float4 a = w1[0], b = w1[1];
    float z = W[kB1 + u];
    z += a.x * f[0];
    z += a.y * f[1];
    z += a.z * f[2];
    z += a.w * f[3];
    z += b.x * f[4];
    z += b.y * f[5];
    z += b.z * f[6];
    z += b.w * f[7];
    xc[u] = z > 0.0f ? z : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const float4 *x4 = (const float4 *)xc;
  float part = 0.0f;
  for (int u = lane; u < kPropH; u += G) {
    const float4 *w2 = (const float4 *)(W + kW2 + u * kPropH);
    float z = W[kB2 + u];
#pragma unroll
    for (int k = 0; k < kPropH / 4; k++) {
      const float4 w = w2[k], h = x4[k];
      z += w.x * h.x;
      z += w.y * h.y;
      z += w.z * h.z;
      z += w.w * h.w;
    }
    const float h = z > 0.0f ? z : 0.0f;
    part += W[kW3 + u] * h;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  float s = W[kB3] + group_sum_f32(part, G);
  return s < kPropSMin ? kPropSMin : (s > kPropSMax ? kPropSMax : s);
}

// Three standard normal draws by Box-Muller from four uniforms, with the
// hardware log2, square root and sin/cos (whose argument is in revolutions,
// so cos(2 pi u) is one instruction).  The Hastings term is computed from the
// draws actually used.
__device__ __forceinline__ void gaussian3(uint64_t &s0, uint64_t &s1, float *xi) {
  const float u1 = (float)gen_f64(s0, s1), u2 = (float)gen_f64(s0, s1);
  const float u3 = (float)gen_f64(s0, s1), u4 = (float)gen_f64(s0, s1);
  const float m2ln2 = -1.3862943611f;  // -2 ln 2
  const float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(1.0f - u1));
  const float r2 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(1.0f - u3));
  xi[0] = r1 * __builtin_amdgcn_cosf(u2);
  xi[1] = r1 * __builtin_amdgcn_sinf(u2);
  xi[2] = r2 * __builtin_amdgcn_cosf(u4);
}

SADMC_DF double prop_lnh(float s, float s_new, float xi2) {
  const double ds = (double)s - (double)s_new;
  return 3.0 * ds + 0.5 * (double)xi2 * (1.0 - exp(2.0 * ds));
}

// Load the BF16 weights into LDS as float (all 64 lanes of the workgroup).
__device__ __forceinline__ void load_weights(const uint16_t *w, float *lw) {
  for (int i = threadIdx.x; i < kPropNP; i += 64) lw[i] = bf16_to_f32(w[i]);
  __syncthreads();
}

}  // namespace sadmc
