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

// Copy the BF16 weights into LDS as they are (kPropLds halfwords, the
// device array's padded length, as 32-bit words; all lanes of a 64-lane
// workgroup).
__device__ __forceinline__ void load_weights_bf16(const uint16_t *w, uint16_t *lw) {
  const uint32_t *src = (const uint32_t *)w;
  uint32_t *dst = (uint32_t *)lw;
  for (int i = threadIdx.x; i < kPropLds / 2; i += 64) dst[i] = src[i];
  __syncthreads();
}

// The two BF16 values of a 32-bit word as float: the one at the lower
// address (low half) and the other.
__device__ __forceinline__ float bf16_lo(uint32_t v) { return __builtin_bit_cast(float, v << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t v) { return __builtin_bit_cast(float, v & 0xFFFF0000u); }

// Forward pass for one walker whose G lanes all hold the features f; lane l
// computes hidden units l, l+G, ... and the group exchanges the first layer
// through `xc` (32 floats of LDS per walker).  W: the BF16 weights in LDS,
// widened to float in registers (exact), FP32 arithmetic.
__device__ __forceinline__ float mlp_forward_bf16(const uint16_t *W, float *xc, const float *f, int lane, int G) {
  for (int u = lane; u < kPropH; u += G) {
    const uint4 a = *(const uint4 *)(W + kW1 + u * kPropNF);
    float z = bf16_to_f32(W[kB1 + u]);
    z += bf16_lo(a.x) * f[0];
    z += bf16_hi(a.x) * f[1];
    z += bf16_lo(a.y) * f[2];
    z += bf16_hi(a.y) * f[3];
    z += bf16_lo(a.z) * f[4];
    z += bf16_hi(a.z) * f[5];
    z += bf16_lo(a.w) * f[6];
    z += bf16_hi(a.w) * f[7];
    xc[u] = z > 0.0f ? z : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const float4 *x4 = (const float4 *)xc;
  float part = 0.0f;
  for (int u = lane; u < kPropH; u += G) {
    const uint4 *w2 = (const uint4 *)(W + kW2 + u * kPropH);
    float z = bf16_to_f32(W[kB2 + u]);
#pragma unroll
    for (int k = 0; k < kPropH / 8; k++) {
      const uint4 w = w2[k];
      const float4 h0 = x4[2 * k], h1 = x4[2 * k + 1];
      z += bf16_lo(w.x) * h0.x;
      z += bf16_hi(w.x) * h0.y;
      z += bf16_lo(w.y) * h0.z;
      z += bf16_hi(w.y) * h0.w;
      z += bf16_lo(w.z) * h1.x;
      z += bf16_hi(w.z) * h1.y;
      z += bf16_lo(w.w) * h1.z;
      z += bf16_hi(w.w) * h1.w;
    }
    const float h = z > 0.0f ? z : 0.0f;
    part += bf16_to_f32(W[kW3 + u]) * h;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  float s = bf16_to_f32(W[kB3]) + group_sum_f32(part, G);
  return s < kPropSMin ? kPropSMin : (s > kPropSMax ? kPropSMax : s);
}

// The network on the matrix cores, for a wave whose 64 lanes all run this
// call and whose walkers have G >= 4 lanes each (at most 16 walkers: the
// columns of 16x16x4 FP32 products; v_mfma_f32_16x16x4f32 takes A[i][k] from
// lane i + 16k, B[k][j] from lane j + 16k and leaves D[4q + e][j] in element
// e of lane j + 16q).  Layer 1 is H1 = W1 X^T (two row blocks of 16 units, two
// steps of 4 features), layer 2 H2 = W2 H1 (the hidden layer stays where
// layer 1 left it: step (m, e) takes unit 16m + 4q + e from element e of
// lane j + 16q), then s = b3 + w3 . relu(H2 + b2) summed over the four lanes
// of a column.  FP32 products and sums as in mlp_forward_bf16, in another
// order.
// W: the BF16 weights in LDS; f: the walker's features in every lane of its
// group; l: the lane in the wave.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float mlp_wave(const uint16_t *W, const float *f, int l, int G) {
  const int i = l & 15, q = l >> 4, j = l & 15;  // A row / B and D column, step
  const int g = l & (G - 1);
  const bool col = j < 64 / G;  // column j is a walker of this wave (else its source lane is wrapped and unused)
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  // both row blocks' products side by side (independent accumulators)
  f32x4 acc0 = zero, acc1 = zero;
#pragma unroll
  for (int st = 0; st < 2; st++) {
    // feature 4 st + q of walker j, from the lane of its group that holds
    // it at in-group index (4 st + q) mod G
    const float mine = G >= 8 ? f[g & 7] : f[(4 * st + g) & 7];
    const float got = __shfl(mine, (G * j + ((4 * st + q) & (G - 1))) & 63, 64);
    const float b = col ? got : 0.0f;
    const float a0 = bf16_to_f32(W[kW1 + i * kPropNF + 4 * st + q]);
    const float a1 = bf16_to_f32(W[kW1 + (16 + i) * kPropNF + 4 * st + q]);
    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b, acc1, 0, 0, 0);
  }
  float h1[2][4];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const float z0 = acc0[e] + bf16_to_f32(W[kB1 + 4 * q + e]);
    const float z1 = acc1[e] + bf16_to_f32(W[kB1 + 16 + 4 * q + e]);
    h1[0][e] = z0 > 0.0f ? z0 : 0.0f;
    h1[1][e] = z1 > 0.0f ? z1 : 0.0f;
  }
  f32x4 d0 = zero, d1 = zero;
#pragma unroll
  for (int m = 0; m < 2; m++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const float a0 = bf16_to_f32(W[kW2 + i * kPropH + 16 * m + 4 * q + e]);
      const float a1 = bf16_to_f32(W[kW2 + (16 + i) * kPropH + 16 * m + 4 * q + e]);
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, h1[m][e], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, h1[m][e], d1, 0, 0, 0);
    }
  }
  float part = 0.0f;
#pragma unroll
  for (int mo = 0; mo < 2; mo++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int u = 16 * mo + 4 * q + e;
      const float z = (mo == 0 ? d0[e] : d1[e]) + bf16_to_f32(W[kB2 + u]);
      part += bf16_to_f32(W[kW3 + u]) * (z > 0.0f ? z : 0.0f);
    }
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  float s = bf16_to_f32(W[kB3]) + part;
  s = s < kPropSMin ? kPropSMin : (s > kPropSMax ? kPropSMax : s);
  // walker l / G reads its column
  return __shfl(s, l / G, 64);
}

// The network for the walker of this lane's group: on the matrix cores when
// every lane of the wave is here and G >= 4, else mlp_forward_bf16.
__device__ __forceinline__ float mlp_eval(const uint16_t *W, float *xc, const float *f, int lane, int G) {
  if (G >= 4 && __builtin_amdgcn_read_exec() == ~0ull) return mlp_wave(W, f, (int)(threadIdx.x & 63), G);
  return mlp_forward_bf16(W, xc, f, lane, G);
}

}  // namespace sadmc
