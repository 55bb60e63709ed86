// Training of the learned proposal's network on probe records.
//
// Loss: L = -(1/B) sum_n Jhat_n(s_n) / max_m J_n,m, where s_n is the network
// output (ln sigma, clamped to the grid) for record n's features and Jhat_n
// the linear interpolation of the record's acceptance-weighted squared jumps
// on the ln sigma grid: the mean relative expected squared jump of the
// proposal.  At a clamp the gradient is kept only when it points back into
// the range.
//
// k_prop_grad: one 64-lane workgroup per 64 records.  Each lane runs one
// record forward (BF16 weights converted to float in LDS, FP32 arithmetic)
// and backward to the per-layer deltas, which go to LDS; then each lane sums
// its share of the 1377 parameter gradients (and the loss) over the 64
// records, writing one partial row per workgroup.  No atomics, so the sums
// are deterministic.
// k_prop_update: sums the partial rows in workgroup order, divides by B, and
// either returns the gradient or applies one Adam step to the FP32 master
// weights and refreshes their BF16 copy.
#include "sadmc_kernels.h"

namespace sadmc {

__global__ __launch_bounds__(64) void k_prop_grad(const float *rec, int64_t nrec, int64_t first, int count,
                                                  const uint16_t *wbf, float *partial) {
  __shared__ float lw[kPropLds];
  __shared__ float sh1[64][kPropH + 1], sh2[64][kPropH + 1], sd1[64][kPropH + 1], sd2[64][kPropH + 1];
  __shared__ float sf[64][kPropNF + 1], sd3[64], sl[64];
  load_weights(wbf, lw);
  const int t = threadIdx.x;
  const int64_t idx = (int64_t)blockIdx.x * 64 + t;
  float f[kPropNF], h1[kPropH], h2[kPropH], d2[kPropH];
  float d3 = 0.0f, loss = 0.0f;
#pragma unroll
  for (int k = 0; k < kPropNF; k++) f[k] = 0.0f;
#pragma unroll
  for (int u = 0; u < kPropH; u++) h1[u] = h2[u] = d2[u] = 0.0f;
  if (idx < count) {
    const float *r = rec + ((first + idx) % nrec) * kPropRec;
#pragma unroll
    for (int k = 0; k < kPropNF; k++) f[k] = r[k];
#pragma unroll
    for (int u = 0; u < kPropH; u++) {
      float z = lw[kB1 + u];
#pragma unroll
      for (int k = 0; k < kPropNF; k++) z += lw[kW1 + u * kPropNF + k] * f[k];
      h1[u] = z > 0.0f ? z : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kPropH; u++) {
      float z = lw[kB2 + u];
#pragma unroll
      for (int k = 0; k < kPropH; k++) z += lw[kW2 + u * kPropH + k] * h1[k];
      h2[u] = z > 0.0f ? z : 0.0f;
    }
    float raw = lw[kB3];
#pragma unroll
    for (int u = 0; u < kPropH; u++) raw += lw[kW3 + u] * h2[u];
    // target: interpolated J on the grid, normalised by its maximum
    const double dS = ((double)kPropSMax - (double)kPropSMin) / (double)(kPropNS - 1);
    double mx = 0.0;
#pragma unroll
    for (int k = 0; k < kPropNS; k++) mx = (double)r[kPropNF + k] > mx ? (double)r[kPropNF + k] : mx;
    if (mx > 0.0) {
      const double sc = raw < kPropSMin ? (double)kPropSMin : (raw > kPropSMax ? (double)kPropSMax : (double)raw);
      const double q = (sc - (double)kPropSMin) / dS;
      int k = (int)floor(q);
      k = k < 0 ? 0 : (k > kPropNS - 2 ? kPropNS - 2 : k);
      const double tt = q - (double)k;
      const double j0 = r[kPropNF + k], j1 = r[kPropNF + k + 1];
      loss = (float)(-(j0 * (1.0 - tt) + j1 * tt) / mx);
      const double slope = (j1 - j0) / dS;
      // projected at the clamps: kept only when it points back into range
      if ((raw > kPropSMin && raw < kPropSMax) || (raw >= kPropSMax && slope < 0.0) ||
          (raw <= kPropSMin && slope > 0.0))
        d3 = (float)(-slope / mx);
    }
#pragma unroll
    for (int u = 0; u < kPropH; u++) d2[u] = h2[u] > 0.0f ? d3 * lw[kW3 + u] : 0.0f;
  }
  // d1 = W2^T d2 masked by the first layer's ReLU
#pragma unroll
  for (int k = 0; k < kPropH; k++) {
    float acc = 0.0f;
#pragma unroll
    for (int u = 0; u < kPropH; u++) acc += d2[u] * lw[kW2 + u * kPropH + k];
    sd1[t][k] = h1[k] > 0.0f ? acc : 0.0f;
    sh1[t][k] = h1[k];
    sh2[t][k] = h2[k];
    sd2[t][k] = d2[k];
  }
#pragma unroll
  for (int k = 0; k < kPropNF; k++) sf[t][k] = f[k];
  sd3[t] = d3;
  sl[t] = loss;
  __syncthreads();
  float *row = partial + (int64_t)blockIdx.x * (kPropNP + 1);
  for (int p = t; p <= kPropNP; p += 64) {
    float acc = 0.0f;
    if (p < kB1) {
      const int u = p / kPropNF, k = p % kPropNF;
      for (int n = 0; n < 64; n++) acc += sd1[n][u] * sf[n][k];
    } else if (p < kW2) {
      const int u = p - kB1;
      for (int n = 0; n < 64; n++) acc += sd1[n][u];
    } else if (p < kB2) {
      const int u = (p - kW2) / kPropH, k = (p - kW2) % kPropH;
      for (int n = 0; n < 64; n++) acc += sd2[n][u] * sh1[n][k];
    } else if (p < kW3) {
      const int u = p - kB2;
      for (int n = 0; n < 64; n++) acc += sd2[n][u];
    } else if (p < kB3) {
      const int u = p - kW3;
      for (int n = 0; n < 64; n++) acc += sd3[n] * sh2[n][u];
    } else if (p == kB3) {
      for (int n = 0; n < 64; n++) acc += sd3[n];
    } else {
      for (int n = 0; n < 64; n++) acc += sl[n];
    }
    row[p] = acc;
  }
}

__global__ __launch_bounds__(256) void k_prop_update(const float *partial, int nblk, int count, float *master,
                                                     float *m, float *v, uint16_t *wbf, float lr, int step,
                                                     float *grad_out, float *loss_out) {
  const float inv = 1.0f / (float)count;
  for (int p = threadIdx.x; p <= kPropNP; p += 256) {
    float g = 0.0f;
    for (int b = 0; b < nblk; b++) g += partial[(int64_t)b * (kPropNP + 1) + p];
    g *= inv;
    if (p == kPropNP) {
      if (loss_out) *loss_out = g;
      continue;
    }
    if (grad_out) {
      grad_out[p] = g;
      continue;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(b1, (float)step));
    const float vh = vv / (1.0f - powf(b2, (float)step));
    const float w = master[p] - lr * mh / (sqrtf(vh) + eps);
    master[p] = w;
    wbf[p] = f32_to_bf16(w);
  }
}

hipError_t launch_prop_grad(const float *rec, int64_t nrec, int64_t first, int count, const uint16_t *wbf,
                            float *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_prop_grad, dim3((count + 63) / 64), dim3(64), 0, s, rec, nrec, first, count, wbf, partial);
  return hipGetLastError();
}

hipError_t launch_prop_update(const float *partial, int nblk, int count, float *master, float *m, float *v,
                              uint16_t *wbf, float lr, int step, float *grad_out, float *loss_out, hipStream_t s) {
  hipLaunchKernelGGL(k_prop_update, dim3(1), dim3(256), 0, s, partial, nblk, count, master, m, v, wbf, lr, step,
                     grad_out, loss_out);
  return hipGetLastError();
}

}  // namespace sadmc
