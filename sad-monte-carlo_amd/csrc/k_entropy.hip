// Mean entropy S(E) of a walker set, reduced on the device.
//
// Per walker (plotting/parse-binning.py:150-171): inside [too_lo, too_hi]
// S = ln w; below too_lo S = ln w(i_lo) + (E - too_lo)/min_T + ln(H/<H>),
// above too_hi S = ln w(i_hi) + ln(H/<H>), <H> the mean histogram over
// [i_lo, i_hi] and i_lo, i_hi the bins whose centres are nearest too_lo and
// too_hi.  Bins outside the range with H = 0 carry no estimate and are
// skipped (the reference's script sets them to 0 for its one walker).  Each
// walker's free constant is removed by subtracting its mean S over a common
// energy window [ga, gb) of global bins (a walker without an estimate in
// every window bin is left out), then S is averaged over the walkers
// covering each bin.
//
// k_entropy_walker: one wavefront per walker computes i_lo, i_hi, <H> and the
// window mean.  k_entropy_bins: one lane per global bin loops over the
// walkers; consecutive lanes read consecutive bins of the same walker, so
// every load of the 16-byte hot records is coalesced.
#include "sadmc_kernels.h"

namespace sadmc {

struct EntropyWalker {
  double too_lo, too_hi, min_T, lnw_lo, lnw_hi, mean_hist, c, bmin;
  int64_t off, base, len;
  int32_t ok, pad;
};

__global__ __launch_bounds__(64) void k_entropy_walker(const sadmc_walker_state *ws, const HotBin *hot, int64_t cap,
                                                       double width, double gmin, int64_t ga, int64_t gb,
                                                       EntropyWalker *out) {
  const int w = blockIdx.x;
  const sadmc_walker_state &s = ws[w];
  const int64_t len = s.len, base = s.base;
  const double bmin = s.bins_min;
  const HotBin *H = hot + (int64_t)w * cap + base;
  const int64_t off = (int64_t)llround((bmin - gmin) / width);
  if (len <= 0 || s.status != 0 || s.method != SADMC_SAD) {
    if (threadIdx.x == 0) {
      EntropyWalker e = {};
      e.ok = 0;
      out[w] = e;
    }
    return;
  }
  // bin nearest a given energy: first index of the smallest |E_i - e|
  auto nearest = [&](double e) {
    int64_t i = (int64_t)floor((e - bmin) / width);
    i = i < 0 ? 0 : (i > len - 1 ? len - 1 : i);
    int64_t best = i;
    double bd = fabs(bmin + ((double)i + 0.5) * width - e);
    for (int64_t j = i - 1; j <= i + 1; j++) {
      if (j < 0 || j >= len) continue;
      double d = fabs(bmin + ((double)j + 0.5) * width - e);
      if (d < bd || (d == bd && j < best)) {
        bd = d;
        best = j;
      }
    }
    return best;
  };
  const int64_t ilo = nearest(s.too_lo), ihi = nearest(s.too_hi);
  // four loads in flight per lane; the additions keep the i order
  double hsum = 0.0;
  for (int64_t i0 = ilo + threadIdx.x; i0 <= ihi; i0 += 256) {
    uint64_t h[4];
#pragma unroll
    for (int u = 0; u < 4; u++) h[u] = i0 + 64 * u <= ihi ? H[i0 + 64 * u].hist : 0;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + 64 * u <= ihi) hsum += (double)h[u];
  }
  for (int d = 32; d >= 1; d >>= 1) hsum += __shfl_xor(hsum, d, 64);
  const double mean_hist = hsum / (double)(ihi - ilo + 1);
  const double lnw_lo = H[ilo].lnw, lnw_hi = H[ihi].lnw;
  // the window mean of the extended S; every window bin must carry an estimate
  double csum = 0.0;
  int64_t ccount = 0;
  for (int64_t g = ga + threadIdx.x; g < gb; g += 64) {
    const int64_t i = g - off;
    if (i < 0 || i >= len) continue;
    const HotBin b = H[i];
    const double E = bmin + ((double)i + 0.5) * width;
    if (E < s.too_lo) {
      if (b.hist == 0) continue;
      csum += lnw_lo + (E - s.too_lo) / s.min_T + log((double)b.hist / mean_hist);
    } else if (E > s.too_hi) {
      if (b.hist == 0) continue;
      csum += lnw_hi + log((double)b.hist / mean_hist);
    } else {
      csum += b.lnw;
    }
    ccount += 1;
  }
  for (int d = 32; d >= 1; d >>= 1) {
    csum += __shfl_xor(csum, d, 64);
    ccount += __shfl_xor(ccount, d, 64);
  }
  if (threadIdx.x == 0) {
    EntropyWalker e;
    e.too_lo = s.too_lo;
    e.too_hi = s.too_hi;
    e.min_T = s.min_T;
    e.lnw_lo = lnw_lo;
    e.lnw_hi = lnw_hi;
    e.mean_hist = mean_hist;
    e.c = ccount > 0 ? csum / (double)ccount : 0.0;
    e.bmin = bmin;
    e.off = off;
    e.base = base;
    e.len = len;
    e.ok = (s.status == 0 && s.method == SADMC_SAD && ccount == gb - ga && len > 0) ? 1 : 0;
    e.pad = 0;
    out[w] = e;
  }
}

// One lane per grid bin and one block row per chunk of walkers: partial
// sums over the chunk's walkers, so that the grid fills the device.
__global__ __launch_bounds__(256) void k_entropy_bins(const EntropyWalker *ew, int W, const HotBin *hot, int64_t cap,
                                                      double width, int64_t ng, int chunks, double *psum,
                                                      int64_t *pcnt) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ng) return;
  const int c = blockIdx.y;
  const int w0 = (int)((int64_t)W * c / chunks), w1 = (int)((int64_t)W * (c + 1) / chunks);
  double sum = 0.0;
  int64_t n = 0;
  for (int w4 = w0; w4 < w1; w4 += 4) {
    // the records of up to four walkers in flight, used in walker order
    HotBin bq[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int w = w4 + u;
      bq[u].lnw = 0.0;
      bq[u].hist = 0;
      if (w < w1) {
        const int64_t i = g - ew[w].off;
        if (ew[w].ok && i >= 0 && i < ew[w].len) bq[u] = hot[(int64_t)w * cap + ew[w].base + i];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int w = w4 + u;
      if (w >= w1) continue;
      const EntropyWalker e = ew[w];
      const int64_t i = g - e.off;
      if (!e.ok || i < 0 || i >= e.len) continue;
      const HotBin b = bq[u];
      const double E = e.bmin + ((double)i + 0.5) * width;
      double v;
      if (E < e.too_lo) {
        if (b.hist == 0) continue;
        v = e.lnw_lo + (E - e.too_lo) / e.min_T + log((double)b.hist / e.mean_hist);
      } else if (E > e.too_hi) {
        if (b.hist == 0) continue;
        v = e.lnw_hi + log((double)b.hist / e.mean_hist);
      } else {
        v = b.lnw;
      }
      sum += v - e.c;
      n += 1;
    }
  }
  psum[(int64_t)c * ng + g] = sum;
  pcnt[(int64_t)c * ng + g] = n;
}

// The chunks' partial sums, combined in chunk order.
__global__ __launch_bounds__(256) void k_entropy_combine(int64_t ng, int chunks, const double *psum,
                                                         const int64_t *pcnt, double *S, int64_t *count) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ng) return;
  double sum = 0.0;
  int64_t n = 0;
  for (int c = 0; c < chunks; c++) {
    sum += psum[(int64_t)c * ng + g];
    n += pcnt[(int64_t)c * ng + g];
  }
  S[g] = n ? sum / (double)n : 0.0;
  count[g] = n;
}

// The set's common grid from the walker states, on the device: gmin = the
// smallest bins_min, lo / hi = the largest too_lo / smallest too_hi over the
// walkers still running with bins, then ng = the largest
// round((bins_min - gmin) / width) + len.  kEntropyRangeBlocks blocks reduce
// slices of the walkers into partial records, one block combines them
// (twice: ng needs gmin).  Minima and maxima do not depend on the order.
template <typename T, typename F>
__device__ __forceinline__ T block_reduce(T v, T *sh, F op) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = blockDim.x / 2; d > 0; d >>= 1) {
    if (t < d) sh[t] = op(sh[t], sh[t + d]);
    __syncthreads();
  }
  const T r = sh[0];
  __syncthreads();
  return r;
}
struct DMin {
  __device__ double operator()(double a, double b) const { return b < a ? b : a; }
};
struct DMax {
  __device__ double operator()(double a, double b) const { return b > a ? b : a; }
};
struct IMax {
  __device__ int64_t operator()(int64_t a, int64_t b) const { return b > a ? b : a; }
};

__global__ __launch_bounds__(256) void k_entropy_range_part(const sadmc_walker_state *ws, int W, EntropyRange *part) {
  __shared__ double sd[256];
  __shared__ int64_t si[256];
  double g = __builtin_inf(), lo = -__builtin_inf(), hi = __builtin_inf();
  int64_t any = 0;
  for (int w = blockIdx.x * 256 + threadIdx.x; w < W; w += gridDim.x * 256) {
    const sadmc_walker_state &x = ws[w];
    if (x.status || x.len <= 0) continue;
    g = x.bins_min < g ? x.bins_min : g;
    lo = x.too_lo > lo ? x.too_lo : lo;
    hi = x.too_hi < hi ? x.too_hi : hi;
    any = 1;
  }
  g = block_reduce(g, sd, DMin());
  lo = block_reduce(lo, sd, DMax());
  hi = block_reduce(hi, sd, DMin());
  any = block_reduce(any, si, IMax());
  if (threadIdx.x == 0) {
    EntropyRange r;
    r.gmin = g;
    r.lo = lo;
    r.hi = hi;
    r.ng = 0;
    r.any = (int32_t)any;
    r.pad = 0;
    part[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void k_entropy_range_ng(const sadmc_walker_state *ws, int W, double width,
                                                          const EntropyRange *out, EntropyRange *part) {
  __shared__ int64_t si[256];
  const double gmin = out->gmin;
  int64_t ng = 0;
  for (int w = blockIdx.x * 256 + threadIdx.x; w < W; w += gridDim.x * 256) {
    const sadmc_walker_state &x = ws[w];
    if (x.status || x.len <= 0) continue;
    const int64_t e = (int64_t)llround((x.bins_min - gmin) / width) + x.len;
    ng = e > ng ? e : ng;
  }
  ng = block_reduce(ng, si, IMax());
  if (threadIdx.x == 0) part[blockIdx.x].ng = ng;
}

// One block of kEntropyRangeBlocks lanes combines the partial records; with
// `ng_only` only their ng.
__global__ __launch_bounds__(kEntropyRangeBlocks) void k_entropy_range_join(const EntropyRange *part,
                                                                            EntropyRange *out, int ng_only) {
  __shared__ double sd[kEntropyRangeBlocks];
  __shared__ int64_t si[kEntropyRangeBlocks];
  const EntropyRange p = part[threadIdx.x];
  if (ng_only) {
    const int64_t ng = block_reduce(p.ng, si, IMax());
    if (threadIdx.x == 0) out->ng = ng;
    return;
  }
  const double g = block_reduce(p.gmin, sd, DMin());
  const double lo = block_reduce(p.lo, sd, DMax());
  const double hi = block_reduce(p.hi, sd, DMin());
  const int64_t any = block_reduce((int64_t)p.any, si, IMax());
  if (threadIdx.x == 0) {
    EntropyRange r;
    r.gmin = g;
    r.lo = lo;
    r.hi = hi;
    r.ng = 0;
    r.any = (int32_t)any;
    r.pad = 0;
    *out = r;
  }
}

hipError_t launch_entropy_range(const sadmc_walker_state *ws, int W, double width, EntropyRange *out,
                                EntropyRange *part, hipStream_t s) {
  const dim3 nb(kEntropyRangeBlocks), nt(256);
  hipLaunchKernelGGL(k_entropy_range_part, nb, nt, 0, s, ws, W, part);
  hipLaunchKernelGGL(k_entropy_range_join, dim3(1), nb, 0, s, (const EntropyRange *)part, out, 0);
  hipLaunchKernelGGL(k_entropy_range_ng, nb, nt, 0, s, ws, W, width, (const EntropyRange *)out, part);
  hipLaunchKernelGGL(k_entropy_range_join, dim3(1), nb, 0, s, (const EntropyRange *)part, out, 1);
  return hipGetLastError();
}

int entropy_chunks(int W, int64_t ng) {
  const int64_t blocks = (ng + 255) / 256;
  int64_t c = (2048 + blocks - 1) / blocks;  // about 8 blocks per CU
  if (c > 64) c = 64;
  if (c > W) c = W;
  return c < 1 ? 1 : (int)c;
}

hipError_t launch_entropy(const sadmc_walker_state *ws, const HotBin *hot, int W, int64_t cap, double width,
                          double gmin, int64_t ga, int64_t gb, int64_t ng, void *ew, int chunks, double *psum,
                          int64_t *pcnt, double *S, int64_t *count, hipStream_t s) {
  hipLaunchKernelGGL(k_entropy_walker, dim3(W), dim3(64), 0, s, ws, hot, cap, width, gmin, ga, gb,
                     (EntropyWalker *)ew);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned bx = (unsigned)((ng + 255) / 256);
  hipLaunchKernelGGL(k_entropy_bins, dim3(bx, (unsigned)chunks), dim3(256), 0, s, (const EntropyWalker *)ew, W, hot,
                     cap, width, ng, chunks, psum, pcnt);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_entropy_combine, dim3(bx), dim3(256), 0, s, ng, chunks, (const double *)psum,
                     (const int64_t *)pcnt, S, count);
  return hipGetLastError();
}

size_t entropy_walker_bytes() { return sizeof(EntropyWalker); }

}  // namespace sadmc
