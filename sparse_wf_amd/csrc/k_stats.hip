// Local-energy statistics and clipping of one walker batch (the consumer of the local energy
// in a VMC step): mean and standard deviation, the clip centre (median or mean), the mean
// absolute deviation from it, the clipped energies and their deviation from their own mean.
// Reference: clip_local_energies / local_energy_diff (src/sparse_wf/update.py:45-69) and the
// E_mean / E_std of Trainer.step (update.py:166-168).
//
// One workgroup of 1024 threads reads the batch a few times (B is a walker batch, thousands of
// doubles).  Sums are deterministic: each thread adds its strided elements in order, then a
// fixed-order tree in LDS.  The median is exact: an MSB-first radix select (8 rounds of 8-bit
// digits) over an order-preserving 64-bit key of the doubles finds the elements of rank
// (B-1)/2 and B/2, whose midpoint is the median (jnp.median's "midpoint" rule).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fire_kernels.hpp"

namespace fire {

namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ uint64_t order_key(double x) {
  uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double from_key(uint64_t k) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

// fixed-order sum over the workgroup; every thread gets the total
__device__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = red[t] + red[t + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kThreads) void k_energy_stats(const double* __restrict__ e, int B, double clip, int stat,
                                                           double* __restrict__ clipped, double* __restrict__ diff,
                                                           double* __restrict__ stats) {
  __shared__ double red[kThreads];
  __shared__ unsigned hist[kWaves][2][256];
  __shared__ unsigned long long prefix[2];
  __shared__ unsigned rank[2];
  __shared__ int any_nan;
  const int t = threadIdx.x;
  const int wv = t >> 6;

  if (t == 0) any_nan = 0;
  __syncthreads();
  double s = 0.0;
  int nan_here = 0;
  for (int i = t; i < B; i += kThreads) {
    const double x = e[i];
    s += x;
    nan_here |= (x != x);
  }
  if (nan_here) atomicOr(&any_nan, 1);
  const double mean = block_sum(s, red) / B;
  s = 0.0;
  for (int i = t; i < B; i += kThreads) {
    const double d = e[i] - mean;
    s += d * d;
  }
  const double var = block_sum(s, red) / B;
  const double nan = __longlong_as_double(0x7ff8000000000000ll);

  double lo = 0.0, hi = 0.0, center = nan, mad = nan;
  const bool do_clip = clip > 0.0;
  if (do_clip) {
    if (stat == 0) {
      // radix select of the ranks (B-1)/2 and B/2
      if (t < 2) {
        prefix[t] = 0ull;
        rank[t] = t == 0 ? (unsigned)((B - 1) / 2) : (unsigned)(B / 2);
      }
      unsigned long long mask = 0ull;
      for (int shift = 56; shift >= 0; shift -= 8) {
        for (int j = t; j < kWaves * 2 * 256; j += kThreads) (&hist[0][0][0])[j] = 0u;
        __syncthreads();
        const unsigned long long p0 = prefix[0], p1 = prefix[1];
        for (int i = t; i < B; i += kThreads) {
          const unsigned long long u = order_key(e[i]);
          const unsigned dg = (unsigned)(u >> shift) & 255u;
          if ((u & mask) == p0) atomicAdd(&hist[wv][0][dg], 1u);
          if ((u & mask) == p1) atomicAdd(&hist[wv][1][dg], 1u);
        }
        __syncthreads();
        // per-digit totals over the waves (512 entries, one thread each)
        if (t < 512) {
          unsigned c = 0;
          for (int w = 0; w < kWaves; ++w) c += (&hist[w][0][0])[t];
          (&hist[0][0][0])[t] = c;
        }
        __syncthreads();
        if (t < 2) {
          unsigned c = 0;
          int b = 0;
          for (; b < 255; ++b) {
            const unsigned n = hist[0][t][b];
            if (c + n > rank[t]) break;
            c += n;
          }
          rank[t] -= c;
          prefix[t] |= (unsigned long long)b << shift;
        }
        mask |= 0xffull << shift;
        __syncthreads();
      }
      const double a = from_key(prefix[0]), b = from_key(prefix[1]);
      center = any_nan ? nan : (a + b) * 0.5;
    } else {
      center = mean;
    }
    s = 0.0;
    for (int i = t; i < B; i += kThreads) s += fabs(e[i] - center);
    mad = block_sum(s, red) / B;
    const double max_dev = clip * mad;
    lo = center - max_dev;
    hi = center + max_dev;
  }
  // clipped values (jnp.clip: a NaN stays NaN) and their mean
  s = 0.0;
  for (int i = t; i < B; i += kThreads) {
    double x = e[i];
    if (do_clip) {
      x = x < lo ? lo : x;
      x = x > hi ? hi : x;
    }
    s += x;
  }
  const double mean_c = block_sum(s, red) / B;
  for (int i = t; i < B; i += kThreads) {
    double x = e[i];
    if (do_clip) {
      x = x < lo ? lo : x;
      x = x > hi ? hi : x;
    }
    if (clipped) clipped[i] = x;
    if (diff) diff[i] = x - mean_c;
  }
  if (t == 0 && stats) {
    stats[0] = mean;
    stats[1] = sqrt(var);
    stats[2] = center;
    stats[3] = mad;
    stats[4] = mean_c;
  }
}

}  // namespace

void launch_energy_stats(hipStream_t s, const double* e, int B, double clip, int stat, double* clipped, double* diff,
                         double* stats) {
  hipLaunchKernelGGL(k_energy_stats, dim3(1), dim3(kThreads), 0, s, e, B, clip, stat, clipped, diff, stats);
}

}  // namespace fire
