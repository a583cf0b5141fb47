// Shared device helpers for the FiRE kernels (gfx950).
//
// Tangent-row convention ("FL5"): a quantity that depends on one 3-vector (an electron
// position, or a pair difference r_ctr - r_nb) is carried as 5 rows:
//   row 0 = value, row 1 = Laplacian, rows 2..4 = d/dx, d/dy, d/dz.
// An embedding group (one centre electron i) extends this to 2 + 3 (1 + nb_i) rows:
//   row 0 = value, row 1 = Laplacian over all 3N coordinates, rows 2..4 = d/dr_i,
//   rows 5+3k..7+3k = d/dr_j for the k-th neighbour j of i (CSR order).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FIRE_W0MAX 32
#define FIRE_W1MAX 16
#define FIRE_EEMAX 64
#define FIRE_SMAX 64  // max size of {n} U N(n) handled by the determinant-Laplacian kernel

namespace fire {

struct FilterParams {  // PairwiseFilter (model/utils.py:393-433) with precomputed softplus
  const float* K;   // [4, w0]
  const float* b;   // [w0]
  const float* sp;  // [E] softplus(scale)
  const float* Wd;  // [w0, w1]
  const float* bd;  // [w1]
  const float* We;  // [E, w1]
  int w0, w1, E;
};

// Store of a tangent-row output that is far larger than the caches and read once by a later
// kernel: streaming (nontemporal) when NT, plain otherwise (value rows, sampler lists).
template <bool NT>
__device__ __forceinline__ void store_rows_out(float* p, float v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw, SC'11): 10 rounds
// on a 4 x 32-bit counter under a 2 x 32-bit key.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = M0 * c.x, hi0 = __umulhi(M0, c.x);
    const uint32_t lo1 = M1 * c.z, hi1 = __umulhi(M1, c.z);
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += W0;
    k.y += W1;
  }
  return c;
}

// (x >> 8 + 1/2) / 2^24: a uniform variate strictly inside (0, 1)
__device__ __forceinline__ double u01(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + __expf(-x)); }

// silu and its first two derivatives
__device__ __forceinline__ void silu3(float x, float& f, float& f1, float& f2) {
  float s = 1.0f / (1.0f + expf(-x));
  f = x * s;
  f1 = s * (1.0f + x * (1.0f - s));
  f2 = s * (1.0f - s) * (2.0f + x * (1.0f - 2.0f * s));
}

// Forward-Laplacian of beta (PairwiseFilter output, w1 values) and of the log-scaled
// features g = [log1p d, diff * log1p(d)/d] with respect to the 3-vector `diff`.
// Output rows follow the FL5 convention.  cut_c: cutoff of the polynomial cutoff_function
// (model/utils.py:103-112, p = 4).
// W1 is the (zero-padded) filter width; every loop over it is unrolled so the per-thread
// arrays stay in registers.  NRF = 1 computes the value row only (value-only evaluation).
template <int W1, int NRF = 5>
__device__ inline void pair_filter_fl(const float diff[3], const FilterParams& P, float cut_c, float beta[NRF][W1],
                                      float g[NRF][4]) {
  static_assert(NRF == 1 || NRF == 5, "value row only, or all forward-Laplacian rows");
  const float dx = diff[0], dy = diff[1], dz = diff[2];
  const float d = sqrtf(dx * dx + dy * dy + dz * dz);
  const float inv_d = 1.0f / d;
  const float n[3] = {dx * inv_d, dy * inv_d, dz * inv_d};
  const float lap_d = 2.0f * inv_d;

  float dir[NRF][W1];
  float rad[NRF][W1];
#pragma unroll
  for (int j = 0; j < W1; ++j) {
#pragma unroll
    for (int r = 0; r < NRF; ++r) { dir[r][j] = 0.f; rad[r][j] = 0.f; }
    dir[0][j] = P.bd[j];
  }
  // directional MLP: silu([d, diff] @ K + b) @ Wd + bd
  for (int k = 0; k < P.w0; ++k) {
    const float k0 = P.K[k], k1 = P.K[P.w0 + k], k2 = P.K[2 * P.w0 + k], k3 = P.K[3 * P.w0 + k];
    const float pre = P.b[k] + d * k0 + dx * k1 + dy * k2 + dz * k3;
    const float pt0 = n[0] * k0 + k1, pt1 = n[1] * k0 + k2, pt2 = n[2] * k0 + k3;
    const float pl = lap_d * k0;
    float f, f1, f2;
    silu3(pre, f, f1, f2);
    const float a[5] = {f, f1 * pl + f2 * (pt0 * pt0 + pt1 * pt1 + pt2 * pt2), f1 * pt0, f1 * pt1, f1 * pt2};
    const float* wrow = P.Wd + k * W1;
#pragma unroll
    for (int j = 0; j < W1; ++j) {
      const float w = wrow[j];
#pragma unroll
      for (int r = 0; r < NRF; ++r) dir[r][j] += a[r] * w;
    }
  }
  // radial envelopes exp(-(d/sp)^2) * cutoff(d / cut_c), then Dense (no bias)
  float cu = 0.f, cu1 = 0.f, cu2 = 0.f;
  {
    const float u = d / cut_c;
    if (u < 1.0f) {
      const float u2 = u * u, u3 = u2 * u, u4 = u2 * u2;
      cu = 1.0f + u4 * (-15.0f + u * (24.0f - 10.0f * u));
      cu1 = u3 * (-60.0f + u * (120.0f - 60.0f * u)) / cut_c;
      cu2 = u2 * (-180.0f + u * (480.0f - 300.0f * u)) / (cut_c * cut_c);
    }
  }
  for (int e = 0; e < P.E; ++e) {
    const float sp = P.sp[e];
    const float x = d / sp;
    const float q = expf(-x * x);
    const float q1 = -2.0f * x / sp * q;
    const float q2 = (4.0f * x * x - 2.0f) / (sp * sp) * q;
    const float ev = q * cu, e1 = q1 * cu + q * cu1, e2 = q2 * cu + 2.0f * q1 * cu1 + q * cu2;
    const float a[5] = {ev, e2 + e1 * lap_d, e1 * n[0], e1 * n[1], e1 * n[2]};
    const float* wrow = P.We + e * W1;
#pragma unroll
    for (int j = 0; j < W1; ++j) {
      const float w = wrow[j];
#pragma unroll
      for (int r = 0; r < NRF; ++r) rad[r][j] += a[r] * w;
    }
  }
#pragma unroll
  for (int j = 0; j < W1; ++j) {
    const float dxv = dir[0][j], rx = rad[0][j];
    beta[0][j] = dxv * rx;
    if constexpr (NRF == 5) {
      beta[1][j] = dir[1][j] * rx + dxv * rad[1][j] +
                   2.0f * (dir[2][j] * rad[2][j] + dir[3][j] * rad[3][j] + dir[4][j] * rad[4][j]);
      beta[2][j] = dir[2][j] * rx + dxv * rad[2][j];
      beta[3][j] = dir[3][j] * rx + dxv * rad[3][j];
      beta[4][j] = dir[4][j] * rx + dxv * rad[4][j];
    }
  }
  // log-scaled features
  const float L = log1pf(d), L1 = 1.0f / (1.0f + d), L2 = -L1 * L1;
  const float w = L * inv_d;
  g[0][0] = L;
  for (int c = 0; c < 3; ++c) g[0][1 + c] = diff[c] * w;
  if constexpr (NRF == 5) {
    g[1][0] = L2 + L1 * lap_d;
    g[2][0] = L1 * n[0];
    g[3][0] = L1 * n[1];
    g[4][0] = L1 * n[2];
    const float w1 = L1 * inv_d - L * inv_d * inv_d;
    const float w2 = L2 * inv_d - 2.0f * L1 * inv_d * inv_d + 2.0f * L * inv_d * inv_d * inv_d;
    for (int c = 0; c < 3; ++c) {
      g[1][1 + c] = 2.0f * w1 * n[c] + diff[c] * (w2 + w1 * lap_d);
      for (int t = 0; t < 3; ++t) g[2 + t][1 + c] = (t == c ? w : 0.0f) + diff[c] * w1 * n[t];
    }
  }
}

__device__ __forceinline__ float warp_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of NV values per thread (blockDim.x <= 1024, multiple of 64).
template <int NV, typename T>
__device__ inline void block_sum(T (&v)[NV], T* smem /* >= 16*NV */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = warp_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[wid * NV + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T s = 0;
    for (int w = 0; w < nw; ++w) s += smem[w * NV + k];
    v[k] = s;
  }
  __syncthreads();
}

// group index of electron i of walker w: up-spin groups first (walker-major), then down
__device__ __forceinline__ int group_of(int w, int i, int B, int n_up, int n_dn) {
  return i < n_up ? w * n_up + i : B * n_up + w * n_dn + (i - n_up);
}
// Workgroups are dealt round-robin over the 8 XCDs (block b runs on XCD b % 8).  This maps
// block b of n to work item x q + min(x, r) + b / 8 (x = b % 8, q = n / 8, r = n % 8), so each
// XCD takes one contiguous range of work items and data they share (the rows of one walker's
// electrons) stays in that XCD's L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ void group_inv(int g, int B, int n_up, int n_dn, int& w, int& i) {
  if (g < B * n_up) { w = g / n_up; i = g - w * n_up; }
  else { int h = g - B * n_up; w = h / n_dn; i = n_up + (h - w * n_dn); }
}

}  // namespace fire
