// Dense contractions over tangent rows: C[M,N] = A[M,K] @ W[K,N] (+ bias on value rows).
//
// These are the Linear layers of the reference applied to NodeWithFwdLap (value, every
// jacobian row and the Laplacian row go through the same matrix, the bias only on the
// value row: sparse_fwd_lap.py:160-181).  Value rows are marked by a per-row flag or by
// row % period == 0.
//
// MFMA kernels: 256 threads = 4 wavefronts, 128x128 output tile, each wavefront a 64x64
// quarter made of 2x2 32x32 MFMA tiles; K staged through LDS 32 at a time with the next
// chunk prefetched into registers while the current one is multiplied.  Tiles are mapped
// so that the N-tiles of one M-tile run back to back on the same XCD (its L2 then serves the
// repeated reads of the A rows).
//   FP32 path : v_mfma_f32_32x32x2f32, A and W in FP32.
//   BF16 path : v_mfma_f32_32x32x16_bf16, FP32 accumulation.  NSPLIT = 3 carries each
//               operand as hi + lo BF16 pair (x = hi + lo to ~2^-17 relative) and
//               accumulates hi*hi + hi*lo + lo*hi ("split BF16 x3"); NSPLIT = 6 carries
//               hi + mid + lo (x to ~2^-26) and accumulates the six products whose order is
//               at most 2^-16 relative (x6, FP32-level products); NSPLIT = 1 is plain BF16.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "fire_common.hpp"
#include "fire_kernels.hpp"

namespace fire {

typedef float float16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float float2v_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// Split-plane product terms (plane of A, plane of B), smallest first: for NP = 3 the six
// products of order >= 2^-18 relative, for NP = 2 the three of order >= 2^-9, for NP = 1 one.
__host__ __device__ constexpr int n_products(int np) { return np == 3 ? 6 : (np == 2 ? 3 : 1); }
__host__ __device__ constexpr int prod_a(int np, int t) {
  return np == 3 ? (t == 0 ? 1 : t == 1 ? 2 : t == 2 ? 0 : t == 3 ? 1 : 0) : (np == 2 ? (t == 0 ? 1 : 0) : 0);
}
__host__ __device__ constexpr int prod_b(int np, int t) {
  return np == 3 ? (t == 0 ? 1 : t == 1 ? 0 : t == 2 ? 2 : t == 3 ? 0 : t == 4 ? 1 : 0) : (np == 2 ? (t == 1 ? 1 : 0) : 0);
}

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------------------
// Two FP16 planes with power-of-two scales ("F16x3").  A row r of A is scaled by
// s_r = 2^(141 - e_r), e_r = biased exponent of max_k |A[r,k]| (so the row maximum lands in
// [2^14, 2^15)), a column n of W by its own s_n chosen the same way at upload; each scaled
// operand is x = hi + lo in FP16 (11 + 11 significand bits, lo exact in FP32 before its own
// rounding) and the contraction accumulates lo*hi + hi*lo + hi*hi in FP32 (v_mfma_f32_32x32x16_f16,
// every product exact).  The accumulators are unscaled by 1 / (s_r s_n) before the epilogue.
// Row maxima arrive as the float bits of max |A[r,:]| (non-negative floats order as integers),
// accumulated with atomicMax by the kernel that produced A.
// ---------------------------------------------------------------------------------------
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));

template <bool F16>
using OpT = std::conditional_t<F16, f16x8_t, bf16x8_t>;

__device__ __forceinline__ int f16_exponent(uint32_t mbits) {
  const int e = (int)((mbits >> 23) & 0xffu);
  return e < 15 ? 15 : (e > 254 ? 254 : e);
}
// s = 2^(141 - e) and 1 / s = 2^(e - 141) as float bits (biased exponents 268 - e and e - 14)
__device__ __forceinline__ float f16_row_scale(uint32_t mbits) {
  return __uint_as_float((uint32_t)(268 - f16_exponent(mbits)) << 23);
}
__device__ __forceinline__ float f16_row_unscale(uint32_t mbits) {
  return __uint_as_float((uint32_t)(f16_exponent(mbits) - 14) << 23);
}
__device__ __forceinline__ float row_scale_of(const uint32_t* amax, int64_t row, int64_t M) {
  return row < M ? f16_row_scale(amax[row]) : 1.f;
}

// The NP planes of four consecutive A values (two 32-bit words per plane): BF16 planes
// (x = hi + mid + lo, each remainder exact in FP32), or with F16 the FP16 planes of s * x.
template <int NP, bool F16>
__device__ __forceinline__ void split_planes(const float4 x, float s, uint2 (&out)[NP]) {
  float2v_t x01 = {x.x, x.y}, x23 = {x.z, x.w};
  if constexpr (F16) {
    x01 = x01 * s;
    x23 = x23 * s;
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    if constexpr (F16) {
      const f16x2_t h01 = __builtin_convertvector(x01, f16x2_t);
      const f16x2_t h23 = __builtin_convertvector(x23, f16x2_t);
      out[p] = make_uint2(__builtin_bit_cast(uint32_t, h01), __builtin_bit_cast(uint32_t, h23));
      if (p + 1 < NP) {
        x01 = x01 - __builtin_convertvector(h01, float2v_t);
        x23 = x23 - __builtin_convertvector(h23, float2v_t);
      }
    } else {
      const bf16x2_t h01 = __builtin_convertvector(x01, bf16x2_t);
      const bf16x2_t h23 = __builtin_convertvector(x23, bf16x2_t);
      out[p] = make_uint2(__builtin_bit_cast(uint32_t, h01), __builtin_bit_cast(uint32_t, h23));
      if (p + 1 < NP) {
        x01 = x01 - __builtin_convertvector(h01, float2v_t);
        x23 = x23 - __builtin_convertvector(h23, float2v_t);
      }
    }
  }
}

__device__ __forceinline__ float16_t mfma16(const bf16x8_t& a, const bf16x8_t& b, const float16_t& c, int, int, int) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float16_t mfma16(const f16x8_t& a, const f16x8_t& b, const float16_t& c, int, int, int) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// Row maxima of a wavefront's 64-row block for the next contraction: mx[a][v] is the lane's
// max |value| of row 32 a + 8 (v / 4) + 4 (lane / 32) + v % 4 over its columns; the 32 lanes of
// each half combine theirs and lane t of the half adds row t of the half with atomicMax.
__device__ __forceinline__ void rows_max_out(float (&mx)[2][16], int64_t rb, int lane, int64_t M,
                                             uint32_t* __restrict__ omax) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      float m = mx[a][v];
#pragma unroll
      for (int off = 1; off < 32; off <<= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
      mx[a][v] = m;
    }
  const int t = lane & 31;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int v = 0; v < 16; ++v)
      if (t == a * 16 + v) {
        const int64_t row = rb + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row < M && mx[a][v] > 0.f) atomicMax(omax + row, __float_as_uint(mx[a][v]));
      }
}

#define GT_M 128
#define GT_N 128
#define GT_K 32
#define LDA_F (GT_K + 4)  // fp32 LDS row stride (floats)
#define LDB_H (GT_K + 8)  // bf16 LDS row stride (halves): 80 bytes, keeps 16-byte alignment
#define N_XCD 8

template <int V>
using IC = std::integral_constant<int, V>;

// f(IC<I>()), f(IC<I + 1>()), ..., f(IC<N - 1>()): indices into register arrays stay compile-time
// constants from the front end on (a #pragma unroll loop is unrolled only after the arrays it
// indexes may already have been placed in scratch memory)
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(IC<I>());
    static_for<I + 1, N>(f);
  }
}

// Number of the wavefront's 32-column MFMA tiles (of nb) that hold columns below N: the last
// N tile of a contraction whose N is not a multiple of the tile width multiplies only those.
__device__ __forceinline__ int live_col_tiles(int c0, int N, int nb) {
  const int n = (N - c0 + 31) >> 5;
  return n < 0 ? 0 : (n > nb ? nb : n);
}

// Offset of weight element (k, n) in a K-blocked plane of N columns (the 16-wide K chunk of a
// column is contiguous, consecutive columns follow each other): ((k / 16) N + n) 16 + k % 16.
__device__ __forceinline__ int64_t wpl_off(int64_t n, int k, int N) {
  return ((int64_t)(k >> 4) * N + n) * 16 + (k & 15);
}

__device__ __forceinline__ bool is_bias_row(int64_t row, const uint8_t* rowflag, int period) {
  if (rowflag) return rowflag[row] != 0;
  if (period > 0) return (row % period) == 0;
  return false;
}

// Linear block id -> (M tile, N tile).  Workgroups are dispatched round-robin over the
// XCDs; tile t of the remapped order lives on XCD t / (T/8), and consecutive t share the M
// tile, so the N tiles of an M tile are served by one L2.
__device__ __forceinline__ void tile_of(int64_t& m_tile, int& n_tile, int nT) {
  const int64_t bid = blockIdx.x;
  const int64_t T = (int64_t)gridDim.x;
  int64_t t = bid;
  if (T % N_XCD == 0) t = (bid % N_XCD) * (T / N_XCD) + bid / N_XCD;
  m_tile = t / nT;
  n_tile = (int)(t - m_tile * nT);
}

// Writes a wavefront's 64 x (32 NB) block of 32x32 MFMA accumulators (lane l, register v of
// tile (a, b): row 32 a + 8 (v / 4) + 4 (l / 32) + v % 4, column 32 b + l % 32), either plain
// (+ bias on value rows) or through the fused forward-Laplacian silu described at ActEpi.
// With F16 every accumulator is first unscaled by 1 / (s_r s_n): lane t of the 64-row block
// loads the row maximum of row t once and the rows' factors travel by shuffle, so no row
// address is held across the block.  C goes out with streaming (nontemporal) stores: its rows
// are far larger than the caches and a later kernel reads them once.
template <int NB, bool FUSE, bool F16 = false, typename AccT>
__device__ __forceinline__ void store_rows(const AccT (&acc)[2][NB], int64_t m0, int n0, int wm, int wn, int lane,
                                           float* __restrict__ C, int64_t ldc, int64_t M, int N,
                                           const float* __restrict__ bias, const uint8_t* __restrict__ rowflag,
                                           int period, const ActEpi& ep) {
  float myrs = 1.f;
  if constexpr (F16) {
    const int64_t r = m0 + wm + lane;
    myrs = r < M ? f16_row_unscale(ep.amax[r]) : 0.f;
  }
  if constexpr (!FUSE && F16) {
    float cs[NB], bcol[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int col = n0 + wn + b * 32 + (lane & 31);
      cs[b] = col < N ? ep.winv[col] : 0.f;
      bcol[b] = (bias && col < N) ? bias[col] : 0.f;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rl = a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const float rs = __shfl(myrs, rl);
        const int64_t row = m0 + wm + rl;
        if (row >= M) continue;
        const bool br = bias && is_bias_row(row, rowflag, period);
        float* crow = C + row * ldc;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int col = n0 + wn + b * 32 + (lane & 31);
          if (col >= N) continue;
          float val = acc[a][b][v] * (rs * cs[b]);
          if (br) val += bcol[b];
          __builtin_nontemporal_store(val, crow + col);
        }
      }
  } else if constexpr (!FUSE) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int col = n0 + wn + b * 32 + (lane & 31);
      if (col >= N) continue;
      const float bcol = bias ? bias[col] : 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int64_t row = m0 + wm + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          if (row >= M) continue;
          float val = acc[a][b][v];
          if (bias && is_bias_row(row, rowflag, period)) val += bcol;
          __builtin_nontemporal_store(val, C + row * ldc + col);
        }
    }
  } else {
    // fused silu: the lane's 32 rows ascend with (a, v), so the jacobian rows of one group form
    // one run per lane (the same for each of its NB columns); each run's sum of squares per
    // column goes to its own slot.  Lane l holds the code of row rb + l; every lane takes part
    // in the shuffles.
    const int h = lane >> 5;
    const int64_t rb = m0 + wm;  // first row of this 64-row block (blocks are 64-aligned)
    const int mycode = rb + lane < M ? ep.rowgrp[rb + lane] : -3;
    int col[NB];
    bool colok[NB], fz[NB];
    float bcol[NB], ss[NB], f1[NB], cs[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      col[b] = n0 + wn + b * 32 + (lane & 31);
      colok[b] = col[b] < N;
      fz[b] = col[b] < ep.nfuse;
      bcol[b] = (bias && colok[b]) ? bias[col[b]] : 0.f;
      ss[b] = 0.f;
      f1[b] = 0.f;
      cs[b] = 1.f;
      if constexpr (F16) cs[b] = colok[b] ? ep.winv[col[b]] : 0.f;
    }
    int cur = -1;
    float mx[2][16];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int v = 0; v < 16; ++v) mx[a][v] = 0.f;
    auto flush = [&]() {
      const int64_t first = (ep.gd.row(cur) + 2) >> 6;  // block of the group's first jacobian row
      const int slot = 2 * (int)((rb >> 6) - first) + h;
      if (slot >= ACT_EPI_SLOTS) return;  // not within the neighbour limit (at most 4 blocks)
      float* dst = ep.part + ((int64_t)cur * ACT_EPI_SLOTS + slot) * ep.ldf;
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (colok[b] && fz[b]) dst[col[b]] = ss[b];
    };
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rl = a * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
        const int cd = __shfl(mycode, rl);
        float rs = 1.f;
        if constexpr (F16) rs = __shfl(myrs, rl);
        if (cd == -3) continue;
        if (cd >= 0 && cd != cur) {
          if (cur >= 0) flush();
          cur = cd;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            ss[b] = 0.f;
            f1[b] = colok[b] && fz[b] ? ep.f1[(int64_t)cd * ep.ldf + col[b]] : 0.f;
          }
        }
        float* crow = C + (rb + rl) * ldc;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (!colok[b]) continue;
          float z = acc[a][b][v];
          if constexpr (F16) z *= rs * cs[b];
          if (!fz[b]) {
            if (cd == -1) z += bcol[b];
          } else if (cd >= 0) {
            ss[b] += z * z;
            z *= f1[b];
          } else if (cd == -1) {
            z += bcol[b];
            z *= 1.0f / (1.0f + expf(-z));
          }
          __builtin_nontemporal_store(z, crow + col[b]);
          // row maxima for the next contraction (Laplacian rows are completed, and their
          // maxima taken, by k_act_lapfix)
          if (ep.omax && cd != -2 && col[b] < ep.omax_cols) mx[a][v] = fmaxf(mx[a][v], fabsf(z));
        }
      }
    if (cur >= 0) flush();
    if (ep.omax) rows_max_out(mx, rb, lane, M, ep.omax);
  }
}

template <bool FUSE, bool F16 = false, typename AccT>
__device__ __forceinline__ void store_tile(const AccT (&acc)[2][2], int64_t m0, int n0, int wm, int wn, int lane,
                                           float* __restrict__ C, int64_t ldc, int64_t M, int N,
                                           const float* __restrict__ bias, const uint8_t* __restrict__ rowflag,
                                           int period, const ActEpi& ep) {
  store_rows<2, FUSE, F16>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
}

// ---------------------------------------------------------------------------------------
// FP32 MFMA GEMM.  Wt is [N,K] FP32 (W transposed).  Vector path: K % 4 == 0, lda % 4 == 0.
// ---------------------------------------------------------------------------------------
template <bool VEC, bool FUSE = false>
__global__ void __launch_bounds__(256) k_gemm_f32(const float* __restrict__ A, int64_t lda,
                                                  const float* __restrict__ Wt, const float* __restrict__ bias,
                                                  float* __restrict__ C, int64_t ldc, int64_t M, int N, int K,
                                                  const uint8_t* __restrict__ rowflag, int period, int nT,
                                                  DevRows dr, int64_t sA, int64_t sW, int64_t sC,
                                                  ActEpi ep) {
  __shared__ __attribute__((aligned(16))) float As[GT_M][LDA_F];
  __shared__ __attribute__((aligned(16))) float Bs[GT_N][LDA_F];  // Bs[n][k]
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // batch entry blockIdx.z: operands offset by the batch strides
  A += blockIdx.z * sA;
  Wt += blockIdx.z * sW;
  C += blockIdx.z * sC;
  int64_t mt;
  int nt;
  tile_of(mt, nt, nT);
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  if (mt * GT_M >= M) return;
  const int64_t m0 = mt * GT_M;
  const int n0 = nt * GT_N;
  const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
  float16_t acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  float4 ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + q * 256;
      const int rr = e >> 3, kq = (e & 7) * 4;
      const int64_t gr = m0 + rr;
      const int gn = n0 + rr;
      if (VEC) {  // K % 4 == 0: a float4 lies wholly inside or wholly past K
        const bool kin = k0 + kq < K;
        ra[q] = gr < M && kin ? *reinterpret_cast<const float4*>(A + gr * lda + k0 + kq) : make_float4(0.f, 0.f, 0.f, 0.f);
        rb[q] = gn < N && kin ? *reinterpret_cast<const float4*>(Wt + (int64_t)gn * K + k0 + kq)
                              : make_float4(0.f, 0.f, 0.f, 0.f);
      } else {
        float va[4], vb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gk = k0 + kq + j;
          va[j] = (gr < M && gk < K) ? A[gr * lda + gk] : 0.f;
          vb[j] = (gn < N && gk < K) ? Wt[(int64_t)gn * K + gk] : 0.f;
        }
        ra[q] = make_float4(va[0], va[1], va[2], va[3]);
        rb[q] = make_float4(vb[0], vb[1], vb[2], vb[3]);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + q * 256;
      const int rr = e >> 3, kq = (e & 7) * 4;
      *reinterpret_cast<float4*>(&As[rr][kq]) = ra[q];
      *reinterpret_cast<float4*>(&Bs[rr][kq]) = rb[q];
    }
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += GT_K) {
    store();
    __syncthreads();
    if (k0 + GT_K < K) load(k0 + GT_K);
#pragma unroll
    for (int kk = 0; kk < GT_K; kk += 2) {
      const int ka = kk + (lane >> 5);
      float av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[wm + a * 32 + (lane & 31)][ka];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = Bs[wn + b * 32 + (lane & 31)][ka];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  store_tile<FUSE>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
}

// ---------------------------------------------------------------------------------------
// BF16 / split-BF16 MFMA GEMM, FP32 accumulation.  W planes are [N,K] BF16 (W transposed):
// NP = 1 plain BF16; NP = 2 hi + lo with the products hi*hi + hi*lo + lo*hi; NP = 3
// hi + mid + lo with the six products of order >= 2^-18 relative (smallest first).  A is
// split the same way while it is staged into LDS.  Output tile BM x BN, one 64 x 64
// quarter (2 x 2 MFMA 32x32x16 tiles) per wavefront; K staged 32 at a time with the next
// chunk prefetched into registers.  Requires K % 32 == 0 and lda % 4 == 0.
// ---------------------------------------------------------------------------------------
template <int NP, int BM, int BN, bool FUSE = false, bool F16 = false>
__global__ void __launch_bounds__(BM * BN / 64) k_gemm_bf16(const float* __restrict__ A, int64_t lda,
                                                           const uint16_t* __restrict__ Wp0,
                                                           const uint16_t* __restrict__ Wp1,
                                                           const uint16_t* __restrict__ Wp2,
                                                           const float* __restrict__ bias, float* __restrict__ C,
                                                           int64_t ldc, int64_t M, int N, int K,
                                                           const uint8_t* __restrict__ rowflag, int period, int nT,
                                                           DevRows dr, ActEpi ep) {
  constexpr int NWN = BN / 64;
  constexpr int NT = (BM / 64) * NWN * 64;  // threads
  constexpr int A_PER = BM * 8 / NT;        // float4 of the A chunk per thread
  constexpr int B_PER = BN * 4 / NT;        // uint4 of each W plane chunk per thread
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small for its thread count");
  extern __shared__ __attribute__((aligned(16))) unsigned char gemm_lds[];
  uint16_t(*As)[BM][LDB_H] = reinterpret_cast<uint16_t(*)[BM][LDB_H]>(gemm_lds);
  uint16_t(*Bs)[BN][LDB_H] = reinterpret_cast<uint16_t(*)[BN][LDB_H]>(gemm_lds + (size_t)NP * BM * LDB_H * 2);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t mt;
  int nt;
  tile_of(mt, nt, nT);
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
    if (ep.amax) ep.amax += o;
    if (ep.omax) ep.omax += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  if (mt * BM >= M) return;
  const int64_t m0 = mt * BM;
  const int n0 = nt * BN;
  const int wm = (wid / NWN) * 64, wn = (wid % NWN) * 64;
  float16_t acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  const uint16_t* Wpl[3] = {Wp0, Wp1, Wp2};
  // FP16 planes: the scale of every A row this thread splits (the same rows in every chunk)
  float rsc[A_PER];
#pragma unroll
  for (int q = 0; q < A_PER; ++q) rsc[q] = F16 ? row_scale_of(ep.amax, m0 + ((tid + q * NT) >> 3), M) : 1.f;

  // A (HBM) is prefetched two chunks ahead through two register sets, the weight planes
  // (L2-resident) one chunk ahead: about twice the A bytes in flight of a single buffer.
  float4 ra0[A_PER], ra1[A_PER];
  uint4 rb[NP][B_PER];
  auto load_a = [&](float4(&dst)[A_PER], int k0) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * NT;
      const int rr = e >> 3, kq = (e & 7) * 4;
      const int64_t gr = m0 + rr;
      dst[q] = gr < M ? *reinterpret_cast<const float4*>(A + gr * lda + k0 + kq) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      const int nn = e >> 2, kq = (e & 3) * 8;
      const int gn = n0 + nn;
#pragma unroll
      for (int p = 0; p < NP; ++p)
        rb[p][q] = gn < N ? *reinterpret_cast<const uint4*>(Wpl[p] + wpl_off(gn, k0 + kq, N)) : make_uint4(0, 0, 0, 0);
    }
  };
  auto store = [&](const float4(&src)[A_PER]) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * NT;
      const int rr = e >> 3, kq = (e & 7) * 4;
      // x = hi + mid + lo with hardware round-to-nearest conversions (v_cvt_pk_bf16_f32);
      // each remainder is exact in FP32
      uint2 pw[NP];
      split_planes<NP, F16>(src[q], rsc[q], pw);
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<uint2*>(&As[p][rr][kq]) = pw[p];
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      const int nn = e >> 2, kq = (e & 3) * 8;
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<uint4*>(&Bs[p][nn][kq]) = rb[p][q];
    }
  };
  const int nbl = live_col_tiles(n0 + wn, N, 2);
  auto compute_nb = [&](auto NBL) {
    constexpr int nbc = decltype(NBL)::value;
#pragma unroll
    for (int kk = 0; kk < GT_K; kk += 16) {
      const int ka = kk + 8 * (lane >> 5);
      OpT<F16> av[NP][2], bv[NP][2];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int a = 0; a < 2; ++a) av[p][a] = *reinterpret_cast<const OpT<F16>*>(&As[p][wm + a * 32 + (lane & 31)][ka]);
#pragma unroll
        for (int b = 0; b < nbc; ++b) bv[p][b] = *reinterpret_cast<const OpT<F16>*>(&Bs[p][wn + b * 32 + (lane & 31)][ka]);
      }
      // product-major order: four independent accumulators between dependent MFMAs
#pragma unroll
      for (int t = 0; t < n_products(NP); ++t)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < nbc; ++b)
            acc[a][b] = mfma16(av[prod_a(NP, t)][a], bv[prod_b(NP, t)][b],
                                                                acc[a][b], 0, 0, 0);
    }
  };
  auto compute = [&]() {
    if (nbl == 2)
      compute_nb(IC<2>());
    else if (nbl == 1)
      compute_nb(IC<1>());
  };

  const int nch = K / GT_K;
  load_a(ra0, 0);
  load_b(0);
  if (nch > 1) load_a(ra1, GT_K);
  for (int c = 0; c < nch; c += 2) {
    store(ra0);  // chunk c
    __syncthreads();
    if (c + 2 < nch) load_a(ra0, (c + 2) * GT_K);
    if (c + 1 < nch) load_b((c + 1) * GT_K);
    compute();
    __syncthreads();
    if (c + 1 >= nch) break;
    store(ra1);  // chunk c + 1
    __syncthreads();
    if (c + 3 < nch) load_a(ra1, (c + 3) * GT_K);
    if (c + 2 < nch) load_b((c + 2) * GT_K);
    compute();
    __syncthreads();
  }
  store_tile<FUSE, F16>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
}

// Small-N FP32 GEMM (N <= 16): one wavefront per row, lanes across K.  W is [K,N].
__global__ void k_gemm_smalln(const float* __restrict__ A, int64_t lda, const float* __restrict__ W,
                              const float* __restrict__ bias, float* __restrict__ C, int64_t ldc, int64_t M, int N,
                              int K, const uint8_t* __restrict__ rowflag, int period, DevRows dr) {
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= M) return;
  float acc[16];
  for (int n = 0; n < 16; ++n) acc[n] = 0.f;
  for (int k = lane; k < K; k += 64) {
    const float a = A[row * lda + k];
#pragma unroll
    for (int n = 0; n < 16; ++n)
      if (n < N) acc[n] += a * W[(int64_t)k * N + n];
  }
  const bool b = bias && is_bias_row(row, rowflag, period);
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    if (n >= N) break;
    float v = warp_sum(acc[n]);
    if (lane == 0) C[row * ldc + n] = v + (b ? bias[n] : 0.f);
  }
}

// Column tail C[:, n0 .. n0 + nr) (nr <= 16) of a product whose other columns go through
// 128-wide tiles, FP32 MFMA 16x16x4.  Workgroup = 4 wavefronts x 64 rows (four 16-row tiles);
// lane l holds row / column l % 16 and, in a chunk of 16 k starting at k0, the four
// consecutive k0 + 4 (l / 16) + t, and MFMA step t contracts exactly those k (the same
// permutation of k for both operands).  VEC: K % 16 == 0 and lda % 4 == 0.
template <bool VEC>
__global__ void __launch_bounds__(256) k_gemm_f32_tail(const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ Wt, const float* __restrict__ bias,
                                                       float* __restrict__ C, int64_t ldc, int64_t M, int n0, int nr,
                                                       int K, const uint8_t* __restrict__ rowflag, int period,
                                                       DevRows dr, int64_t sA, int64_t sW, int64_t sC) {
  typedef float f4_t __attribute__((ext_vector_type(4)));
  A += blockIdx.z * sA;
  Wt += blockIdx.z * sW;
  C += blockIdx.z * sC;
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, li = lane & 15, kg = lane >> 4;
  const int64_t rbase = (int64_t)blockIdx.x * 256 + wid * 64;
  if (rbase >= M) return;
  const float* ar[4];
  bool aok[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int64_t r = rbase + t * 16 + li;
    aok[t] = r < M;
    ar[t] = A + (aok[t] ? r : 0) * lda + 4 * kg;
  }
  const bool wok = li < nr;
  const float* wr = Wt + (int64_t)(n0 + (wok ? li : 0)) * K + 4 * kg;
  f4_t acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = (f4_t){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    f4_t a[4], w;
    if (VEC) {
#pragma unroll
      for (int t = 0; t < 4; ++t) a[t] = aok[t] ? *reinterpret_cast<const f4_t*>(ar[t] + k0) : (f4_t)(0.f);
      w = wok ? *reinterpret_cast<const f4_t*>(wr + k0) : (f4_t)(0.f);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool kok = k0 + 4 * kg + j < K;
#pragma unroll
        for (int t = 0; t < 4; ++t) a[t][j] = aok[t] && kok ? ar[t][k0 + j] : 0.f;
        w[j] = wok && kok ? wr[k0 + j] : 0.f;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][j], w[j], acc[t], 0, 0, 0);
  }
  if (!wok) return;
  const int col = n0 + li;
  const float bcol = bias ? bias[col] : 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = rbase + t * 16 + 4 * kg + v;
      if (row >= M) continue;
      float val = acc[t][v];
      if (bias && is_bias_row(row, rowflag, period)) val += bcol;
      C[row * ldc + col] = val;
    }
}

void launch_gemm_f32_tail(hipStream_t s, const float* A, int64_t lda, const float* Wt, const float* bias, float* C,
                          int64_t ldc, int64_t M, int n0, int nr, int K, const uint8_t* rowflag, int period,
                          DevRows dr, int64_t sA, int64_t sW, int64_t sC, int nb) {
  if (M == 0 || nr <= 0 || nb == 0) return;
  const dim3 grid((unsigned)((M + 255) / 256), 1, (unsigned)nb);
  if ((K % 16) == 0 && (lda % 4) == 0)
    k_gemm_f32_tail<true><<<grid, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, n0, nr, K, rowflag, period, dr, sA, sW, sC);
  else
    k_gemm_f32_tail<false><<<grid, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, n0, nr, K, rowflag, period, dr, sA, sW,
                                                sC);
}

void launch_gemm_f32(hipStream_t s, const float* A, int64_t lda, const float* W, const float* Wt, const float* bias,
                     float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period, DevRows dr,
                     ActEpi ep) {
  if (M == 0) return;
  if (ep.rowgrp && N <= 16) abort();  // the fused epilogue is in the MFMA kernel only
  if (N <= 16) {
    const int64_t threads = M * 64;
    k_gemm_smalln<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(A, lda, W, bias, C, ldc, M, N, K, rowflag, period,
                                                                      dr);
    return;
  }
  // a remainder of at most 16 columns past whole 128-wide tiles goes to the tail kernel (not
  // with the fused epilogue, whose columns all go through the MFMA kernel)
  const int rem = N % GT_N, Nm = rem > 0 && rem <= 16 && !ep.rowgrp ? N - rem : N;
  const int nT = (Nm + GT_N - 1) / GT_N;
  const int64_t blocks = ((M + GT_M - 1) / GT_M) * nT;
  if (ep.rowgrp) {
    if ((K % 4) == 0 && (lda % 4) == 0)
      k_gemm_f32<true, true><<<(unsigned)blocks, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, Nm, K, rowflag, period, nT,
                                                              dr, 0, 0, 0, ep);
    else
      k_gemm_f32<false, true><<<(unsigned)blocks, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, Nm, K, rowflag, period, nT,
                                                               dr, 0, 0, 0, ep);
    return;
  }
  if ((K % 4) == 0 && (lda % 4) == 0)
    k_gemm_f32<true><<<(unsigned)blocks, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, Nm, K, rowflag, period, nT, dr, 0,
                                                      0, 0, ep);
  else
    k_gemm_f32<false><<<(unsigned)blocks, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, Nm, K, rowflag, period, nT, dr, 0,
                                                       0, 0, ep);
  if (Nm < N) launch_gemm_f32_tail(s, A, lda, Wt, bias, C, ldc, M, Nm, N - Nm, K, rowflag, period, dr, 0, 0, 0, 1);
}

// nb independent products C_b = A_b W_b (b < nb; operands at strides sA, sW (of Wt), sC) in
// one launch; every product is computed exactly as by launch_gemm_f32 (N > 16).
void launch_gemm_f32_batched(hipStream_t s, const float* A, int64_t lda, int64_t sA, const float* Wt, int64_t sW,
                             float* C, int64_t ldc, int64_t sC, int64_t M, int N, int K, int nb) {
  if (M == 0 || nb == 0) return;
  const int rem = N % GT_N, Nm = rem > 0 && rem <= 16 && N > GT_N ? N - rem : N;
  const int nT = (Nm + GT_N - 1) / GT_N;
  const dim3 grid((unsigned)(((M + GT_M - 1) / GT_M) * nT), 1, (unsigned)nb);
  if ((K % 4) == 0 && (lda % 4) == 0 && (sA % 4) == 0 && (sW % 4) == 0)
    k_gemm_f32<true><<<grid, 256, 0, s>>>(A, lda, Wt, nullptr, C, ldc, M, Nm, K, nullptr, 0, nT, DevRows(), sA, sW, sC, ActEpi());
  else
    k_gemm_f32<false><<<grid, 256, 0, s>>>(A, lda, Wt, nullptr, C, ldc, M, Nm, K, nullptr, 0, nT, DevRows(), sA, sW,
                                           sC, ActEpi());
  if (Nm < N)
    launch_gemm_f32_tail(s, A, lda, Wt, nullptr, C, ldc, M, Nm, N - Nm, K, nullptr, 0, DevRows(), sA, sW, sC, nb);
}

// Double-buffered form: two LDS stages of KC (16 or 32) k each; chunk c is multiplied from
// stage c % 2 while chunk c + 1 (already in registers) is split into stage (c + 1) % 2 and
// chunk c + 2 is loaded: one barrier per chunk.  With KC = 16 a 128 x 128 block needs 74 KB
// (NP = 3), so two blocks share a CU.
template <int NP, int BM, int BN, int KC, int PD>
__global__ void __launch_bounds__(BM * BN / 64) k_gemm_bf16_db(const float* __restrict__ A, int64_t lda,
                                                              const uint16_t* __restrict__ Wp0,
                                                              const uint16_t* __restrict__ Wp1,
                                                              const uint16_t* __restrict__ Wp2,
                                                              const float* __restrict__ bias, float* __restrict__ C,
                                                              int64_t ldc, int64_t M, int N, int K,
                                                              const uint8_t* __restrict__ rowflag, int period, int nT,
                                                              DevRows dr, ActEpi ep) {
  constexpr int NWN = BN / 64;
  constexpr int NT = (BM / 64) * NWN * 64;
  constexpr int LDH = KC + 8;                // halves per LDS row (16-byte aligned, conflict-free reads)
  constexpr int A_UNITS = BM * (KC / 4), B_UNITS = BN * (KC / 8);  // float4 / uint4 per chunk
  constexpr int A_PER = (A_UNITS + NT - 1) / NT;                    // per thread (guarded)
  constexpr int B_PER = (B_UNITS + NT - 1) / NT;
  constexpr size_t STAGE = (size_t)NP * (BM + BN) * LDH;  // halves per stage
  extern __shared__ __attribute__((aligned(16))) unsigned char gemm_lds[];
  uint16_t* lds = reinterpret_cast<uint16_t*>(gemm_lds);
  auto As = [&](int st, int p, int r) { return lds + st * STAGE + ((size_t)p * BM + r) * LDH; };
  auto Bs = [&](int st, int p, int r) { return lds + st * STAGE + (size_t)NP * BM * LDH + ((size_t)p * BN + r) * LDH; };
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t mt;
  int nt;
  tile_of(mt, nt, nT);
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  if (mt * BM >= M) return;
  const int64_t m0 = mt * BM;
  const int n0 = nt * BN;
  const int wm = (wid / NWN) * 64, wn = (wid % NWN) * 64;
  float16_t acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  const uint16_t* Wpl[3] = {Wp0, Wp1, Wp2};
  // PD register buffers: chunk j >= 1 waits in buffer (j - 1) % PD until it is split into LDS
  float4 ra[PD][A_PER];
  uint4 rb[PD][NP][B_PER];
  auto load = [&](int k0, float4(&xa)[A_PER], uint4(&xb)[NP][B_PER]) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * NT;
      const int rr = e / (KC / 4), kq = (e % (KC / 4)) * 4;
      const int64_t gr = m0 + rr;
      xa[q] = (e < A_UNITS && gr < M) ? *reinterpret_cast<const float4*>(A + gr * lda + k0 + kq)
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      const int nn = e / (KC / 8), kq = (e % (KC / 8)) * 8;
      const int gn = n0 + nn;
#pragma unroll
      for (int p = 0; p < NP; ++p)
        xb[p][q] = (e < B_UNITS && gn < N) ? *reinterpret_cast<const uint4*>(Wpl[p] + wpl_off(gn, k0 + kq, N))
                                           : make_uint4(0, 0, 0, 0);
    }
  };
  auto store = [&](int st, const float4(&ra)[A_PER], const uint4(&rb)[NP][B_PER]) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * NT;
      if (A_UNITS % NT != 0 && e >= A_UNITS) continue;
      const int rr = e / (KC / 4), kq = (e % (KC / 4)) * 4;
      float2v_t x01 = {ra[q].x, ra[q].y}, x23 = {ra[q].z, ra[q].w};
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const bf16x2_t h01 = __builtin_convertvector(x01, bf16x2_t);
        const bf16x2_t h23 = __builtin_convertvector(x23, bf16x2_t);
        *reinterpret_cast<uint2*>(As(st, p, rr) + kq) =
            make_uint2(__builtin_bit_cast(uint32_t, h01), __builtin_bit_cast(uint32_t, h23));
        if (p + 1 < NP) {
          x01 = x01 - __builtin_convertvector(h01, float2v_t);
          x23 = x23 - __builtin_convertvector(h23, float2v_t);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      if (B_UNITS % NT != 0 && e >= B_UNITS) continue;
      const int nn = e / (KC / 8), kq = (e % (KC / 8)) * 8;
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<uint4*>(Bs(st, p, nn) + kq) = rb[p][q];
    }
  };
  auto compute = [&](int st) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 16) {
      const int ka = kk + 8 * (lane >> 5);
      bf16x8_t av[NP][2], bv[NP][2];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int a = 0; a < 2; ++a) av[p][a] = *reinterpret_cast<const bf16x8_t*>(As(st, p, wm + a * 32 + (lane & 31)) + ka);
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[p][b] = *reinterpret_cast<const bf16x8_t*>(Bs(st, p, wn + b * 32 + (lane & 31)) + ka);
      }
#pragma unroll
      for (int t = 0; t < n_products(NP); ++t)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[prod_a(NP, t)][a], bv[prod_b(NP, t)][b],
                                                                acc[a][b], 0, 0, 0);
    }
  };
  const int nch = K / KC;
  load(0, ra[0], rb[0]);
  store(0, ra[0], rb[0]);
  if (nch > 1) load(KC, ra[0], rb[0]);
  if constexpr (PD == 1) {
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
      compute(c & 1);
      if (c + 1 < nch) {
        store((c + 1) & 1, ra[0], rb[0]);
        if (c + 2 < nch) load((c + 2) * KC, ra[0], rb[0]);
      }
      __syncthreads();
    }
  } else {
    // loads run two chunks ahead of the split into LDS (two compute phases of latency cover)
    if (nch > 2) load(2 * KC, ra[1], rb[1]);
    __syncthreads();
    for (int c = 0; c < nch; c += 2) {
      compute(0);  // chunk c (even) in stage 0
      if (c + 1 < nch) {
        store(1, ra[0], rb[0]);
        if (c + 3 < nch) load((c + 3) * KC, ra[0], rb[0]);
      }
      __syncthreads();
      if (c + 1 >= nch) break;
      compute(1);  // chunk c + 1 in stage 1
      if (c + 2 < nch) {
        store(0, ra[1], rb[1]);
        if (c + 4 < nch) load((c + 4) * KC, ra[1], rb[1]);
      }
      __syncthreads();
    }
  }
  store_tile<false>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
}

// Deep-A form of the double-buffered kernel: chunk c + PA (its A values from HBM and its weight
// planes from L2) is loaded into a ring of register buffers while chunk c is multiplied, and
// split into LDS one chunk ahead.  Every chunk issues the same loads in the same order, all of
// them unconditionally (rows past M read row M - 1, columns past N column N - 1, chunks past
// the end the last chunk again; none of these is ever stored), so the wait before a chunk's
// split is a memory counter of (PA - 1) chunks' loads and PA - 1 chunks stay in flight: the
// bytes in flight per CU are what bounds the staged forms on the tall, narrow contractions
// of the tangent rows (a few hundred columns, K = 256).  A load under a lane condition, or a
// weight load issued after the A loads it has to be waited with, would force a wait for every
// outstanding load at each chunk.
template <int NP, int BM, int BN, int KC, int PA, bool FUSE = false, bool F16 = false, bool PERSIST = false>
__global__ void __launch_bounds__(BM * BN / 64) k_gemm_bf16_dq(const float* __restrict__ A, int64_t lda,
                                                              const uint16_t* __restrict__ Wp0,
                                                              const uint16_t* __restrict__ Wp1,
                                                              const uint16_t* __restrict__ Wp2,
                                                              const float* __restrict__ bias, float* __restrict__ C,
                                                              int64_t ldc, int64_t M, int N, int K,
                                                              const uint8_t* __restrict__ rowflag, int period, int nT,
                                                              DevRows dr, ActEpi ep) {
  constexpr int NWN = BN / 64;
  constexpr int NT = (BM / 64) * NWN * 64;
  constexpr int LDH = KC + 8;
  constexpr int A_UNITS = BM * (KC / 4), B_UNITS = BN * (KC / 8);
  constexpr int A_PER = (A_UNITS + NT - 1) / NT;
  constexpr int B_PER = (B_UNITS + NT - 1) / NT;
  constexpr size_t STAGE = (size_t)NP * (BM + BN) * LDH;
  extern __shared__ __attribute__((aligned(16))) unsigned char gemm_lds[];
  uint16_t* lds = reinterpret_cast<uint16_t*>(gemm_lds);
  auto As = [&](int st, int p, int r) { return lds + st * STAGE + ((size_t)p * BM + r) * LDH; };
  auto Bs = [&](int st, int p, int r) { return lds + st * STAGE + (size_t)NP * BM * LDH + ((size_t)p * BN + r) * LDH; };
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t mt;
  int nt;
  if constexpr (PERSIST) {
    mt = blockIdx.x;
    nt = 0;
  } else {
    tile_of(mt, nt, nT);
  }
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
    if (ep.amax) ep.amax += o;
    if (ep.omax) ep.omax += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  const int64_t MT = (M + BM - 1) / BM;  // row tiles
  if (mt >= MT) return;
  int64_t m0 = mt * BM;
  int n0 = nt * BN;
  const int wm = (wid / NWN) * 64, wn = (wid % NWN) * 64;
  float16_t acc[2][2];
  const uint16_t* Wpl[3] = {Wp0, Wp1, Wp2};
  // the A rows (with their FP16 scales) and weight columns this thread loads for the tile at
  // (m0, n0), the same in every chunk, clamped into the operands; M >= 1 and N >= 1 here
  const float* arow[A_PER];
  float rsc[A_PER];
  int64_t bcol[B_PER];
  auto setup = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = (tid + q * NT) % A_UNITS;
      const int64_t gr = m0 + e / (KC / 4) < M ? m0 + e / (KC / 4) : M - 1;
      arow[q] = A + gr * lda + (e % (KC / 4)) * 4;
      rsc[q] = F16 ? f16_row_scale(ep.amax[gr]) : 1.f;
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = (tid + q * NT) % B_UNITS;
      const int gn = n0 + e / (KC / 8) < N ? n0 + e / (KC / 8) : N - 1;
      bcol[q] = wpl_off(gn, (e % (KC / 8)) * 8, N);
    }
  };
  setup();
  const int nch = K / KC;
  float4 ra[PA][A_PER];
  u32x4_t rb[PA][NP][B_PER];
  auto load = [&](int c, float4(&xa)[A_PER], u32x4_t(&xb)[NP][B_PER]) __attribute__((always_inline)) {
    const int k0 = (c < nch ? c : nch - 1) * KC;
#pragma unroll
    for (int q = 0; q < A_PER; ++q) xa[q] = *reinterpret_cast<const float4*>(arow[q] + k0);
#pragma unroll
    for (int q = 0; q < B_PER; ++q)
#pragma unroll
      for (int p = 0; p < NP; ++p) xb[p][q] = *reinterpret_cast<const u32x4_t*>(Wpl[p] + bcol[q] + (int64_t)k0 * N);
  };
  auto store = [&](int st, const float4(&xa)[A_PER], const u32x4_t(&xb)[NP][B_PER]) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * NT;
      if (A_UNITS % NT != 0 && e >= A_UNITS) continue;
      const int rr = e / (KC / 4), kq = (e % (KC / 4)) * 4;
      uint2 pw[NP];
      split_planes<NP, F16>(xa[q], rsc[q], pw);
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<uint2*>(As(st, p, rr) + kq) = pw[p];
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      if (B_UNITS % NT != 0 && e >= B_UNITS) continue;
      const int nn = e / (KC / 8), kq = (e % (KC / 8)) * 8;
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<u32x4_t*>(Bs(st, p, nn) + kq) = xb[p][q];
    }
  };
  auto compute = [&](int st) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 16) {
      const int ka = kk + 8 * (lane >> 5);
      OpT<F16> av[NP][2], bv[NP][2];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int a = 0; a < 2; ++a) av[p][a] = *reinterpret_cast<const OpT<F16>*>(As(st, p, wm + a * 32 + (lane & 31)) + ka);
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[p][b] = *reinterpret_cast<const OpT<F16>*>(Bs(st, p, wn + b * 32 + (lane & 31)) + ka);
      }
#pragma unroll
      for (int t = 0; t < n_products(NP); ++t)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[a][b] = mfma16(av[prod_a(NP, t)][a], bv[prod_b(NP, t)][b],
                                                                acc[a][b], 0, 0, 0);
    }
  };
  // chunk j lives in ring slot j % PA: chunks 0 .. PA - 1 first, chunk 0 split into stage 0,
  // chunk PA into its slot
  static_for<0, PA>([&](auto J) __attribute__((always_inline)) { load(J.value, ra[J.value], rb[J.value]); });
  // PERSIST: the block walks row tiles blockIdx.x, + gridDim.x, ..., all N tiles of each in
  // turn; the first PA chunks of the next tile are loaded before the current tile's C stores
  while (true) {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    store(0, ra[0], rb[0]);
    load(PA, ra[0], rb[0]);
    __syncthreads();
    for (int c = 0; c < nch; c += PA) {
      static_for<0, PA>([&](auto Q) __attribute__((always_inline)) {
        constexpr int slot = (decltype(Q)::value + 1) % PA;
        const int cc = c + decltype(Q)::value;  // chunk cc in stage cc & 1; chunk cc + 1 in `slot`
        if (cc < nch) compute(cc & 1);
        if (cc + 1 < nch) store((cc + 1) & 1, ra[slot], rb[slot]);
        load(cc + 1 + PA, ra[slot], rb[slot]);
        __syncthreads();
      });
    }
    if constexpr (!PERSIST) {
      store_tile<FUSE, F16>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
      return;
    } else {
      const int64_t m0c = m0;
      const int n0c = n0;
      if (nt + 1 < nT) {
        ++nt;
      } else {
        nt = 0;
        mt += gridDim.x;
      }
      const bool more = mt < MT;
      if (more) {
        m0 = mt * BM;
        n0 = nt * BN;
        setup();
        static_for<0, PA>([&](auto J) __attribute__((always_inline)) { load(J.value, ra[J.value], rb[J.value]); });
      }
      store_tile<FUSE, F16>(acc, m0c, n0c, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
      if (!more) return;
    }
  }
}

// Wide-wavefront form of the deep-A kernel: each wavefront computes 64 x 128 (2 x 4 MFMA
// tiles, 128 accumulators), so a K step reads 6 NP operand fragments from LDS for 8 tiles
// instead of 4 NP for 4, and a 256 x 256 block stages each weight chunk once for 256 rows.
// Both cut the LDS traffic per output, which bounds the 64 x 64 forms.
template <int NP, int BM, int BN, int KC, int PA, bool FUSE = false, bool F16 = false, bool PERSIST = false>
__global__ void __launch_bounds__(BM * BN / 128) __attribute__((amdgpu_waves_per_eu(2))) k_gemm_bf16_dw(const float* __restrict__ A, int64_t lda,
                                                              const uint16_t* __restrict__ Wp0,
                                                              const uint16_t* __restrict__ Wp1,
                                                              const uint16_t* __restrict__ Wp2,
                                                              const float* __restrict__ bias, float* __restrict__ C,
                                                              int64_t ldc, int64_t M, int N, int K,
                                                              const uint8_t* __restrict__ rowflag, int period, int nT,
                                                              DevRows dr, ActEpi ep) {
  constexpr int NWN = BN / 128;
  constexpr int NT = (BM / 64) * NWN * 64;
  constexpr int LDH = KC + 8;
  constexpr int A_UNITS = BM * (KC / 4), B_UNITS = BN * (KC / 8);
  constexpr int A_PER = (A_UNITS + NT - 1) / NT;
  constexpr int B_PER = (B_UNITS + NT - 1) / NT;
  constexpr size_t STAGE = (size_t)NP * (BM + BN) * LDH;
  extern __shared__ __attribute__((aligned(16))) unsigned char gemm_lds[];
  uint16_t* lds = reinterpret_cast<uint16_t*>(gemm_lds);
  auto As = [&](int st, int p, int r) { return lds + st * STAGE + ((size_t)p * BM + r) * LDH; };
  auto Bs = [&](int st, int p, int r) { return lds + st * STAGE + (size_t)NP * BM * LDH + ((size_t)p * BN + r) * LDH; };
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t mt;
  int nt;
  if constexpr (PERSIST) {
    mt = blockIdx.x;
    nt = 0;
  } else {
    tile_of(mt, nt, nT);
  }
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
    if (ep.amax) ep.amax += o;
    if (ep.omax) ep.omax += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  const int64_t MT = (M + BM - 1) / BM;  // row tiles
  if (mt >= MT) return;
  int64_t m0 = mt * BM;
  int n0 = nt * BN;
  const int wm = (wid / NWN) * 64, wn = (wid % NWN) * 128;
  float16_t acc[2][4];
  const uint16_t* Wpl[3] = {Wp0, Wp1, Wp2};
  // the A rows (with their FP16 scales) and weight columns this thread loads for the tile at
  // (m0, n0), the same in every chunk, clamped into the operands; M >= 1 and N >= 1 here
  const float* arow[A_PER];
  float rsc[A_PER];
  int64_t bcol[B_PER];
  auto setup = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = (tid + q * NT) % A_UNITS;
      const int64_t gr = m0 + e / (KC / 4) < M ? m0 + e / (KC / 4) : M - 1;
      arow[q] = A + gr * lda + (e % (KC / 4)) * 4;
      rsc[q] = F16 ? f16_row_scale(ep.amax[gr]) : 1.f;
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = (tid + q * NT) % B_UNITS;
      const int gn = n0 + e / (KC / 8) < N ? n0 + e / (KC / 8) : N - 1;
      bcol[q] = wpl_off(gn, (e % (KC / 8)) * 8, N);
    }
  };
  setup();
  const int nch = K / KC;
  float4 ra[PA][A_PER];
  u32x4_t rb[PA][NP][B_PER];
  auto load = [&](int c, float4(&xa)[A_PER], u32x4_t(&xb)[NP][B_PER]) __attribute__((always_inline)) {
    const int k0 = (c < nch ? c : nch - 1) * KC;
#pragma unroll
    for (int q = 0; q < A_PER; ++q) xa[q] = *reinterpret_cast<const float4*>(arow[q] + k0);
#pragma unroll
    for (int q = 0; q < B_PER; ++q)
#pragma unroll
      for (int p = 0; p < NP; ++p) xb[p][q] = *reinterpret_cast<const u32x4_t*>(Wpl[p] + bcol[q] + (int64_t)k0 * N);
  };
  auto store = [&](int st, const float4(&xa)[A_PER], const u32x4_t(&xb)[NP][B_PER]) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * NT;
      if (A_UNITS % NT != 0 && e >= A_UNITS) continue;
      const int rr = e / (KC / 4), kq = (e % (KC / 4)) * 4;
      uint2 pw[NP];
      split_planes<NP, F16>(xa[q], rsc[q], pw);
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<uint2*>(As(st, p, rr) + kq) = pw[p];
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      if (B_UNITS % NT != 0 && e >= B_UNITS) continue;
      const int nn = e / (KC / 8), kq = (e % (KC / 8)) * 8;
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<u32x4_t*>(Bs(st, p, nn) + kq) = xb[p][q];
    }
  };
  const int nbl = live_col_tiles(n0 + wn, N, 4);
  auto compute_nb = [&](int st, auto NBL) {
    constexpr int nbc = decltype(NBL)::value;
#pragma unroll
    for (int kk = 0; kk < KC; kk += 16) {
      const int ka = kk + 8 * (lane >> 5);
      OpT<F16> av[NP][2];
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int a = 0; a < 2; ++a) av[p][a] = *reinterpret_cast<const OpT<F16>*>(As(st, p, wm + a * 32 + (lane & 31)) + ka);
#pragma unroll
      for (int b = 0; b < nbc; ++b) {
        OpT<F16> bv[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) bv[p] = *reinterpret_cast<const OpT<F16>*>(Bs(st, p, wn + b * 32 + (lane & 31)) + ka);
#pragma unroll
        for (int t = 0; t < n_products(NP); ++t)
#pragma unroll
          for (int a = 0; a < 2; ++a)
            acc[a][b] = mfma16(av[prod_a(NP, t)][a], bv[prod_b(NP, t)], acc[a][b], 0,
                                                                0, 0);
      }
    }
  };
  auto compute = [&](int st) {
    if (nbl == 4)
      compute_nb(st, IC<4>());
    else if (nbl == 3)
      compute_nb(st, IC<3>());
    else if (nbl == 2)
      compute_nb(st, IC<2>());
    else if (nbl == 1)
      compute_nb(st, IC<1>());
  };
  // chunk j lives in ring slot j % PA: chunks 0 .. PA - 1 first, chunk 0 split into stage 0,
  // chunk PA into its slot
  static_for<0, PA>([&](auto J) __attribute__((always_inline)) { load(J.value, ra[J.value], rb[J.value]); });
  // PERSIST: the block walks row tiles blockIdx.x, + gridDim.x, ..., all N tiles of each in
  // turn (the A rows come back from the cache for the second); the first PA chunks of the
  // next tile are loaded before the C stores of the current one, so those stores overlap loads
  while (true) {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 4; ++b)
        for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    store(0, ra[0], rb[0]);
    load(PA, ra[0], rb[0]);
    __syncthreads();
    for (int c = 0; c < nch; c += PA) {
      static_for<0, PA>([&](auto Q) __attribute__((always_inline)) {
        constexpr int slot = (decltype(Q)::value + 1) % PA;
        const int cc = c + decltype(Q)::value;  // chunk cc in stage cc & 1; chunk cc + 1 in `slot`
        if (cc < nch) compute(cc & 1);
        if (cc + 1 < nch) store((cc + 1) & 1, ra[slot], rb[slot]);
        load(cc + 1 + PA, ra[slot], rb[slot]);
        __syncthreads();
      });
    }
    if constexpr (!PERSIST) {
      store_rows<4, FUSE, F16>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
      return;
    } else {
      const int64_t m0c = m0;
      const int n0c = n0;
      if (nt + 1 < nT) {
        ++nt;
      } else {
        nt = 0;
        mt += gridDim.x;
      }
      const bool more = mt < MT;
      if (more) {
        m0 = mt * BM;
        n0 = nt * BN;
        setup();
        static_for<0, PA>([&](auto J) __attribute__((always_inline)) { load(J.value, ra[J.value], rb[J.value]); });
      }
      store_rows<4, FUSE, F16>(acc, m0c, n0c, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
      if (!more) return;
    }
  }
}

// A-in-registers form: each wavefront computes a 64 x WN tile; it loads its A fragments
// straight from memory into MFMA operand registers (lane l: row l % 32 of a 32-row tile,
// k = 8 (l / 32) .. +7 of each 16-k step), two chunks ahead, and splits them into the NP planes
// in registers.  Only the weight planes go through LDS (two stages of 16 k, one barrier per
// chunk), shared by the NWM x NWN wavefronts of the block (BM = 64 NWM rows, BN = WN NWN
// columns); the NWN wavefronts of a row range read the same A rows, served once from memory
// through the CU's cache.  Halves the LDS reads per MFMA of the staged forms, whose operands
// both come from LDS.
template <int NP, int NWM, int NWN, int WN>
__global__ void __launch_bounds__(NWM * NWN * 64) __attribute__((amdgpu_waves_per_eu(2))) k_gemm_bf16_ar(const float* __restrict__ A, int64_t lda,
                                                           const uint16_t* __restrict__ Wp0,
                                                           const uint16_t* __restrict__ Wp1,
                                                           const uint16_t* __restrict__ Wp2,
                                                           const float* __restrict__ bias, float* __restrict__ C,
                                                           int64_t ldc, int64_t M, int N, int K,
                                                           const uint8_t* __restrict__ rowflag, int period, int nT,
                                                           DevRows dr, ActEpi ep) {
  constexpr int NT = NWM * NWN * 64, BM = 64 * NWM, BN = WN * NWN, NB = WN / 32;
  constexpr int KC = 16, LDH = KC + 8;     // halves per LDS row: 48 bytes, 16-byte aligned
  constexpr int B_UNITS = BN * (KC / 8);   // uint4 per plane and chunk
  constexpr int B_PER = (B_UNITS + NT - 1) / NT;
  constexpr int STAGE = NP * BN * LDH;     // halves per stage
  constexpr int PA = 3;                    // A chunks in registers (loads two chunks ahead)
  __shared__ __attribute__((aligned(16))) uint16_t Ws[2 * STAGE];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t mt;
  int nt;
  tile_of(mt, nt, nT);
  if (dr.offset) {
    const int64_t o = *dr.offset;
    A += o * lda;
    C += o * ldc;
    if (rowflag) rowflag += o;
  }
  if (dr.count) M = M < (int64_t)*dr.count ? M : (int64_t)*dr.count;
  if (mt * BM >= M) return;
  const int64_t m0 = mt * BM;
  const int n0 = nt * BN;
  const int wm = (wid / NWN) * 64, wn = (wid % NWN) * WN;
  const uint16_t* Wpl[3] = {Wp0, Wp1, Wp2};
  // this lane's A rows (two 32-row tiles) and k offset inside a 16-k step
  const float* arow[2];
  bool aok[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t r = m0 + wm + a * 32 + (lane & 31);
    aok[a] = r < M;
    arow[a] = A + (aok[a] ? r : 0) * lda + 8 * (lane >> 5);
  }
  float4 ra[PA][2][2];  // [buffer][row tile][half of the 8 k]
  auto load_a = [&](int c, float4(&x)[2][2]) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        x[a][h] = aok[a] ? *reinterpret_cast<const float4*>(arow[a] + c * KC + 4 * h) : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  uint4 rb[NP][B_PER];
  auto load_b = [&](int c) {
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      const int nn = e >> 1, kq = (e & 1) * 8;
      const int gn = n0 + nn;
#pragma unroll
      for (int p = 0; p < NP; ++p)
        rb[p][q] = (e < B_UNITS && gn < N) ? *reinterpret_cast<const uint4*>(Wpl[p] + wpl_off(gn, c * KC + kq, N))
                                           : make_uint4(0, 0, 0, 0);
    }
  };
  auto store_b = [&](int st) {
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * NT;
      if (B_UNITS % NT != 0 && e >= B_UNITS) continue;
      const int nn = e >> 1, kq = (e & 1) * 8;
#pragma unroll
      for (int p = 0; p < NP; ++p) *reinterpret_cast<uint4*>(Ws + st * STAGE + (p * BN + nn) * LDH + kq) = rb[p][q];
    }
  };
  float16_t acc[2][NB];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  auto compute = [&](int st, const float4(&x)[2][2]) {
    // x = hi + mid + lo with hardware round-to-nearest conversions; each remainder is exact
    bf16x8_t av[NP][2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      float2v_t r[4] = {{x[a][0].x, x[a][0].y}, {x[a][0].z, x[a][0].w}, {x[a][1].x, x[a][1].y}, {x[a][1].z, x[a][1].w}};
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        bf16x2_t h[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          h[j] = __builtin_convertvector(r[j], bf16x2_t);
          if (p + 1 < NP) r[j] = r[j] - __builtin_convertvector(h[j], float2v_t);
        }
        av[p][a] = (bf16x8_t){h[0][0], h[0][1], h[1][0], h[1][1], h[2][0], h[2][1], h[3][0], h[3][1]};
      }
    }
    const int kb = 8 * (lane >> 5);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      bf16x8_t bv[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p)
        bv[p] = *reinterpret_cast<const bf16x8_t*>(Ws + st * STAGE + (p * BN + wn + b * 32 + (lane & 31)) * LDH + kb);
#pragma unroll
      for (int t = 0; t < n_products(NP); ++t)
#pragma unroll
        for (int a = 0; a < 2; ++a)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[prod_a(NP, t)][a], bv[prod_b(NP, t)], acc[a][b], 0, 0, 0);
    }
  };
  const int nch = K / KC;
  load_b(0);
  load_a(0, ra[0]);
  if (nch > 1) load_a(1, ra[1]);
  store_b(0);
  if (nch > 1) load_b(1);
  __syncthreads();
  // chunk c: A in ra[c % 3], W in stage c % 2; chunk c + 1's W is in registers, chunk c + 2's A
  // is loaded now
  for (int c = 0; c < nch; c += PA) {
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      const int cc = c + q;
      if (cc < nch) {
        if (cc + 2 < nch) load_a(cc + 2, ra[(q + 2) % PA]);
        compute(cc & 1, ra[q]);
        if (cc + 1 < nch) {
          store_b((cc + 1) & 1);
          if (cc + 2 < nch) load_b(cc + 2);
        }
        __syncthreads();
      }
    }
  }
  store_rows<NB, false>(acc, m0, n0, wm, wn, lane, C, ldc, M, N, bias, rowflag, period, ep);
}

// Tile shapes offered to the autotuner (index = variant).
struct Bf16Tile {
  int bm, bn;
};
// variants 0-6: single LDS stage, K chunk 32; variants 7-10: two LDS stages (kc = 16 or 32);
// variants 11-14: two LDS stages with loads two chunks ahead; variants 15-17: A in registers;
// variants 18-22 and 26: two LDS stages with the loads PA chunks ahead; variants 23-25: the same
// with 64 x 128 wavefront tiles; 27-28: those with one persistent block per CU; 29-30: 128 x 256
// two-stage tiles with persistent blocks
static const Bf16Tile kBf16Tiles[] = {{128, 128}, {128, 256}, {256, 128}, {64, 256}, {256, 64},  {64, 128},
                                      {128, 64},  {128, 128}, {128, 128}, {256, 128}, {128, 256}, {256, 128},
                                      {128, 256}, {128, 128}, {128, 128}, {128, 256}, {256, 128}, {64, 256},
                                      {128, 256}, {128, 256}, {256, 128}, {128, 256}, {128, 128}, {256, 256},
                                      {128, 256}, {256, 256}, {128, 256}, {256, 256}, {256, 256}, {128, 256},
                                      {128, 256}};

// forms that carry FP16 planes (the others are offered to BF16 planes only)
bool gemm_bf16_f16_ok(int variant) { return variant < 3 || variant > 17; }

// forms whose registers spill to scratch for this plane count (hipcc resource report): the
// 2 x 2 A-in-registers form with two or three planes; with three planes also the 4 x 1
// A-in-registers form and every wide-wavefront form.  They are not offered: a spilling kernel
// is slower, and every stream running it reserves scratch.
bool gemm_bf16_spills(int variant, int nsplit) {
  if (variant == 15) return nsplit != 1;
  return nsplit == 6 && (variant == 16 || variant == 23 || variant == 24 || variant == 25 || variant >= 27);
}

bool gemm_bf16_fusable(int variant) {
  return variant == 0 || variant == 2 || variant == 18 || variant == 20 || (variant >= 23 && variant != 24);
}

int gemm_bf16_variants() { return (int)(sizeof(kBf16Tiles) / sizeof(kBf16Tiles[0])); }

// forms whose blocks walk several row tiles (one block per CU or per resident slot)
bool gemm_bf16_persistent(int variant) { return variant >= 27 && variant <= 30; }
void gemm_bf16_tile(int variant, int* bm, int* bn) {
  const Bf16Tile t = kBf16Tiles[(variant >= 0 && variant < gemm_bf16_variants()) ? variant : 0];
  *bm = t.bm;
  *bn = t.bn;
}

constexpr size_t kMaxLds = 160 * 1024;  // LDS of one CU

template <int NP, int BM, int BN, bool FUSE = false, bool F16 = false>
static void launch_tile(hipStream_t s, const float* A, int64_t lda, const uint16_t* const* W, const float* bias,
                        float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period,
                        DevRows dr, ActEpi ep) {
  static bool attr_set = false;  // LDS above 64 KB must be enabled once per kernel
  const size_t lds = (size_t)NP * (BM + BN) * LDB_H * 2;
  if (lds > kMaxLds) {  // does not fit one CU: the base tile (same result) is launched instead
    launch_tile<NP, 128, 128, FUSE, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
    return;
  }
  if (!attr_set) {
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_gemm_bf16<NP, BM, BN, FUSE, F16>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_set = true;
  }
  const int nT = (N + BN - 1) / BN;
  const int64_t blocks = ((M + BM - 1) / BM) * nT;
  k_gemm_bf16<NP, BM, BN, FUSE, F16><<<(unsigned)blocks, BM * BN / 64, lds, s>>>(A, lda, W[0], W[1], W[2], bias, C, ldc,
                                                                              M, N, K, rowflag, period, nT, dr, ep);
}

template <int NP, int BM, int BN, int KC, int PD = 1>
static void launch_tile_db(hipStream_t s, const float* A, int64_t lda, const uint16_t* const* W, const float* bias,
                           float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period,
                           DevRows dr, ActEpi ep) {
  static bool attr_set = false;
  const size_t lds = (size_t)2 * NP * (BM + BN) * (KC + 8) * 2;
  if (lds > kMaxLds) {
    launch_tile<NP, 128, 128>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
    return;
  }
  if (!attr_set) {
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_gemm_bf16_db<NP, BM, BN, KC, PD>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_set = true;
  }
  const int nT = (N + BN - 1) / BN;
  const int64_t blocks = ((M + BM - 1) / BM) * nT;
  k_gemm_bf16_db<NP, BM, BN, KC, PD><<<(unsigned)blocks, BM * BN / 64, lds, s>>>(A, lda, W[0], W[1], W[2], bias, C,
                                                                               ldc, M, N, K, rowflag, period, nT, dr, ep);
}

// compute units of the current device (cached per device; launches come from several host
// threads, which all store the same value)
static int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) dev = 0;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

// PERSIST: per CU as many blocks as its registers and LDS hold, each walking row tiles
template <int NP, int BM, int BN, int KC, int PA, bool FUSE = false, bool F16 = false, bool PERSIST = false>
static void launch_tile_dq(hipStream_t s, const float* A, int64_t lda, const uint16_t* const* W, const float* bias,
                           float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period,
                           DevRows dr, ActEpi ep) {
  static bool attr_set = false;
  const size_t lds = (size_t)2 * NP * (BM + BN) * (KC + 8) * 2;
  if (lds > kMaxLds) {
    launch_tile<NP, 128, 128, FUSE, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
    return;
  }
  if (!attr_set) {
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_gemm_bf16_dq<NP, BM, BN, KC, PA, FUSE, F16, PERSIST>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_set = true;
  }
  const int nT = (N + BN - 1) / BN;
  const int64_t mtiles = (M + BM - 1) / BM;
  static std::atomic<int> resident{0};  // blocks per CU (registers and LDS), for the persistent grid
  int per_cu = resident.load(std::memory_order_relaxed);
  if (PERSIST && per_cu == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gemm_bf16_dq<NP, BM, BN, KC, PA, FUSE, F16, PERSIST>,
                                                     BM * BN / 64, lds) != hipSuccess || per_cu <= 0)
      per_cu = 1;
    resident.store(per_cu, std::memory_order_relaxed);
  }
  const int64_t blocks = PERSIST ? std::min<int64_t>(mtiles, (int64_t)per_cu * device_cus()) : mtiles * nT;
  k_gemm_bf16_dq<NP, BM, BN, KC, PA, FUSE, F16, PERSIST><<<(unsigned)blocks, BM * BN / 64, lds, s>>>(
      A, lda, W[0], W[1], W[2], bias, C, ldc, M, N, K, rowflag, period, nT, dr, ep);
}

// PERSIST: one block per CU (the form's LDS and registers admit one) walking the row tiles
template <int NP, int BM, int BN, int KC, int PA, bool FUSE = false, bool F16 = false, bool PERSIST = false>
static void launch_tile_dw(hipStream_t s, const float* A, int64_t lda, const uint16_t* const* W, const float* bias,
                           float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period,
                           DevRows dr, ActEpi ep) {
  static bool attr_set = false;
  const size_t lds = (size_t)2 * NP * (BM + BN) * (KC + 8) * 2;
  if (lds > kMaxLds) {
    launch_tile<NP, 128, 128, FUSE, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
    return;
  }
  if (!attr_set) {
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_gemm_bf16_dw<NP, BM, BN, KC, PA, FUSE, F16, PERSIST>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_set = true;
  }
  const int nT = (N + BN - 1) / BN;
  const int64_t mtiles = (M + BM - 1) / BM;
  const int64_t blocks = PERSIST ? std::min<int64_t>(mtiles, device_cus()) : mtiles * nT;
  k_gemm_bf16_dw<NP, BM, BN, KC, PA, FUSE, F16, PERSIST><<<(unsigned)blocks, BM * BN / 128, lds, s>>>(
      A, lda, W[0], W[1], W[2], bias, C, ldc, M, N, K, rowflag, period, nT, dr, ep);
}

template <int NP, int NWM, int NWN, int WN>
static void launch_tile_ar(hipStream_t s, const float* A, int64_t lda, const uint16_t* const* W, const float* bias,
                           float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period,
                           DevRows dr, ActEpi ep) {
  const int nT = (N + WN * NWN - 1) / (WN * NWN);
  const int64_t blocks = ((M + 64 * NWM - 1) / (64 * NWM)) * nT;
  k_gemm_bf16_ar<NP, NWM, NWN, WN><<<(unsigned)blocks, NWM * NWN * 64, 0, s>>>(A, lda, W[0], W[1], W[2], bias, C, ldc,
                                                                              M, N, K, rowflag, period, nT, dr, ep);
}

template <int NP, bool F16 = false>
static void launch_np(hipStream_t s, int variant, const float* A, int64_t lda, const uint16_t* const* W,
                      const float* bias, float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag,
                      int period, DevRows dr, ActEpi ep) {
  if (ep.rowgrp) {  // fused silu epilogue: the shapes of gemm_bf16_fusable
    switch (variant) {
      case 2: launch_tile<NP, 256, 128, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 18: launch_tile_dq<NP, 128, 256, 16, 4, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 20: launch_tile_dq<NP, 256, 128, 16, 4, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 23: launch_tile_dw<NP, 256, 256, 16, 4, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 25: launch_tile_dw<NP, 256, 256, 16, 2, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 26: launch_tile_dq<NP, 128, 256, 16, 1, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 27: launch_tile_dw<NP, 256, 256, 16, 4, true, F16, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 28: launch_tile_dw<NP, 256, 256, 16, 2, true, F16, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 29: launch_tile_dq<NP, 128, 256, 16, 1, true, F16, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 30: launch_tile_dq<NP, 128, 256, 16, 2, true, F16, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      default: launch_tile<NP, 128, 128, true, F16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
    }
    return;
  }
  if constexpr (F16) {  // the forms with FP16 planes: single stage, deep A, wide wavefront
    switch (variant) {
      case 18: launch_tile_dq<NP, 128, 256, 16, 4, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 19: launch_tile_dq<NP, 128, 256, 16, 2, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 20: launch_tile_dq<NP, 256, 128, 16, 4, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 21: launch_tile_dq<NP, 128, 256, 32, 2, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 22: launch_tile_dq<NP, 128, 128, 16, 3, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 23: launch_tile_dw<NP, 256, 256, 16, 4, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 24: launch_tile_dw<NP, 128, 256, 16, 3, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 25: launch_tile_dw<NP, 256, 256, 16, 2, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 26: launch_tile_dq<NP, 128, 256, 16, 1, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 27: launch_tile_dw<NP, 256, 256, 16, 4, false, true, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 28: launch_tile_dw<NP, 256, 256, 16, 2, false, true, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 29: launch_tile_dq<NP, 128, 256, 16, 1, false, true, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 30: launch_tile_dq<NP, 128, 256, 16, 2, false, true, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 1: launch_tile<NP, 128, 256, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 2: launch_tile<NP, 256, 128, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      default: launch_tile<NP, 128, 128, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
    }
    return;
  } else {
    switch (variant) {
      case 7: launch_tile_db<NP, 128, 128, 16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 8: launch_tile_db<NP, 128, 128, 32>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 9: launch_tile_db<NP, 256, 128, 16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 10: launch_tile_db<NP, 128, 256, 16>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 11: launch_tile_db<NP, 256, 128, 16, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 12: launch_tile_db<NP, 128, 256, 16, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 13: launch_tile_db<NP, 128, 128, 32, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 14: launch_tile_db<NP, 128, 128, 16, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 15: launch_tile_ar<NP, 2, 2, 128>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 16: launch_tile_ar<NP, 4, 1, 128>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 17: launch_tile_ar<NP, 1, 4, 64>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 18: launch_tile_dq<NP, 128, 256, 16, 4>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 19: launch_tile_dq<NP, 128, 256, 16, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 20: launch_tile_dq<NP, 256, 128, 16, 4>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 21: launch_tile_dq<NP, 128, 256, 32, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 22: launch_tile_dq<NP, 128, 128, 16, 3>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 23: launch_tile_dw<NP, 256, 256, 16, 4>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 24: launch_tile_dw<NP, 128, 256, 16, 3>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 25: launch_tile_dw<NP, 256, 256, 16, 2>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 26: launch_tile_dq<NP, 128, 256, 16, 1>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 27: launch_tile_dw<NP, 256, 256, 16, 4, false, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 28: launch_tile_dw<NP, 256, 256, 16, 2, false, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 29: launch_tile_dq<NP, 128, 256, 16, 1, false, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 30: launch_tile_dq<NP, 128, 256, 16, 2, false, false, true>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 1: launch_tile<NP, 128, 256>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 2: launch_tile<NP, 256, 128>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 3: launch_tile<NP, 64, 256>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 4: launch_tile<NP, 256, 64>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 5: launch_tile<NP, 64, 128>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      case 6: launch_tile<NP, 128, 64>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
      default: launch_tile<NP, 128, 128>(s, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep); break;
    }
  }
}

void launch_gemm_bf16(hipStream_t s, const float* A, int64_t lda, const uint16_t* Whi, const uint16_t* Wlo,
                      const uint16_t* Wlo2, const float* bias, float* C, int64_t ldc, int64_t M, int N, int K,
                      const uint8_t* rowflag, int period, int nsplit, int variant, DevRows dr, ActEpi ep) {
  if (M == 0) return;
  const uint16_t* W[3] = {Whi, Wlo ? Wlo : Whi, Wlo2 ? Wlo2 : Whi};
  if (nsplit == 4)  // two FP16 planes with scales (ep.amax, ep.winv), three products
    launch_np<2, true>(s, variant, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
  else if (nsplit == 6)
    launch_np<3>(s, variant, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
  else if (nsplit == 3)
    launch_np<2>(s, variant, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
  else
    launch_np<1>(s, variant, A, lda, W, bias, C, ldc, M, N, K, rowflag, period, dr, ep);
}

}  // namespace fire
