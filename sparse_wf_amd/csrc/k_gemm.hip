// Dense contractions over tangent rows: C[M,N] = A[M,K] @ W[K,N] (+ bias on value rows).
//
// These are the Linear layers of the reference applied to NodeWithFwdLap (value, every
// jacobian row and the Laplacian row go through the same matrix, the bias only on the
// value row: sparse_fwd_lap.py:160-181).  Value rows are marked by a per-row flag or by
// row % period == 0.
//
// MFMA kernel: 256 threads = 4 wavefronts, 128x128 output tile, each wavefront a 64x64
// quarter made of 2x2 32x32 MFMA tiles; K staged through LDS 32 at a time.
//   FP32 path : v_mfma_f32_32x32x2f32 (A, W in FP32)
//   MIXED path: v_mfma_f32_32x32x8bf16_1k (A rounded to BF16 on load, W pre-rounded), FP32 accumulate.
#include "fire_common.hpp"
#include "fire_kernels.hpp"

namespace fire {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

#define GT_M 128
#define GT_N 128
#define GT_K 32
#define LDA_S (GT_K + 4)   // fp32 LDS row stride (floats)
#define LDB_S (GT_K + 8)   // bf16 LDS row stride (halves)

__device__ __forceinline__ bool is_bias_row(int64_t row, const uint8_t* rowflag, int period) {
  if (rowflag) return rowflag[row] != 0;
  if (period > 0) return (row % period) == 0;
  return false;
}

// ---------------------------------------------------------------------------------------
// FP32 MFMA GEMM.  W is [K,N] FP32 row-major.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gemm_f32(const float* __restrict__ A, int64_t lda,
                                                  const float* __restrict__ W, const float* __restrict__ bias,
                                                  float* __restrict__ C, int64_t ldc, int64_t M, int N, int K,
                                                  const uint8_t* __restrict__ rowflag, int period) {
  __shared__ float As[GT_M][LDA_S];
  __shared__ float Bs[GT_N][LDA_S];  // transposed: Bs[n][k]
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * GT_M;
  const int n0 = blockIdx.y * GT_N;
  const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
  float16_t acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

  for (int k0 = 0; k0 < K; k0 += GT_K) {
    // A tile 128x32: 4096 floats, 16 per thread
    for (int e = tid; e < GT_M * GT_K; e += 256) {
      const int rr = e / GT_K, kk = e % GT_K;
      const int64_t gr = m0 + rr;
      const int gk = k0 + kk;
      As[rr][kk] = (gr < M && gk < K) ? A[gr * lda + gk] : 0.f;
    }
    // W tile 32x128 -> Bs[n][k]
    for (int e = tid; e < GT_N * GT_K; e += 256) {
      const int kk = e / GT_N, nn = e % GT_N;
      const int gk = k0 + kk, gn = n0 + nn;
      Bs[nn][kk] = (gk < K && gn < N) ? W[(int64_t)gk * N + gn] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GT_K; kk += 2) {
      const int ka = kk + (lane >> 5);
      float av[2], bv[2];
      for (int a = 0; a < 2; ++a) av[a] = As[wm + a * 32 + (lane & 31)][ka];
      for (int b = 0; b < 2; ++b) bv[b] = Bs[wn + b * 32 + (lane & 31)][ka];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int col = n0 + wn + b * 32 + (lane & 31);
      if (col >= N) continue;
      const float bcol = bias ? bias[col] : 0.f;
      for (int v = 0; v < 16; ++v) {
        const int64_t row = m0 + wm + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row >= M) continue;
        float val = acc[a][b][v];
        if (bias && is_bias_row(row, rowflag, period)) val += bcol;
        C[row * ldc + col] = val;
      }
    }
}

// ---------------------------------------------------------------------------------------
// BF16 MFMA GEMM, FP32 accumulation.  Wt is [N,K] BF16 (W transposed, pre-rounded).
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gemm_bf16(const float* __restrict__ A, int64_t lda,
                                                   const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                                   float* __restrict__ C, int64_t ldc, int64_t M, int N, int K,
                                                   const uint8_t* __restrict__ rowflag, int period) {
  __shared__ uint16_t As[GT_M][LDB_S];
  __shared__ uint16_t Bs[GT_N][LDB_S];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * GT_M;
  const int n0 = blockIdx.y * GT_N;
  const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
  float16_t acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

  for (int k0 = 0; k0 < K; k0 += GT_K) {
    // A tile: 128 rows x 32 k, float4 loads (K is a multiple of 32, lda of 4)
    for (int e = tid; e < GT_M * GT_K / 4; e += 256) {
      const int rr = e / (GT_K / 4), kq = (e % (GT_K / 4)) * 4;
      const int64_t gr = m0 + rr;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gr < M) v = *reinterpret_cast<const float4*>(A + gr * lda + k0 + kq);
      As[rr][kq + 0] = f2bf(v.x);
      As[rr][kq + 1] = f2bf(v.y);
      As[rr][kq + 2] = f2bf(v.z);
      As[rr][kq + 3] = f2bf(v.w);
    }
    // W tile: 128 n x 32 k from Wt[n][k], 8-byte loads
    for (int e = tid; e < GT_N * GT_K / 4; e += 256) {
      const int nn = e / (GT_K / 4), kq = (e % (GT_K / 4)) * 4;
      const int gn = n0 + nn;
      uint2 v = make_uint2(0u, 0u);
      if (gn < N) v = *reinterpret_cast<const uint2*>(Wt + (int64_t)gn * K + k0 + kq);
      *reinterpret_cast<uint2*>(&Bs[nn][kq]) = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GT_K; kk += 8) {
      const int ka = kk + 4 * (lane >> 5);
      short4_t av[2], bv[2];
      for (int a = 0; a < 2; ++a) av[a] = *reinterpret_cast<const short4_t*>(&As[wm + a * 32 + (lane & 31)][ka]);
      for (int b = 0; b < 2; ++b) bv[b] = *reinterpret_cast<const short4_t*>(&Bs[wn + b * 32 + (lane & 31)][ka]);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int col = n0 + wn + b * 32 + (lane & 31);
      if (col >= N) continue;
      const float bcol = bias ? bias[col] : 0.f;
      for (int v = 0; v < 16; ++v) {
        const int64_t row = m0 + wm + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row >= M) continue;
        float val = acc[a][b][v];
        if (bias && is_bias_row(row, rowflag, period)) val += bcol;
        C[row * ldc + col] = val;
      }
    }
}

// Small-N FP32 GEMM (N <= 16): one wavefront per row, lanes across K.
__global__ void k_gemm_smalln(const float* __restrict__ A, int64_t lda, const float* __restrict__ W,
                              const float* __restrict__ bias, float* __restrict__ C, int64_t ldc, int64_t M, int N,
                              int K, const uint8_t* __restrict__ rowflag, int period) {
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= M) return;
  float acc[16];
  for (int n = 0; n < N; ++n) acc[n] = 0.f;
  for (int k = lane; k < K; k += 64) {
    const float a = A[row * lda + k];
    for (int n = 0; n < N; ++n) acc[n] += a * W[(int64_t)k * N + n];
  }
  const bool b = bias && is_bias_row(row, rowflag, period);
  for (int n = 0; n < N; ++n) {
    float v = warp_sum(acc[n]);
    if (lane == 0) C[row * ldc + n] = v + (b ? bias[n] : 0.f);
  }
}

void launch_gemm_f32(hipStream_t s, const float* A, int64_t lda, const float* W, const float* bias, float* C,
                     int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period) {
  if (M == 0) return;
  if (N <= 16) {
    const int64_t threads = M * 64;
    k_gemm_smalln<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(A, lda, W, bias, C, ldc, M, N, K, rowflag, period);
    return;
  }
  dim3 grid((unsigned)((M + GT_M - 1) / GT_M), (N + GT_N - 1) / GT_N);
  k_gemm_f32<<<grid, 256, 0, s>>>(A, lda, W, bias, C, ldc, M, N, K, rowflag, period);
}

void launch_gemm_bf16(hipStream_t s, const float* A, int64_t lda, const uint16_t* Wt, const float* bias, float* C,
                      int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period) {
  if (M == 0) return;
  dim3 grid((unsigned)((M + GT_M - 1) / GT_M), (N + GT_N - 1) / GT_N);
  k_gemm_bf16<<<grid, 256, 0, s>>>(A, lda, Wt, bias, C, ldc, M, N, K, rowflag, period);
}

}  // namespace fire
