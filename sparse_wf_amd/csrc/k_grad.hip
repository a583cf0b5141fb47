// Parameter gradient of log|psi| per walker (value path): the jax.grad(log_p) whose batch the
// natural-gradient preconditioner contracts (src/sparse_wf/preconditioner.py:191-194).
//
// The gradient is taken with respect to the packed parameter tensors of fire_upload_params;
// the host maps it onto the reference's parameter tree (softplus, folded normalisation scales,
// padded filter widths, spin-restricted columns).  The forward intermediates come from the
// value path in FP32 and from the recomputations below; every per-walker sum over electrons,
// pairs or e-n edges is a fixed-order loop (no atomics), so results are deterministic.
//
// Row layouts: per-electron rows in group order (up-spin electrons of all walkers first,
// walker-major within a spin); e-e pair rows in the CSR order of the neighbour lists; e-n edge
// rows as group * M + nucleus.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fire_common.hpp"
#include "fire_kernels.hpp"
#include "fire_grad.hpp"

namespace fire {

namespace {

__device__ __forceinline__ float sigm(float x) { return 1.0f / (1.0f + expf(-x)); }
__device__ __forceinline__ float silu_f(float x) { return x * sigm(x); }
__device__ __forceinline__ float silu_d(float x) {
  const float s = sigm(x);
  return s * (1.0f + x * (1.0f - s));
}

// cutoff polynomial 1 - 15 x^4 + 24 x^5 - 10 x^6 (x < 1) and its derivative
__device__ __forceinline__ void cutoff_poly(float x, float& f, float& df) {
  if (x >= 1.0f) {
    f = 0.f;
    df = 0.f;
    return;
  }
  const float x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
  f = 1.0f - 15.0f * x4 + 24.0f * x4 * x - 10.0f * x4 * x2;
  df = -60.0f * x3 + 120.0f * x4 - 60.0f * x4 * x;
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }
__device__ __forceinline__ float4 fma4(float a, float4 b, float4 c) {
  return make_float4(fmaf(a, b.x, c.x), fmaf(a, b.y, c.y), fmaf(a, b.z, c.z), fmaf(a, b.w, c.w));
}
__device__ __forceinline__ float4 mul4(float4 a, float4 b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }

__device__ __forceinline__ void group_walker(int64_t g, int B, int nu, int nd, int& w, int& i) {
  const int64_t up = (int64_t)B * nu;
  if (g < up) {
    w = (int)(g / nu);
    i = (int)(g - (int64_t)w * nu);
  } else {
    const int64_t q = g - up;
    w = (int)(q / nd);
    i = nu + (int)(q - (int64_t)w * nd);
  }
}

__device__ __forceinline__ int64_t group_index(int w, int i, int B, int nu, int nd) {
  return i < nu ? (int64_t)w * nu + i : (int64_t)B * nu + (int64_t)w * nd + (i - nu);
}

}  // namespace

// ---------------------------------------------------------------------------------------
// out[b][k][j] (+)= sum over the rows of walker b of X[row][k] Y[row][j] (X == nullptr: 1).
// Rows of walker b: [rng[4b], rng[4b+1]) and [rng[4b+2], rng[4b+3]), or without rng the fixed
// ranges a0 + b as .. + na and c0 + b cs .. + nc.  rowmask (optional): only rows with
// rowmask[row] == want.  One block per (64 x 64 output tile, walker, slice): with S > 1 slices
// the walker's rows (both ranges, concatenated) are cut into S consecutive parts whose partial
// products go to part[(slice B + b) K J ..] and are summed in slice order by k_wgrad_reduce;
// with S == 1 the block writes out directly.  16 rows per LDS chunk.
// ---------------------------------------------------------------------------------------
template <int TK, int TJ>
__global__ void __launch_bounds__(256) k_wgrad(const float* __restrict__ X, int64_t ldx, const float* __restrict__ Y,
                                               int64_t ldy, int K, int J, WRows wr, const uint8_t* __restrict__ rowmask,
                                               int want, float* __restrict__ out, int64_t so, int accumulate,
                                               float* __restrict__ part) {
  static_assert(TK * TJ == 4096, "16 outputs per thread");
  constexpr int TXN = TJ / 4;  // threads along j; each takes columns tx, tx + TXN, tx + 2 TXN, tx + 3 TXN
  __shared__ float xs[16][TK + 1];
  __shared__ float ys[16][TJ + 1];
  const int b = blockIdx.y, S = gridDim.z, sl = blockIdx.z;
  const int tk = (K + TK - 1) / TK;
  const int k0 = (blockIdx.x % tk) * TK, j0 = (blockIdx.x / tk) * TJ;
  const int tx = threadIdx.x % TXN, ty = threadIdx.x / TXN;
  int64_t s[2], e[2];
  if (wr.rng) {
    s[0] = wr.rng[4 * b];
    e[0] = wr.rng[4 * b + 1];
    s[1] = wr.rng[4 * b + 2];
    e[1] = wr.rng[4 * b + 3];
  } else {
    s[0] = wr.a0 + (int64_t)b * wr.as;
    e[0] = s[0] + wr.na;
    s[1] = wr.c0 + (int64_t)b * wr.cs;
    e[1] = s[1] + wr.nc;
  }
  if (S > 1) {  // this slice's part of the concatenated ranges
    const int64_t n0 = e[0] - s[0], n = n0 + (e[1] - s[1]);
    const int64_t f0 = n * sl / S, f1 = n * (sl + 1) / S;
    const int64_t a = s[0], c = s[1];
    s[0] = a + (f0 < n0 ? f0 : n0);
    e[0] = a + (f1 < n0 ? f1 : n0);
    s[1] = c + (f0 > n0 ? f0 - n0 : 0);
    e[1] = c + (f1 > n0 ? f1 - n0 : 0);
  }
  float acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.f;
  for (int pt = 0; pt < 2; ++pt) {
    for (int64_t r0 = s[pt]; r0 < e[pt]; r0 += 16) {
      for (int t = threadIdx.x; t < 16 * TK; t += 256) {
        const int rr = t / TK, cc = t - rr * TK;
        const int64_t row = r0 + rr;
        bool ok = row < e[pt];
        if (ok && rowmask) ok = rowmask[row] == want;
        const int k = k0 + cc;
        xs[rr][cc] = ok && k < K ? (X ? X[row * ldx + k] : 1.f) : 0.f;
      }
      for (int t = threadIdx.x; t < 16 * TJ; t += 256) {
        const int rr = t / TJ, cc = t - rr * TJ;
        const int64_t row = r0 + rr;
        bool ok = row < e[pt];
        if (ok && rowmask) ok = rowmask[row] == want;
        const int j = j0 + cc;
        ys[rr][cc] = ok && j < J ? Y[row * ldy + j] : 0.f;
      }
      __syncthreads();
#pragma unroll 4
      for (int rr = 0; rr < 16; ++rr) {
        float xv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xs[rr][ty * 4 + u];
#pragma unroll
        for (int v = 0; v < 4; ++v) yv[v] = ys[rr][tx + v * TXN];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] += xv[u] * yv[v];
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + ty * 4 + u, j = j0 + tx + v * TXN;
      if (k < K && j < J) {
        if (S > 1) {
          part[((int64_t)sl * gridDim.y + b) * K * J + (int64_t)k * J + j] = acc[u][v];
        } else {
          float* o = out + (int64_t)b * so + (int64_t)k * J + j;
          *o = accumulate ? *o + acc[u][v] : acc[u][v];
        }
      }
    }
}

// Column sums per walker over fixed row ranges (the bias gradients of the dense layers):
// out[b][j] = sum over the rows of walker b (first range, then second, in order) of Y[row][j].
__global__ void k_wcolsum(const float* __restrict__ Y, int64_t ldy, int J, WRows wr, float* __restrict__ out,
                          int64_t so) {
  const int b = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  float acc = 0.f;
  const int64_t a0 = wr.a0 + (int64_t)b * wr.as, c0 = wr.c0 + (int64_t)b * wr.cs;
  for (int r = 0; r < wr.na; ++r) acc += Y[(a0 + r) * ldy + j];
  for (int r = 0; r < wr.nc; ++r) acc += Y[(c0 + r) * ldy + j];
  out[(int64_t)b * so + j] = acc;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Dense per-walker products on the FP32 matrix cores (X non-null, fixed row ranges, no row
// mask): out[b][k][j] = sum over the walker's rows of X[row][k] Y[row][j].  Block = (128 x 128
// output tile, walker, slice), 16 rows of X[:, k0 .. +127] and Y[:, j0 .. +127] per LDS stage;
// wavefront w computes a 64 x 64 quarter as 2 x 2 tiles of v_mfma_f32_32x32x2f32, two rows per
// step: lane l supplies X[row 2s + l/32][k0 + 32 a + l%32] as A and Y[row 2s + l/32][j0 + 32 b +
// l%32] as B, so the contraction index is the row.  Slices and partial sums as k_wgrad.
__global__ void __launch_bounds__(256) k_wgrad_mfma(const float* __restrict__ X, int64_t ldx,
                                                    const float* __restrict__ Y, int64_t ldy, int K, int J, WRows wr,
                                                    float* __restrict__ out, int64_t so, float* __restrict__ part) {
  constexpr int PITCH = 128 + 4;
  __shared__ float xs[16][PITCH];
  __shared__ float ys[16][PITCH];
  const int b = blockIdx.y, S = gridDim.z, sl = blockIdx.z;
  const int tk = (K + 127) / 128;
  const int k0 = (blockIdx.x % tk) * 128, j0 = (blockIdx.x / tk) * 128;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
  int64_t s[2], e[2];
  s[0] = wr.a0 + (int64_t)b * wr.as;
  e[0] = s[0] + wr.na;
  s[1] = wr.c0 + (int64_t)b * wr.cs;
  e[1] = s[1] + wr.nc;
  if (S > 1) {
    const int64_t n0 = e[0] - s[0], n = n0 + (e[1] - s[1]);
    const int64_t f0 = n * sl / S, f1 = n * (sl + 1) / S;
    const int64_t a = s[0], c = s[1];
    s[0] = a + (f0 < n0 ? f0 : n0);
    e[0] = a + (f1 < n0 ? f1 : n0);
    s[1] = c + (f0 > n0 ? f0 - n0 : 0);
    e[1] = c + (f1 > n0 ? f1 - n0 : 0);
  }
  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][c][v] = 0.f;
  // chunks of 16 rows over both ranges; chunk q + 1 is loaded into registers while chunk q is
  // multiplied from LDS
  const int64_t nc0 = (e[0] - s[0] + 15) / 16, nch = nc0 + (e[1] - s[1] + 15) / 16;
  float rx[8], ry[8];
  auto load = [&](int64_t q) {
    const int pt = q < nc0 ? 0 : 1;
    const int64_t r0 = s[pt] + 16 * (q < nc0 ? q : q - nc0);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = threadIdx.x + 256 * u;
      const int rr = t >> 7, cc = t & 127;
      const int64_t row = r0 + rr;
      const bool ok = row < e[pt];
      rx[u] = ok && k0 + cc < K ? X[row * ldx + k0 + cc] : 0.f;
      ry[u] = ok && j0 + cc < J ? Y[row * ldy + j0 + cc] : 0.f;
    }
  };
  if (nch > 0) load(0);
  for (int64_t q = 0; q < nch; ++q) {
    {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = threadIdx.x + 256 * u;
        xs[t >> 7][t & 127] = rx[u];
        ys[t >> 7][t & 127] = ry[u];
      }
      __syncthreads();
      if (q + 1 < nch) load(q + 1);
#pragma unroll
      for (int kk = 0; kk < 16; kk += 2) {
        const int rr = kk + (lane >> 5);
        float av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = xs[rr][wm + a * 32 + (lane & 31)];
#pragma unroll
        for (int c = 0; c < 2; ++c) bv[c] = ys[rr][wn + c * 32 + (lane & 31)];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < 2; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[c], acc[a][c], 0, 0, 0);
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int j = j0 + wn + c * 32 + (lane & 31);
      if (j >= J) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int k = k0 + wm + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (k >= K) continue;
        if (S > 1)
          part[((int64_t)sl * gridDim.y + b) * K * J + (int64_t)k * J + j] = acc[a][c][v];
        else
          out[(int64_t)b * so + (int64_t)k * J + j] = acc[a][c][v];
      }
    }
}

// out[b][kj] (+)= sum over slices (in order) of part[(slice B + b) K J + kj]
__global__ void k_wgrad_reduce(const float* __restrict__ part, int S, int B, int64_t KJ, float* __restrict__ out,
                               int64_t so, int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)B * KJ) return;
  const int b = (int)(e / KJ);
  const int64_t kj = e - (int64_t)b * KJ;
  float acc = 0.f;
  for (int sl = 0; sl < S; ++sl) acc += part[((int64_t)sl * B + b) * KJ + kj];
  float* o = out + (int64_t)b * so + kj;
  *o = accumulate ? *o + acc : acc;
}

void launch_wgrad(hipStream_t s, const float* X, int64_t ldx, const float* Y, int64_t ldy, int K, int J, int B,
                  const WRows& wr, float* out, int64_t so, bool accumulate, const uint8_t* rowmask, int want,
                  float* part, int64_t part_cap) {
  if (B == 0 || K == 0 || J == 0) return;
  if (!X && K == 1 && !wr.rng && !rowmask && !accumulate) {  // column sums over fixed ranges
    k_wcolsum<<<dim3((unsigned)((J + 255) / 256), (unsigned)B), 256, 0, s>>>(Y, ldy, J, wr, out, so);
    return;
  }
  if (X && !wr.rng && !rowmask && !accumulate && K >= 32 && J >= 32) {
    // dense product on the matrix cores; rows split in two or more slices when few blocks
    const int tiles = ((K + 127) / 128) * ((J + 127) / 128);
    const int64_t rows = (int64_t)wr.na + wr.nc;
    int S = 1;
    if (part) {
      const int64_t want_blocks = (4096 + (int64_t)tiles * B - 1) / ((int64_t)tiles * B);
      S = (int)std::max<int64_t>(1, std::min<int64_t>({8, rows / 32, want_blocks}));
      while (S > 1 && (int64_t)S * B * K * J > part_cap) --S;
    }
    k_wgrad_mfma<<<dim3((unsigned)tiles, (unsigned)B, (unsigned)S), 256, 0, s>>>(X, ldx, Y, ldy, K, J, wr, out, so,
                                                                                 part);
    if (S > 1) {
      const int64_t n = (int64_t)B * K * J;
      k_wgrad_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, S, B, (int64_t)K * J, out, so, 0);
    }
    return;
  }
  // tile shape: 16 x 256 for narrow products (K <= 16 against J > 64), else 64 x 64
  const int TK = K <= 16 && J > 64 ? 16 : 64, TJ = 4096 / TK;
  const int tiles = ((K + TK - 1) / TK) * ((J + TJ - 1) / TJ);
  // slices: at least about 96 rows each, enough blocks for 8192 in flight, at most 32, and
  // the partial products within the scratch
  const int64_t rows = wr.rng ? wr.avg_rows : (int64_t)wr.na + wr.nc;
  int S = 1;
  if (part) {
    const int64_t want_blocks = (8192 + (int64_t)tiles * B - 1) / ((int64_t)tiles * B);
    S = (int)std::max<int64_t>(1, std::min<int64_t>({32, rows / 96, want_blocks}));
    while (S > 1 && (int64_t)S * B * K * J > part_cap) --S;
  }
  const dim3 grid((unsigned)tiles, (unsigned)B, (unsigned)S);
  const int acc = accumulate ? 1 : 0;
  if (TK == 16)
    k_wgrad<16, 256><<<grid, 256, 0, s>>>(X, ldx, Y, ldy, K, J, wr, rowmask, want, out, so, acc, part);
  else
    k_wgrad<64, 64><<<grid, 256, 0, s>>>(X, ldx, Y, ldy, K, J, wr, rowmask, want, out, so, acc, part);
  if (S > 1) {
    const int64_t n = (int64_t)B * K * J;
    k_wgrad_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, S, B, (int64_t)K * J, out, so, accumulate ? 1 : 0);
  }
}

// ---------------------------------------------------------------------------------------
// element-wise helpers
// ---------------------------------------------------------------------------------------
__global__ void k_silu_bwd(float* __restrict__ dz, const float* __restrict__ dy, int64_t ldy, const float* __restrict__ z,
                           int64_t ldz, int64_t rows, int cols, int64_t ldo) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t r = e / cols;
  const int c = (int)(e - r * cols);
  dz[r * ldo + c] = dy[r * ldy + c] * silu_d(z[r * ldz + c]);
}

void launch_silu_bwd(hipStream_t s, float* dz, int64_t ldo, const float* dy, int64_t ldy, const float* z, int64_t ldz,
                     int64_t rows, int cols) {
  const int64_t n = rows * cols;
  if (n == 0) return;
  k_silu_bwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dz, dy, ldy, z, ldz, rows, cols, ldo);
}

__global__ void k_silu_fwd(float* __restrict__ y, const float* __restrict__ z, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = silu_f(z[e]);
}

void launch_silu_fwd(hipStream_t s, float* y, const float* z, int64_t n) {
  if (n) k_silu_fwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(y, z, n);
}

__global__ void k_axpy(float* __restrict__ a, const float* __restrict__ b, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] += b[e];
}

void launch_axpy(hipStream_t s, float* a, const float* b, int64_t n) {
  if (n) k_axpy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, b, n);
}

__global__ void k_fill_f32(float* __restrict__ a, int64_t n, float v) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = v;
}

void launch_fill_f32(hipStream_t s, float* a, int64_t n, float v) {
  if (n) k_fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, n, v);
}

__global__ void k_mul(float* __restrict__ o, const float* __restrict__ a, const float* __restrict__ b, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) o[e] = a[e] * b[e];
}

void launch_mul(hipStream_t s, float* o, const float* a, const float* b, int64_t n) {
  if (n) k_mul<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(o, a, b, n);
}

__global__ void k_transpose2d(const float* __restrict__ A, int R, int C, float* __restrict__ AT) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)R * C) return;
  const int r = (int)(e / C), c = (int)(e - (int64_t)r * C);
  AT[(int64_t)c * R + r] = A[e];
}

void launch_transpose2d(hipStream_t s, const float* A, int R, int C, float* AT) {
  const int64_t n = (int64_t)R * C;
  if (n) k_transpose2d<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(A, R, C, AT);
}

// ---------------------------------------------------------------------------------------
// envelopes and Slater matrices
// ---------------------------------------------------------------------------------------
// Per group row g: ex[g][d][m][e] = exp(-sp_s[m][d][e] |r - R_m|) and
// env[g][d N + o] = sum_{m,e} pi_s[d][m E + e][o] ex[g][d][m][e] (s = spin of the electron).
__global__ void __launch_bounds__(256) k_gr_env(const float* __restrict__ r, const float* __restrict__ R, int B, int N,
                                               int nu, int M, int D, int E, const float* __restrict__ sp_up,
                                               const float* __restrict__ pi_up, const float* __restrict__ sp_dn,
                                               const float* __restrict__ pi_dn, float* __restrict__ ex,
                                               float* __restrict__ env, float* __restrict__ dist) {
  extern __shared__ float gx[];  // [D][M][E]
  const int64_t g = blockIdx.x;
  int w, i;
  group_walker(g, B, nu, N - nu, w, i);
  const bool up = i < nu;
  const float* sp = up ? sp_up : sp_dn;
  const float* pi = up ? pi_up : pi_dn;
  const float* ri = r + ((int64_t)w * N + i) * 3;
  const int DME = D * M * E;
  for (int t = threadIdx.x; t < DME; t += blockDim.x) {
    const int d = t / (M * E), me = t - d * M * E, m = me / E, e = me - m * E;
    const float dx = ri[0] - R[3 * m], dy = ri[1] - R[3 * m + 1], dz = ri[2] - R[3 * m + 2];
    const float dm = sqrtf(dx * dx + dy * dy + dz * dz);
    const float x = expf(-sp[((int64_t)m * D + d) * E + e] * dm);
    gx[t] = x;
    ex[g * DME + t] = x;
    if (d == 0 && e == 0) dist[g * M + m] = dm;
  }
  if (!env) return;  // the sums run as a contraction elsewhere
  __syncthreads();
  const int ND = D * N;
  for (int c = threadIdx.x; c < ND; c += blockDim.x) {
    const int d = c / N, o = c - d * N;
    const float* p = pi + (int64_t)d * M * E * N + o;
    const float* xg = gx + d * M * E;
    float acc = 0.f;
    for (int me = 0; me < M * E; ++me) acc += p[(int64_t)me * N] * xg[me];
    env[g * ND + c] = acc;
  }
}

void launch_gr_env(hipStream_t s, const float* r, const float* R, int B, int N, int nu, int M, int D, int E,
                   const float* sp_up, const float* pi_up, const float* sp_dn, const float* pi_dn, float* ex,
                   float* env, float* dist) {
  const int64_t G = (int64_t)B * N;
  if (G == 0) return;
  k_gr_env<<<(unsigned)G, 256, sizeof(float) * D * M * E, s>>>(r, R, B, N, nu, M, D, E, sp_up, pi_up, sp_dn, pi_dn,
                                                                ex, env, dist);
}

// A[w][d][i][o] = orb[g][d N + o] env[g][d N + o] in FP64
__global__ void k_gr_phi(const float* __restrict__ orb, const float* __restrict__ env, int B, int N, int nu, int D,
                         double* __restrict__ A) {
  const int64_t g = blockIdx.x;
  int w, i;
  group_walker(g, B, nu, N - nu, w, i);
  const int ND = D * N;
  for (int c = threadIdx.x; c < ND; c += blockDim.x) {
    const int d = c / N, o = c - d * N;
    A[(((int64_t)w * D + d) * N + i) * N + o] = (double)orb[g * ND + c] * (double)env[g * ND + c];
  }
}

void launch_gr_phi(hipStream_t s, const float* orb, const float* env, int B, int N, int nu, int D, double* A) {
  if (B * N) k_gr_phi<<<(unsigned)((int64_t)B * N), 256, 0, s>>>(orb, env, B, N, nu, D, A);
}

// ---------------------------------------------------------------------------------------
// Per walker: determinant weights, Jastrow sums, attention readout forward and backward,
// e-e cusp; writes the readout / scale / cusp parameter gradients and the per-walker
// cotangents the row kernel needs.
// ---------------------------------------------------------------------------------------
__device__ float block_sum_f(float v, float* red) {
  v = warp_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  float t = 0.f;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
  return t;
}

__device__ double block_sum_d(double v, double* red) {
  v = warp_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
  return t;
}

__global__ void __launch_bounds__(256) k_gr_head(GradHead a) {
  extern __shared__ float hs[];
  __shared__ float red[8];
  __shared__ double redd[8];
  __shared__ float sJ[4];
  const int w = blockIdx.x, t = threadIdx.x;
  const int N = a.N, nu = a.nu, nd = N - nu, B = a.B;
  // determinants: weights w_d = s_d exp(L_d - m) / sum
  if (t == 0) {
    double m = -1e300;
    for (int d = 0; d < a.D; ++d) m = fmax(m, a.ldet[w * a.D + d]);
    double sum = 0.0;
    for (int d = 0; d < a.D; ++d) sum += a.dsgn[w * a.D + d] * exp(a.ldet[w * a.D + d] - m);
    for (int d = 0; d < a.D; ++d) a.wdet[w * a.D + d] = a.dsgn[w * a.D + d] * exp(a.ldet[w * a.D + d] - m) / sum;
    a.logpsi[w] = m + log(fabs(sum));
  }
  float J[2] = {0.f, 0.f};
  float x3s[2] = {0.f, 0.f};
  // MLP Jastrow: sum_i silu(z3_i) * scale + N zb
  if (a.has_mlp) {
    for (int k = 0; k < 2; ++k) {
      float v = 0.f;
      for (int i = t; i < N; i += blockDim.x) v += silu_f(a.z3[group_index(w, i, B, nu, nd) * 2 + k]);
      x3s[k] = block_sum_f(v, red);
      J[k] += x3s[k] * a.mlp_scale[k] + (float)N * a.mlp_zb[k];
    }
  }
  // attention: per-spin sums of exp(-q.key/sqrt(Dk)) and of its products with the values
  float* u = hs;                        // [2 HD] normalised values (readout input)
  float* nrm = u + 2 * a.HD;            // [2 H]
  float* buf0 = nrm + 2 * a.H;          // [max(2HD, W)] layer activations
  float* zst = buf0 + a.lwmax;          // [depth + 2][lwmax]: raw value sums, then the readout pre-activations
  float* dbuf = zst + (a.depth + 2) * a.lwmax;  // [lwmax] cotangent
  float* dbuf2 = dbuf + a.lwmax;
  const int HD = a.HD, Dk = a.Dk;
  const float rs = rsqrtf((float)Dk);
  if (a.has_att) {
    if (t < HD) {
      const int h = t / Dk, k = t - h * Dk;
      const float key = a.keys[t];
      for (int sp = 0; sp < 2; ++sp) {
        float nv = 0.f, vv = 0.f;
        const int i0 = sp == 0 ? 0 : nu, i1 = sp == 0 ? nu : N;
        for (int i = i0; i < i1; ++i) {
          const float* row = a.qv + group_index(w, i, B, nu, nd) * 2 * HD;
          float qk = row[t] * key;
          for (int o = Dk >> 1; o > 0; o >>= 1) qk += __shfl_xor(qk, o, Dk);
          const float at = expf(-qk * rs);
          nv += at;
          vv += at * row[HD + t];
        }
        u[sp * HD + t] = vv;  // values sums; normalised below
        if (k == 0) nrm[sp * a.H + h] = nv;
      }
    }
    __syncthreads();
    for (int c = t; c < 2 * HD; c += blockDim.x) {
      const int sp = c / HD, h = (c - sp * HD) / Dk;
      zst[c] = u[c];  // keep the raw value sums in the first zst row while the layers run
      u[c] = u[c] / nrm[sp * a.H + h];
    }
    __syncthreads();
    // readout MLP forward: layers l = 0 .. depth (silu on hidden layers)
    int fan = 2 * HD;
    const float* in = u;
    for (int l = 0; l <= a.depth; ++l) {
      const int nout = l < a.depth ? a.W : 2;
      const float* Wl = a.ro_W[l];
      const float* bl = a.ro_b[l];
      float* zl = zst + (l + 1) * a.lwmax;
      if (nout > 2) {
        for (int j = t; j < nout; j += blockDim.x) {
          float acc = bl[j];
          for (int q = 0; q < fan; ++q) acc += in[q] * Wl[(int64_t)q * nout + j];
          zl[j] = acc;
        }
      } else {
        for (int j = 0; j < 2; ++j) {
          float v = 0.f;
          for (int q = t; q < fan; q += blockDim.x) v += in[q] * Wl[(int64_t)q * 2 + j];
          v = block_sum_f(v, red);
          if (t == 0) zl[j] = v + bl[j];
        }
      }
      __syncthreads();
      if (l < a.depth) {
        for (int j = t; j < nout; j += blockDim.x) buf0[j + (l & 1) * 0] = silu_f(zl[j]);
        __syncthreads();
        // copy the activation next to its pre-activation for the backward pass
        for (int j = t; j < nout; j += blockDim.x) dbuf2[j] = buf0[j];
        __syncthreads();
        in = dbuf2;
        // keep activations of layer l in zst's spare row? recomputed in the backward from zl
      }
      fan = nout;
    }
    __syncthreads();
    for (int k = 0; k < 2; ++k) J[k] += zst[(a.depth + 1) * a.lwmax + k] * a.att_scale[k] + a.att_zb[k];
  }
  // e-e cusp value and its parameter derivatives
  double cusp = 0.0;
  if (a.cusp) {
    double v[6] = {0, 0, 0, 0, 0, 0};  // sum_same q, sum_diff q, d/dalpha_same, d/dalpha_diff
    const float* rw = a.r + (int64_t)w * N * 3;
    const double als = a.cusp_p[0], ald = a.cusp_p[1];
    const int64_t npair = (int64_t)N * (N - 1) / 2;
    for (int64_t p = t; p < npair; p += blockDim.x) {
      // p -> (i, j), i < j
      int i = (int)((2.0 * N - 1 - sqrt((2.0 * N - 1) * (2.0 * N - 1) - 8.0 * (double)p)) / 2);
      while ((int64_t)(i + 1) * (2 * N - i - 2) / 2 <= p) ++i;
      while ((int64_t)i * (2 * N - i - 1) / 2 > p) --i;
      const int j = (int)(p - (int64_t)i * (2 * N - i - 1) / 2) + i + 1;
      const double dx = (double)rw[3 * i] - rw[3 * j], dy = (double)rw[3 * i + 1] - rw[3 * j + 1],
                   dz = (double)rw[3 * i + 2] - rw[3 * j + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const bool same = (i < nu) == (j < nu);
      const double al = same ? als : ald;
      const double q = al * al / (al + d);
      const double dq = (al * al + 2.0 * al * d) / ((al + d) * (al + d));
      v[same ? 0 : 1] += q;
      v[same ? 2 : 3] += dq;
    }
    for (int k = 0; k < 4; ++k) v[k] = block_sum_d(v[k], redd);
    cusp = a.cusp_p[2] * v[0] + a.cusp_p[3] * v[1];
    if (t == 0) {
      float* gc = a.gout + (int64_t)w * a.gstride + a.o_cusp;
      gc[0] = (float)(a.cusp_p[2] * v[2]);
      gc[1] = (float)(a.cusp_p[3] * v[3]);
      gc[2] = (float)v[0];
      gc[3] = (float)v[1];
    }
  }
  // log psi of the Jastrow and its cotangent dJ
  if (t == 0) {
    float dJ0 = a.use_mlp ? 1.f : 0.f, dJ1 = a.use_log ? 1.f / J[1] : 0.f;
    sJ[0] = dJ0;
    sJ[1] = dJ1;
    a.dJ[2 * w] = dJ0;
    a.dJ[2 * w + 1] = dJ1;
    a.logpsi[w] += cusp + (a.use_mlp ? J[0] : 0.f) + (a.use_log ? logf(fabsf(J[1])) : 0.f);
    if (a.has_mlp) {
      float* g = a.gout + (int64_t)w * a.gstride;
      g[a.o_mlp_scale] = x3s[0] * dJ0;
      g[a.o_mlp_scale + 1] = x3s[1] * dJ1;
      g[a.o_mlp_zb] = (float)N * dJ0;
      g[a.o_mlp_zb + 1] = (float)N * dJ1;
    }
  }
  __syncthreads();
  if (!a.has_att) return;
  float* gw = a.gout + (int64_t)w * a.gstride;
  // readout backward: last layer (no silu), then hidden layers with silu
  const float* zlast = zst + (a.depth + 1) * a.lwmax;
  if (t == 0) {
    gw[a.o_att_scale] = zlast[0] * sJ[0];
    gw[a.o_att_scale + 1] = zlast[1] * sJ[1];
    gw[a.o_att_zb] = sJ[0];
    gw[a.o_att_zb + 1] = sJ[1];
  }
  if (t < 2) dbuf[t] = sJ[t] * a.att_scale[t];
  __syncthreads();
  for (int l = a.depth; l >= 0; --l) {
    const int nout = l < a.depth ? a.W : 2;
    const int fan = l == 0 ? 2 * HD : a.W;
    // input activation of layer l
    const float* in;
    if (l == 0) {
      in = u;
    } else {
      const float* zp = zst + l * a.lwmax;
      for (int j = t; j < fan; j += blockDim.x) buf0[j] = silu_f(zp[j]);
      __syncthreads();
      in = buf0;
    }
    // dW[q][j] = in[q] dz[j], db[j] = dz[j]
    float* gW = gw + a.o_ro_W[l];
    for (int64_t e = t; e < (int64_t)fan * nout; e += blockDim.x) {
      const int q = (int)(e / nout), j = (int)(e - (int64_t)q * nout);
      gW[e] = in[q] * dbuf[j];
    }
    for (int j = t; j < nout; j += blockDim.x) gw[a.o_ro_b[l] + j] = dbuf[j];
    // d in[q] = sum_j W[q][j] dz[j]; then through the silu of layer l - 1
    const float* Wl = a.ro_W[l];
    for (int q = t; q < fan; q += blockDim.x) {
      float acc = 0.f;
      for (int j = 0; j < nout; ++j) acc += Wl[(int64_t)q * nout + j] * dbuf[j];
      dbuf2[q] = l > 0 ? acc * silu_d(zst[l * a.lwmax + q]) : acc;
    }
    __syncthreads();
    for (int q = t; q < fan; q += blockDim.x) dbuf[q] = dbuf2[q];
    __syncthreads();
  }
  // dbuf = d u [2 HD]: d values = du / norm; d norm = -sum_k du values / norm^2
  for (int c = t; c < 2 * HD; c += blockDim.x) {
    const int sp = c / HD, h = (c - sp * HD) / Dk;
    const float n = nrm[sp * a.H + h];
    a.dval[(int64_t)w * 2 * HD + c] = dbuf[c] / n;
  }
  if (t < HD) {
    const int h = t / Dk, k = t - h * Dk;
    for (int sp = 0; sp < 2; ++sp) {
      const float n = nrm[sp * a.H + h];
      float v = -dbuf[sp * HD + t] * zst[sp * HD + t] / (n * n);
      for (int o = Dk >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, Dk);
      if (k == 0) a.dnorm[((int64_t)w * 2 + sp) * a.H + h] = v;
    }
  }
}

size_t gr_head_lds(const GradHead& a) {
  return sizeof(float) * ((size_t)2 * a.HD + 2 * a.H + (size_t)a.lwmax * (a.depth + 5));
}

void launch_gr_head(hipStream_t s, const GradHead& a) {
  if (a.B == 0) return;
  k_gr_head<<<a.B, 256, gr_head_lds(a), s>>>(a);
}

// ---------------------------------------------------------------------------------------
// Per group row: cotangents of the raw orbital and envelope rows, of the last MLP-Jastrow
// pre-activations, and of the attention projections (with the per-row register-key term).
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gr_rows(GradRows a) {
  const int64_t g = blockIdx.x;
  const int t = threadIdx.x;
  int w, i;
  group_walker(g, a.B, a.nu, a.N - a.nu, w, i);
  const int N = a.N, ND = a.D * N;
  for (int c = t; c < ND; c += blockDim.x) {
    const int d = c / N, o = c - d * N;
    const double dphi = a.wdet[w * a.D + d] * a.Ainv[(((int64_t)w * a.D + d) * N + o) * N + i];
    a.dorb[g * ND + c] = (float)(dphi * (double)a.env[g * ND + c]);
    a.denv[g * ND + c] = (float)(dphi * (double)a.orb[g * ND + c]);
  }
  if (a.has_mlp && t < 2) {
    const float z = a.z3[g * 2 + t];
    a.dz3[g * 2 + t] = a.dJ[2 * w + t] * a.mlp_scale[t] * silu_d(z);
  }
  if (a.has_att && t < a.HD) {
    const int HD = a.HD, Dk = a.Dk, h = t / Dk;
    const int sp = i < a.nu ? 0 : 1;
    const float rs = rsqrtf((float)Dk);
    const float* row = a.qv + g * 2 * HD;
    const float q = row[t], v = row[HD + t], key = a.keys[t];
    float qk = q * key;
    for (int o = Dk >> 1; o > 0; o >>= 1) qk += __shfl_xor(qk, o, Dk);
    const float at = expf(-qk * rs);
    const float dv = a.dval[(int64_t)w * 2 * HD + sp * HD + t];
    float datt = dv * v;
    for (int o = Dk >> 1; o > 0; o >>= 1) datt += __shfl_xor(datt, o, Dk);
    datt += a.dnorm[((int64_t)w * 2 + sp) * a.H + h];
    const float dqk = datt * at * (-rs);
    a.dqv[g * 2 * HD + t] = dqk * key;
    a.dqv[g * 2 * HD + HD + t] = dv * at;
    a.dkey[g * HD + t] = dqk * q;
  }
}

void launch_gr_rows(hipStream_t s, const GradRows& a, int64_t G) {
  if (G) k_gr_rows<<<(unsigned)G, 256, 0, s>>>(a);
}

// d sp_s[m][d][e] = -sum over the walker's spin-s electrons of dist_m ex[d][m][e] t[d][m][e]
// (t = d env_d pi_d^T); one block per walker and spin, threads over (m, d, e).
__global__ void k_gr_sp(const float* __restrict__ ex, const float* __restrict__ tt, const float* __restrict__ dist,
                        int B, int N, int nu, int M, int D, int E, float* __restrict__ gout, int64_t gstride,
                        int64_t o_up, int64_t o_dn) {
  const int w = blockIdx.x, sp = blockIdx.y;
  const int nd = N - nu;
  const int i0 = sp == 0 ? 0 : nu, i1 = sp == 0 ? nu : N;
  const int DME = D * M * E;
  float* out = gout + (int64_t)w * gstride + (sp == 0 ? o_up : o_dn);
  for (int c = threadIdx.x; c < DME; c += blockDim.x) {
    const int m = c / (D * E), de = c - m * D * E, d = de / E, e = de - d * E;  // packed [M][D][E]
    const int col = (d * M + m) * E + e;                                         // [D][M][E]
    float acc = 0.f;
    for (int i = i0; i < i1; ++i) {
      const int64_t g = group_index(w, i, B, nu, nd);
      acc -= dist[g * M + m] * ex[g * DME + col] * tt[g * DME + col];
    }
    out[c] = acc;
  }
}

void launch_gr_sp(hipStream_t s, const float* ex, const float* tt, const float* dist, int B, int N, int nu, int M,
                  int D, int E, float* gout, int64_t gstride, int64_t o_up, int64_t o_dn) {
  if (B == 0) return;
  const dim3 grid(B, N - nu > 0 ? 2 : 1);
  k_gr_sp<<<grid, 256, 0, s>>>(ex, tt, dist, B, N, nu, M, D, E, gout, gstride, o_up, o_dn);
}

// ---------------------------------------------------------------------------------------
// e-e edges and messages
// ---------------------------------------------------------------------------------------
// Forward of the pair filter per CSR pair (thread per pair): the small quantities the edge
// backward needs, laid out per pair as PairRec (see fire_grad.hpp), and beta / the log-scaled
// features as separate row arrays for the contractions.
__global__ void k_gr_pair_fwd(const float* __restrict__ r, int B, int N, int nu, const int64_t* __restrict__ ptr,
                              const int32_t* __restrict__ nbr, EdgeParams ps, EdgeParams pd, float cut,
                              int64_t P, float* __restrict__ rec, float* __restrict__ beta, float* __restrict__ lsf,
                              uint8_t* __restrict__ same, int32_t* __restrict__ ctr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  // centre group: binary search of p in ptr[0 .. G]
  const int64_t G = (int64_t)B * N;
  int64_t lo = 0, hi = G;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (ptr[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t g = lo;
  int w, i;
  group_walker(g, B, nu, N - nu, w, i);
  const int j = nbr[p];
  const bool sm = (i < nu) == (j < nu);
  same[p] = sm ? 1 : 0;
  ctr[p] = (int32_t)g;
  const EdgeParams& e = sm ? ps : pd;
  const FilterParams& f = e.filt;
  const float* ri = r + ((int64_t)w * N + i) * 3;
  const float* rj = r + ((int64_t)w * N + j) * 3;
  const float df[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  const float d = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
  const float feat[4] = {d, df[0], df[1], df[2]};
  const float ls = log1pf(d) / d;
  float* rc = rec + p * PR_FLOATS;
  for (int c = 0; c < 4; ++c) {
    rc[PR_FEAT + c] = feat[c];
    lsf[p * 4 + c] = feat[c] * ls;
  }
  rc[PR_D] = d;
  float pre[FIRE_W0MAX], y[FIRE_W0MAX];
  for (int a = 0; a < f.w0; ++a) {
    float v = f.b[a];
    for (int c = 0; c < 4; ++c) v += feat[c] * f.K[c * f.w0 + a];
    pre[a] = v;
    y[a] = silu_f(v);
    rc[PR_PRE + a] = v;
  }
  float fc, dfc;
  cutoff_poly(d / cut, fc, dfc);
  float envr[FIRE_W0MAX];
  for (int q = 0; q < f.E; ++q) {
    const float x = d / f.sp[q];
    envr[q] = expf(-x * x) * fc;
    rc[PR_ENV + q] = envr[q];
  }
  for (int k = 0; k < f.w1; ++k) {
    float dv = f.bd[k], ev = 0.f;
    for (int a = 0; a < f.w0; ++a) dv += y[a] * f.Wd[a * f.w1 + k];
    for (int q = 0; q < f.E; ++q) ev += envr[q] * f.We[q * f.w1 + k];
    rc[PR_DIR + k] = dv;
    rc[PR_ENVD + k] = ev;
    beta[p * f.w1 + k] = dv * ev;
  }
}

// d out5 rows of the messages: d m_ct_s[i] = sum over i's pairs of spin class s of du,
// d m_nb_s[j] = sum over the pairs (k -> j) of spin class s (found through j's own list and the
// reverse slots); thread per (group, feature).
__global__ void k_gr_msg_collect(const float* __restrict__ du, const int64_t* __restrict__ ptr,
                                 const int32_t* __restrict__ nbr, const int32_t* __restrict__ rev,
                                 const uint8_t* __restrict__ same, int B, int N, int nu, int F, float* __restrict__ dM5) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t G = (int64_t)B * N;
  if (e >= G * F) return;
  const int64_t g = e / F;
  const int f = (int)(e - g * F);
  int w, i;
  group_walker(g, B, nu, N - nu, w, i);
  float cs = 0.f, cd = 0.f, ns = 0.f, nd_ = 0.f;
  for (int64_t p = ptr[g]; p < ptr[g + 1]; ++p) {
    const float v = du[p * F + f];
    if (same[p])
      cs += v;
    else
      cd += v;
    // the pair (k -> i) in k's list
    const int k = nbr[p];
    const int64_t gk = group_index(w, k, B, nu, N - nu);
    const int64_t q = ptr[gk] + rev[p];
    const float vr = du[q * F + f];
    if (same[q])
      ns += vr;
    else
      nd_ += vr;
  }
  dM5[g * 5 * F + 1 * F + f] = cs;
  dM5[g * 5 * F + 2 * F + f] = cd;
  dM5[g * 5 * F + 3 * F + f] = ns;
  dM5[g * 5 * F + 4 * F + f] = nd_;
}

// ---------------------------------------------------------------------------------------
// ElecInit: e-n edges (rows g * M + m) and the dense stage
// ---------------------------------------------------------------------------------------
__global__ void k_gr_en_fwd(const float* __restrict__ r, const float* __restrict__ R, int B, int N, int nu, int M,
                            float cutoff_1el, EnParams ep, int64_t rows, float* __restrict__ rec,
                            float* __restrict__ beta, float* __restrict__ lsf, uint8_t* __restrict__ mask) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows) return;
  const int64_t g = q / M;
  const int m = (int)(q - g * M);
  int w, i;
  group_walker(g, B, nu, N - nu, w, i);
  const float* ri = r + ((int64_t)w * N + i) * 3;
  const float df[3] = {ri[0] - R[3 * m], ri[1] - R[3 * m + 1], ri[2] - R[3 * m + 2]};
  const float d = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
  // the e-n mask of ElecInit: the float32 distance strictly inside cutoff_1el
  const float dd = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(df[0], df[0]), __fmul_rn(df[1], df[1])), __fmul_rn(df[2], df[2])));
  mask[q] = dd < cutoff_1el ? 1 : 0;
  const float feat[4] = {d, df[0], df[1], df[2]};
  const float ls = log1pf(d) / d;
  float* rc = rec + q * PR_FLOATS;
  for (int c = 0; c < 4; ++c) {
    rc[PR_FEAT + c] = feat[c];
    lsf[q * 4 + c] = feat[c] * ls;
  }
  rc[PR_D] = d;
  const float* K = ep.K + (int64_t)m * 4 * ep.w0;
  const float* bb = ep.b + (int64_t)m * ep.w0;
  const float* sp = ep.sp + (int64_t)m * ep.E;
  float y[FIRE_W0MAX];
  for (int a = 0; a < ep.w0; ++a) {
    float v = bb[a];
    for (int c = 0; c < 4; ++c) v += feat[c] * K[c * ep.w0 + a];
    rc[PR_PRE + a] = v;
    y[a] = silu_f(v);
  }
  float fc, dfc;
  cutoff_poly(d / cutoff_1el, fc, dfc);
  float envr[FIRE_EEMAX];
  for (int e = 0; e < ep.E; ++e) {
    const float x = d / sp[e];
    envr[e] = expf(-x * x) * fc;
  }
  for (int k = 0; k < ep.w1; ++k) {
    float dv = ep.bd[k], ev = 0.f;
    for (int a = 0; a < ep.w0; ++a) dv += y[a] * ep.Wd[a * ep.w1 + k];
    for (int e = 0; e < ep.E; ++e) ev += envr[e] * ep.We[e * ep.w1 + k];
    rc[PR_DIR + k] = dv;
    rc[PR_ENVD + k] = ev;
    beta[q * ep.w1 + k] = dv * ev;
  }
}

// h_en[g][f] = sum_m mask Gamma_gm[f] (e_gm[f] + h_nuc[m][f]) with Gamma = beta WG and
// e = lsf WF (rows g M + m); thread per (group, 4 features), F % 4 == 0.
__global__ void k_gr_en_sum(const float* __restrict__ beta, const float* __restrict__ lsf, EnParams ep,
                            const uint8_t* __restrict__ mask, int64_t G, int M, int F, float* __restrict__ hen) {
  const int F4 = F >> 2;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G * F4) return;
  const int64_t g = e / F4;
  const int f = (int)(e - g * F4) * 4;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int m = 0; m < M; ++m) {
    const int64_t q = g * M + m;
    if (!mask[q]) continue;
    float4 gam = make_float4(0.f, 0.f, 0.f, 0.f), ev = ld4(ep.hnuc + (int64_t)m * F + f);
    for (int k = 0; k < ep.w1; ++k) gam = fma4(beta[q * ep.w1 + k], ld4(ep.WG + k * F + f), gam);
    for (int c = 0; c < 4; ++c) ev = fma4(lsf[q * 4 + c], ld4(ep.WF + c * F + f), ev);
    acc.x += gam.x * ev.x;
    acc.y += gam.y * ev.y;
    acc.z += gam.z * ev.z;
    acc.w += gam.w * ev.w;
  }
  st4(hen + g * F + f, acc);
}

// GLU forward: y = x silu(gate), [x | gate] = Z2 [G, 2F]
__global__ void k_gr_glu_fwd(const float* __restrict__ Z2, int64_t G, int F, float* __restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G * F) return;
  const int64_t g = e / F;
  const int f = (int)(e - g * F);
  y[e] = Z2[g * 2 * F + f] * silu_f(Z2[g * 2 * F + F + f]);
}

// LayerNorm forward (flax: biased variance, eps 1e-6) keeping n = (h - mean) / sqrt(var + eps)
// and inv per row; one block of 256 threads per row (F <= 256).
__global__ void k_gr_ln_fwd(const float* __restrict__ h, int F, const float* __restrict__ gsc, const float* __restrict__ gb,
                            float* __restrict__ y, float* __restrict__ nrm, float* __restrict__ inv) {
  __shared__ float red[8];
  const int64_t g = blockIdx.x;
  const int t = threadIdx.x;
  const float x = t < F ? h[g * F + t] : 0.f;
  const float mean = block_sum_f(x, red) / F;
  const float c = t < F ? x - mean : 0.f;
  const float var = block_sum_f(c * c, red) / F;
  const float iv = rsqrtf(var + 1e-6f);
  if (t < F) {
    nrm[g * F + t] = c * iv;
    y[g * F + t] = c * iv * gsc[t] + gb[t];
  }
  if (t == 0) inv[g] = iv;
}

// LayerNorm backward: dn = dy * scale; dh = inv (dn - mean(dn) - n mean(dn n))
__global__ void k_gr_ln_bwd(const float* __restrict__ dy, const float* __restrict__ nrm, const float* __restrict__ inv,
                            const float* __restrict__ gsc, int F, float* __restrict__ dh) {
  __shared__ float red[8];
  const int64_t g = blockIdx.x;
  const int t = threadIdx.x;
  const float n = t < F ? nrm[g * F + t] : 0.f;
  const float dn = t < F ? dy[g * F + t] * gsc[t] : 0.f;
  const float m1 = block_sum_f(dn, red) / F;
  const float m2 = block_sum_f(dn * n, red) / F;
  if (t < F) dh[g * F + t] = inv[g] * (dn - m1 - n * m2);
}

// GLU backward: y = x silu(gate), [x | gate] = Z2 [G, 2F]; dZ2 = [dy silu(gate) | dy x silu'(gate)]
__global__ void k_gr_glu_bwd(const float* __restrict__ dy, const float* __restrict__ Z2, int64_t G, int F,
                             float* __restrict__ dZ2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G * F) return;
  const int64_t g = e / F;
  const int f = (int)(e - g * F);
  const float x = Z2[g * 2 * F + f], gt = Z2[g * 2 * F + F + f], d = dy[e];
  dZ2[g * 2 * F + f] = d * silu_f(gt);
  dZ2[g * 2 * F + F + f] = d * x * silu_d(gt);
}

// Where the filter-parameter gradients of one e-e edge class go in the gradient row: output o
// in the order [K (4 w0) | b (w0) | sp (w0) | Wd (w0 w1) | bd (w1) | We (w0 w1)] (the partials
// of k_gr_pair_cot), its packed destination dst, and the row fields (xi, yi) it multiplies.
struct PairFiltOut {
  int w0, w1, n_out;
  int64_t oK, ob, osp, oWd, obd, oWe;
};

__device__ __forceinline__ void pair_filt_slot(const PairFiltOut& d, int o, int& xi, int& yi, int64_t& dst) {
  const int w0 = d.w0, w1 = d.w1;
  const int sENV = 4, sPRE = 4 + w0, sSP = 4 + 2 * w0, sY = 4 + 3 * w0, sDIR = 4 + 4 * w0, sENVD = 4 + 4 * w0 + w1;
  if (o < 4 * w0) {
    xi = o / w0;
    yi = sPRE + o % w0;
    dst = d.oK + o;
  } else if ((o -= 4 * w0) < w0) {
    xi = -1;
    yi = sPRE + o;
    dst = d.ob + o;
  } else if ((o -= w0) < w0) {
    xi = -1;
    yi = sSP + o;
    dst = d.osp + o;
  } else if ((o -= w0) < w0 * w1) {
    xi = sY + o / w1;
    yi = sDIR + o % w1;
    dst = d.oWd + o;
  } else if ((o -= w0 * w1) < w1) {
    xi = -1;
    yi = sDIR + o;
    dst = d.obd + o;
  } else {
    o -= w1;
    xi = sENV + o / w1;
    yi = sENVD + o % w1;
    dst = d.oWe + o;
  }
}

// Sum of v[k] over the 64 lanes for every k < W1, spread over the lanes: halving steps over
// lane bits 5, 4, ... each hand half of the remaining values to the partner lane (W1 - 1
// shuffles), plain butterflies finish the lower bits.  Lane l returns the total for
// k = xr_index<W1>(l); the lanes with the low 6 - log2(W1) bits clear hold each k once.
template <int W1>
__device__ __forceinline__ int xr_index(int lane) {
  int k = 0;
#pragma unroll
  for (int st = 0, n = W1; n > 1; ++st, n >>= 1)
    if (lane & (32 >> st)) k += n >> 1;
  return k;
}

template <int W1>
__device__ __forceinline__ float xreduce(float (&v)[W1], int lane) {
  int st = 0;
#pragma unroll
  for (int n = W1; n > 1; n >>= 1, ++st) {
    const int o = 32 >> st, half = n >> 1;
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      const float send = up ? v[j] : v[half + j];
      const float keep = up ? v[half + j] : v[j];
      v[j] = keep + __shfl_xor(send, o, 64);
    }
  }
  float t = v[0];
  for (int o = 32 >> st; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  return t;
}

// Fused e-n edge backward (rows g M + m), one workgroup per (nucleus m, walker b): the four
// wavefronts take the walker's electrons in turn, lane features f = lane + 64 t (F <= 256).
// Gamma = beta WG and e = lsf WF + h_nuc[m] are rebuilt per edge, dGamma = dh_en e and
// de = dh_en Gamma are used in registers only: the WG / WF products accumulate per lane
// (partials per (walker, nucleus) to part, summed over nuclei by k_gr_en_cot_reduce), the
// h_nuc gradient of (walker, m) is complete here, and d beta[q][k] = sum_f dGamma WG[k][f] is a
// wavefront reduction per edge.  The filter backward of the edge follows in the same wavefront
// (d beta is uniform): lane a < w0 takes pre-activation a (Wd, K[m], b[m] gradients), lane
// e < E envelope e (We, sp[m]); Wd / We / bd partials per (walker, nucleus) are summed over
// nuclei by k_gr_en_filt_reduce.  Masked edges (outside cutoff_1el) contribute nothing.
template <int W1>
__global__ void __launch_bounds__(256) k_gr_en_cot(const float* __restrict__ beta, const float* __restrict__ lsf,
                                                   const uint8_t* __restrict__ mask, const float* __restrict__ dhen,
                                                   const float* __restrict__ rec, float cutoff_1el, EnParams ep, int B,
                                                   int N, int nu, int M, int F, float* __restrict__ part,
                                                   float* __restrict__ part2, float* __restrict__ gout, int64_t gstride,
                                                   int64_t o_hn, int64_t o_K, int64_t o_b, int64_t o_sp) {
  __shared__ float red[4][64 * FIRE_W1MAX];
  const int m = blockIdx.x, b = blockIdx.y;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int T = F >> 6, nd = N - nu;
  float wg[W1][4], wf[4][4], hn[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int f = lane + 64 * t;
    const bool ok = t < T;
#pragma unroll
    for (int k = 0; k < W1; ++k) wg[k][t] = ok ? ep.WG[k * F + f] : 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) wf[c][t] = ok ? ep.WF[c * F + f] : 0.f;
    hn[t] = ok ? ep.hnuc[(int64_t)m * F + f] : 0.f;
  }
  // filter backward: lane a < w0 takes pre-activation a, lane e < E envelope e of the edge
  const bool la = lane < ep.w0, le = lane < ep.E;
  float wd[W1], we[W1], aWd[W1], aWe[W1], abd[W1];
#pragma unroll
  for (int k = 0; k < W1; ++k) {
    wd[k] = la ? ep.Wd[lane * W1 + k] : 0.f;
    we[k] = le ? ep.We[lane * W1 + k] : 0.f;
    aWd[k] = aWe[k] = abd[k] = 0.f;
  }
  const float spl = le ? ep.sp[(int64_t)m * ep.E + lane] : 1.f;
  float aK[4] = {0.f, 0.f, 0.f, 0.f}, ab = 0.f, asp = 0.f;
  float aw[W1][4], af[4][4], ah[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int k = 0; k < W1; ++k) aw[k][t] = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) af[c][t] = 0.f;
    ah[t] = 0.f;
  }
  for (int i = wid; i < N; i += 4) {
    const int64_t g = group_index(b, i, B, nu, nd);
    const int64_t q = g * M + m;
    if (!mask[q]) continue;
    float bk[W1], lc[4], db[W1];
#pragma unroll
    for (int k = 0; k < W1; ++k) {
      bk[k] = beta[q * W1 + k];
      db[k] = 0.f;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) lc[c] = lsf[q * 4 + c];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= T) break;
      const float dh = dhen[g * F + lane + 64 * t];
      float gam = 0.f, ev = hn[t];
#pragma unroll
      for (int k = 0; k < W1; ++k) gam += bk[k] * wg[k][t];
#pragma unroll
      for (int c = 0; c < 4; ++c) ev += lc[c] * wf[c][t];
      const float dgam = dh * ev, de = dh * gam;
#pragma unroll
      for (int k = 0; k < W1; ++k) {
        aw[k][t] += bk[k] * dgam;
        db[k] += dgam * wg[k][t];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) af[c][t] += lc[c] * de;
      ah[t] += de;
    }
    const float tot = xreduce<W1>(db, lane);
    // d beta of every k in every lane, then the filter backward of the edge
    const float* rc = rec + q * PR_FLOATS;
    const float d = rc[PR_D];
    float ddir[W1], denvd[W1];
#pragma unroll
    for (int k = 0; k < W1; ++k) {
      const float dbk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tot), k * (64 / W1)));
      ddir[k] = dbk * rc[PR_ENVD + k];
      denvd[k] = dbk * rc[PR_DIR + k];
      abd[k] += ddir[k];
    }
    if (la) {
      const float pre = rc[PR_PRE + lane];
      float dy = 0.f;
#pragma unroll
      for (int k = 0; k < W1; ++k) dy += wd[k] * ddir[k];
      const float y = silu_f(pre), dpre = dy * silu_d(pre);
#pragma unroll
      for (int k = 0; k < W1; ++k) aWd[k] += y * ddir[k];
#pragma unroll
      for (int c = 0; c < 4; ++c) aK[c] += rc[PR_FEAT + c] * dpre;
      ab += dpre;
    }
    if (le) {
      float fc, dfc;
      cutoff_poly(d / cutoff_1el, fc, dfc);
      float den = 0.f;
#pragma unroll
      for (int k = 0; k < W1; ++k) den += we[k] * denvd[k];
      const float x = d / spl, g0 = expf(-x * x);
#pragma unroll
      for (int k = 0; k < W1; ++k) aWe[k] += g0 * fc * denvd[k];
      asp += den * g0 * fc * 2.f * x * x / spl;
    }
  }
  // the four wavefronts' accumulators, one field (a row of F) at a time through LDS
  const int f_out = (threadIdx.x >> 2) + 64 * (threadIdx.x & 3);  // thread -> (lane, t) slot
  const bool f_ok = (threadIdx.x & 3) < T;
  float* pb = part + ((int64_t)b * M + m) * (W1 + 4) * F;
  for (int r = 0; r < W1 + 5; ++r) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float v;
      if (r < W1) {
        v = aw[0][t];
#pragma unroll
        for (int k = 1; k < W1; ++k) v = r == k ? aw[k][t] : v;
      } else if (r < W1 + 4) {
        v = af[0][t];
#pragma unroll
        for (int c = 1; c < 4; ++c) v = r == W1 + c ? af[c][t] : v;
      } else {
        v = ah[t];
      }
      red[wid][lane * 4 + t] = v;
    }
    __syncthreads();
    const float sum = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    if (f_ok) {
      if (r < W1 + 4)
        pb[(int64_t)r * F + f_out] = sum;
      else
        gout[(int64_t)b * gstride + o_hn + (int64_t)m * F + f_out] = sum;
    }
    __syncthreads();
  }
  // filter parameters: Wd [w0][W1], We [E][W1], bd [W1] partials per (walker, nucleus); K[m],
  // b[m], sp[m] complete here
  const int w0 = ep.w0, E = ep.E;
  const int P2 = W1 * (w0 + E + 1);
  float* p2 = part2 + ((int64_t)b * M + m) * P2;
  float* gb = gout + (int64_t)b * gstride;
  for (int r = 0; r < 3; ++r) {  // 0: Wd, 1: We, 2: bd with K / b / sp
#pragma unroll
    for (int k = 0; k < W1; ++k) red[wid][lane * W1 + k] = r == 0 ? aWd[k] : (r == 1 ? aWe[k] : abd[k]);
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * W1; e += blockDim.x) {
      const int ln = e / W1, k = e - ln * W1;
      const float v = red[0][e] + red[1][e] + red[2][e] + red[3][e];
      if (r == 0 && ln < w0) p2[ln * W1 + k] = v;
      if (r == 1 && ln < E) p2[w0 * W1 + ln * W1 + k] = v;
      if (r == 2 && ln == 0) p2[(w0 + E) * W1 + k] = v;
    }
    __syncthreads();
  }
  // per-nucleus filter gradients: 4 + 1 values per pre-activation lane, 1 per envelope lane
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    red[wid][lane] = c < 4 ? aK[c] : (c == 4 ? ab : asp);
    __syncthreads();
    if (threadIdx.x < 64) {
      const int ln = threadIdx.x;
      const float v = red[0][ln] + red[1][ln] + red[2][ln] + red[3][ln];
      if (c < 4 && ln < w0) gb[o_K + (int64_t)m * 4 * w0 + c * w0 + ln] = v;
      if (c == 4 && ln < w0) gb[o_b + (int64_t)m * w0 + ln] = v;
      if (c == 5 && ln < E) gb[o_sp + (int64_t)m * E + ln] = v;
    }
    __syncthreads();
  }
}

// gout[b][Wd | We | bd] = sum over nuclei (in order) of the filter partials of k_gr_en_cot
__global__ void k_gr_en_filt_reduce(const float* __restrict__ part2, int B, int M, int P2, int n_wd, int n_we,
                                    float* __restrict__ gout, int64_t gstride, int64_t o_wd, int64_t o_we,
                                    int64_t o_bd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)B * P2) return;
  const int b = (int)(e / P2), j = (int)(e - (int64_t)b * P2);
  float acc = 0.f;
  for (int m = 0; m < M; ++m) acc += part2[((int64_t)b * M + m) * P2 + j];
  const int64_t o = j < n_wd ? o_wd + j : (j < n_wd + n_we ? o_we + (j - n_wd) : o_bd + (j - n_wd - n_we));
  gout[(int64_t)b * gstride + o] = acc;
}

// gout[b][WG (r < W1) | WF (r - W1)] = sum over nuclei (in order) of the partials of k_gr_en_cot
__global__ void k_gr_en_cot_reduce(const float* __restrict__ part, int B, int M, int W1, int F, float* __restrict__ gout,
                                   int64_t gstride, int64_t o_wg, int64_t o_wf) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)(W1 + 4) * F;
  if (e >= (int64_t)B * per) return;
  const int b = (int)(e / per);
  const int64_t rf = e - (int64_t)b * per;
  const int r = (int)(rf / F), f = (int)(rf - (int64_t)r * F);
  float acc = 0.f;
  for (int m = 0; m < M; ++m) acc += part[((int64_t)b * M + m) * per + rf];
  gout[(int64_t)b * gstride + (r < W1 ? o_wg + (int64_t)r * F : o_wf + (int64_t)(r - W1) * F) + f] = acc;
}

void launch_gr_en_cot(hipStream_t s, const float* beta, const float* lsf, const uint8_t* mask, const float* dhen,
                      const float* rec, float cutoff_1el, const EnParams& ep, int B, int N, int nu, int M, int F,
                      float* part, float* part2, float* gout, int64_t gstride, int64_t o_wg, int64_t o_wf,
                      int64_t o_hn, int64_t o_K, int64_t o_b, int64_t o_sp, int64_t o_wd, int64_t o_bd, int64_t o_we) {
  if (B == 0 || M == 0) return;
  const dim3 grid((unsigned)M, (unsigned)B);
  switch (ep.w1) {
    case 4: k_gr_en_cot<4><<<grid, 256, 0, s>>>(beta, lsf, mask, dhen, rec, cutoff_1el, ep, B, N, nu, M, F, part, part2, gout, gstride, o_hn, o_K, o_b, o_sp); break;
    case 8: k_gr_en_cot<8><<<grid, 256, 0, s>>>(beta, lsf, mask, dhen, rec, cutoff_1el, ep, B, N, nu, M, F, part, part2, gout, gstride, o_hn, o_K, o_b, o_sp); break;
    default: k_gr_en_cot<16><<<grid, 256, 0, s>>>(beta, lsf, mask, dhen, rec, cutoff_1el, ep, B, N, nu, M, F, part, part2, gout, gstride, o_hn, o_K, o_b, o_sp); break;
  }
  const int64_t n = (int64_t)B * (ep.w1 + 4) * F;
  k_gr_en_cot_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, B, M, ep.w1, F, gout, gstride, o_wg, o_wf);
  const int P2 = ep.w1 * (ep.w0 + ep.E + 1);
  const int64_t n2 = (int64_t)B * P2;
  k_gr_en_filt_reduce<<<(unsigned)((n2 + 255) / 256), 256, 0, s>>>(part2, B, M, P2, ep.w0 * ep.w1, ep.E * ep.w1, gout,
                                                                   gstride, o_wd, o_we, o_bd);
}

// Fused e-e edge backward for one edge class (want = 1: same spin, 0: opposite spin), one
// wavefront per centre electron (four per workgroup), lane features f = lane + 64 t.  For each of
// the centre's pairs of the class: Gamma = beta WG, e = lsf WF + BF, u = e + m_ct[i] + m_nb[j]
// rebuilt, dGamma = dh_i u sigma(u) and du = dh_i Gamma silu'(u) in registers (du is also written:
// the message collection needs it for the neighbour side), WG / WF / BF products accumulated,
// d beta reduced across the wavefront, then the pair filter's backward lane-parallel as in
// k_gr_en_cot (K, b, sp, Wd, bd, We and the learned cutoff).  Partials per (walker, workgroup)
// in part: [WG (W1 F) | WF (4 F) | BF (F) | filter outputs (pair_filt_slot order) | d cutoff].
template <int W1>
__global__ void __launch_bounds__(256) k_gr_pair_cot(const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                                                     const uint8_t* __restrict__ same, int want,
                                                     const float* __restrict__ beta, const float* __restrict__ lsf,
                                                     const float* __restrict__ rec, const float* __restrict__ M5,
                                                     const float* __restrict__ dh, EdgeParams ep, float cut, int B,
                                                     int N, int nu, int F, float* __restrict__ du, PairFiltOut d,
                                                     float* __restrict__ part) {
  __shared__ float red[4][64 * FIRE_W1MAX];
  const int b = blockIdx.y, blk = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int i = blk * 4 + wid;
  const int T = F >> 6, nd = N - nu;
  const FilterParams& fp = ep.filt;
  const int w0 = fp.w0;
  float wg[W1][4], wf[4][4], bf[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int f = lane + 64 * t;
    const bool ok = t < T;
#pragma unroll
    for (int k = 0; k < W1; ++k) wg[k][t] = ok ? ep.WG[k * F + f] : 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) wf[c][t] = ok ? ep.WF[c * F + f] : 0.f;
    bf[t] = ok ? ep.BF[f] : 0.f;
  }
  const bool la = lane < w0;  // pre-activation a = lane; envelope q = lane (E = w0 for e-e edges)
  float wd[W1], we[W1], aWd[W1], aWe[W1], abd[W1];
#pragma unroll
  for (int k = 0; k < W1; ++k) {
    wd[k] = la ? fp.Wd[lane * W1 + k] : 0.f;
    we[k] = la ? fp.We[lane * W1 + k] : 0.f;
    aWd[k] = aWe[k] = abd[k] = 0.f;
  }
  const float spl = la ? fp.sp[lane] : 1.f;
  float aK[4] = {0.f, 0.f, 0.f, 0.f}, abb = 0.f, asp = 0.f, adc = 0.f;
  float aw[W1][4], af[4][4], ab[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int k = 0; k < W1; ++k) aw[k][t] = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) af[c][t] = 0.f;
    ab[t] = 0.f;
  }
  if (i < N) {
    const int64_t g = group_index(b, i, B, nu, nd);
    float dhv[4], mct[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool ok = t < T;
      const int f = lane + 64 * t;
      dhv[t] = ok ? dh[g * F + f] : 0.f;
      mct[t] = ok ? M5[g * 5 * F + (want ? 1 : 2) * F + f] : 0.f;
    }
    const int64_t p1 = ptr[g + 1];
    for (int64_t p = ptr[g]; p < p1; ++p) {
      if ((same[p] != 0) != (want != 0)) continue;
      const int64_t gj = group_index(b, nbr[p], B, nu, nd);
      float bk[W1], lc[4], db[W1];
#pragma unroll
      for (int k = 0; k < W1; ++k) {
        bk[k] = beta[p * W1 + k];
        db[k] = 0.f;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) lc[c] = lsf[p * 4 + c];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= T) break;
        const int f = lane + 64 * t;
        const float mnb = M5[gj * 5 * F + (want ? 3 : 4) * F + f];
        float gam = 0.f, ev = bf[t];
#pragma unroll
        for (int k = 0; k < W1; ++k) gam += bk[k] * wg[k][t];
#pragma unroll
        for (int c = 0; c < 4; ++c) ev += lc[c] * wf[c][t];
        const float u = ev + mct[t] + mnb;
        const float sg = sigm(u);
        const float dgam = dhv[t] * u * sg;
        const float duv = dhv[t] * gam * sg * (1.0f + u * (1.0f - sg));
        du[p * F + f] = duv;
#pragma unroll
        for (int k = 0; k < W1; ++k) {
          aw[k][t] += bk[k] * dgam;
          db[k] += dgam * wg[k][t];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) af[c][t] += lc[c] * duv;
        ab[t] += duv;
      }
      const float tot = xreduce<W1>(db, lane);
      const float* rc = rec + p * PR_FLOATS;
      const float dd = rc[PR_D];
      float ddir[W1], denvd[W1];
#pragma unroll
      for (int k = 0; k < W1; ++k) {
        const float dbk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tot), k * (64 / W1)));
        ddir[k] = dbk * rc[PR_ENVD + k];
        denvd[k] = dbk * rc[PR_DIR + k];
        abd[k] += ddir[k];
      }
      if (la) {
        const float pre = rc[PR_PRE + lane];
        float dy = 0.f, den = 0.f;
#pragma unroll
        for (int k = 0; k < W1; ++k) {
          dy += wd[k] * ddir[k];
          den += we[k] * denvd[k];
        }
        const float y = silu_f(pre), dpre = dy * silu_d(pre);
        const float envr = rc[PR_ENV + lane];
#pragma unroll
        for (int k = 0; k < W1; ++k) {
          aWd[k] += y * ddir[k];
          aWe[k] += envr * denvd[k];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) aK[c] += rc[PR_FEAT + c] * dpre;
        abb += dpre;
        float fc, dfc;
        cutoff_poly(dd / cut, fc, dfc);
        const float x = dd / spl, g0 = expf(-x * x);
        asp += den * g0 * fc * 2.f * x * x / spl;
        adc += den * g0 * dfc * (-dd / (cut * cut));
      }
    }
  }
  // combine the four wavefronts into this workgroup's partials
  const int64_t PL = (int64_t)(W1 + 5) * F + d.n_out + 1;
  float* pb = part + ((int64_t)b * gridDim.x + blk) * PL;
  const int f_out = (threadIdx.x >> 2) + 64 * (threadIdx.x & 3);
  const bool f_ok = (threadIdx.x & 3) < T;
  for (int r = 0; r < W1 + 5; ++r) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float v;
      if (r < W1) {
        v = aw[0][t];
#pragma unroll
        for (int k = 1; k < W1; ++k) v = r == k ? aw[k][t] : v;
      } else if (r < W1 + 4) {
        v = af[0][t];
#pragma unroll
        for (int c = 1; c < 4; ++c) v = r == W1 + c ? af[c][t] : v;
      } else {
        v = ab[t];
      }
      red[wid][lane * 4 + t] = v;
    }
    __syncthreads();
    if (f_ok) pb[(int64_t)r * F + f_out] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    __syncthreads();
  }
  float* pf = pb + (int64_t)(W1 + 5) * F;  // filter outputs in pair_filt_slot order
  const int w1 = W1;
  for (int r = 0; r < 3; ++r) {  // 0: Wd, 1: We, 2: bd
#pragma unroll
    for (int k = 0; k < W1; ++k) red[wid][lane * W1 + k] = r == 0 ? aWd[k] : (r == 1 ? aWe[k] : abd[k]);
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * W1; e += blockDim.x) {
      const int ln = e / W1, k = e - ln * W1;
      const float v = red[0][e] + red[1][e] + red[2][e] + red[3][e];
      if (r == 0 && ln < w0) pf[6 * w0 + ln * w1 + k] = v;
      if (r == 1 && ln < w0) pf[6 * w0 + w0 * w1 + w1 + ln * w1 + k] = v;
      if (r == 2 && ln == 0) pf[6 * w0 + w0 * w1 + k] = v;
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    red[wid][lane] = c < 4 ? aK[c] : (c == 4 ? abb : (c == 5 ? asp : adc));
    __syncthreads();
    if (threadIdx.x < 64) {
      const int ln = threadIdx.x;
      const float v = red[0][ln] + red[1][ln] + red[2][ln] + red[3][ln];
      if (c < 4 && ln < w0) pf[c * w0 + ln] = v;
      if (c == 4 && ln < w0) pf[4 * w0 + ln] = v;
      if (c == 5 && ln < w0) pf[5 * w0 + ln] = v;
      if (c == 6) red[0][64 + ln] = v;  // per-lane cutoff terms, summed below
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float sdc = 0.f;
    for (int ln = 0; ln < w0; ++ln) sdc += red[0][64 + ln];
    pf[d.n_out] = sdc;
  }
}

// Sums the partials of k_gr_pair_cot over the walker's workgroups (in order) into the gradient
// row: WG, WF, BF and the filter outputs of the class; the learned-cutoff term is written
// (first class) or added (second class).
__global__ void k_gr_pair_cot_reduce(const float* __restrict__ part, int nblk, int B, int W1, int F, PairFiltOut d,
                                     float* __restrict__ gout, int64_t gstride, int64_t oWG, int64_t oWF, int64_t oBF,
                                     int64_t ocut, int add_cut) {
  const int64_t PL = (int64_t)(W1 + 5) * F + d.n_out + 1;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)B * PL) return;
  const int b = (int)(e / PL);
  const int64_t x = e - (int64_t)b * PL;
  float acc = 0.f;
  for (int k = 0; k < nblk; ++k) acc += part[((int64_t)b * nblk + k) * PL + x];
  float* g = gout + (int64_t)b * gstride;
  if (x < (int64_t)W1 * F) {
    g[oWG + x] = acc;
  } else if (x < (int64_t)(W1 + 4) * F) {
    g[oWF + x - (int64_t)W1 * F] = acc;
  } else if (x < (int64_t)(W1 + 5) * F) {
    g[oBF + x - (int64_t)(W1 + 4) * F] = acc;
  } else if (x < PL - 1) {
    int xi, yi;
    int64_t dst;
    pair_filt_slot(d, (int)(x - (int64_t)(W1 + 5) * F), xi, yi, dst);
    g[dst] = acc;
  } else {
    g[ocut] = add_cut ? g[ocut] + acc : acc;
  }
}

void launch_gr_pair_cot(hipStream_t s, const int64_t* ptr, const int32_t* nbr, const uint8_t* same, int want,
                        const float* beta, const float* lsf, const float* rec, const float* M5, const float* dh,
                        const EdgeParams& ep, float cut, int B, int N, int nu, int F, float* du, int64_t oK, int64_t ob,
                        int64_t osp, int64_t oWd, int64_t obd, int64_t oWe, int64_t oWG, int64_t oWF, int64_t oBF,
                        int64_t ocut, bool add_cut, float* part, float* gout, int64_t gstride) {
  if (B == 0 || N == 0) return;
  const int W1 = ep.filt.w1, w0 = ep.filt.w0;
  PairFiltOut d{w0, W1, 6 * w0 + 2 * w0 * W1 + W1, oK, ob, osp, oWd, obd, oWe};
  const int nblk = (N + 3) / 4;
  const dim3 grid((unsigned)nblk, (unsigned)B);
  switch (W1) {
    case 4: k_gr_pair_cot<4><<<grid, 256, 0, s>>>(ptr, nbr, same, want, beta, lsf, rec, M5, dh, ep, cut, B, N, nu, F, du, d, part); break;
    case 8: k_gr_pair_cot<8><<<grid, 256, 0, s>>>(ptr, nbr, same, want, beta, lsf, rec, M5, dh, ep, cut, B, N, nu, F, du, d, part); break;
    default: k_gr_pair_cot<16><<<grid, 256, 0, s>>>(ptr, nbr, same, want, beta, lsf, rec, M5, dh, ep, cut, B, N, nu, F, du, d, part); break;
  }
  const int64_t PL = (int64_t)(W1 + 5) * F + d.n_out + 1;
  const int64_t n = (int64_t)B * PL;
  k_gr_pair_cot_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, nblk, B, W1, F, d, gout, gstride, oWG, oWF, oBF,
                                                                   ocut, add_cut ? 1 : 0);
}

void launch_gr_pair_fwd(hipStream_t s, const float* r, int B, int N, int nu, const int64_t* ptr, const int32_t* nbr,
                        const EdgeParams& ps, const EdgeParams& pd, float cut, int64_t P, float* rec, float* beta,
                        float* lsf, uint8_t* same, int32_t* ctr) {
  if (P) k_gr_pair_fwd<<<(unsigned)((P + 127) / 128), 128, 0, s>>>(r, B, N, nu, ptr, nbr, ps, pd, cut, P, rec, beta, lsf, same, ctr);
}
void launch_gr_msg_collect(hipStream_t s, const float* du, const int64_t* ptr, const int32_t* nbr, const int32_t* rev,
                           const uint8_t* same, int B, int N, int nu, int F, float* dM5) {
  const int64_t n = (int64_t)B * N * F;
  if (n) k_gr_msg_collect<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(du, ptr, nbr, rev, same, B, N, nu, F, dM5);
}
void launch_gr_en_fwd(hipStream_t s, const float* r, const float* R, int B, int N, int nu, int M, float cutoff_1el,
                      const EnParams& ep, int64_t rows, float* rec, float* beta, float* lsf, uint8_t* mask) {
  if (rows) k_gr_en_fwd<<<(unsigned)((rows + 127) / 128), 128, 0, s>>>(r, R, B, N, nu, M, cutoff_1el, ep, rows, rec, beta, lsf, mask);
}
void launch_gr_en_sum(hipStream_t s, const float* beta, const float* lsf, const EnParams& ep, const uint8_t* mask,
                      int64_t G, int M, int F, float* hen) {
  const int64_t n = G * (F / 4);
  if (n) k_gr_en_sum<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(beta, lsf, ep, mask, G, M, F, hen);
}
void launch_gr_glu_fwd(hipStream_t s, const float* Z2, int64_t G, int F, float* y) {
  const int64_t n = G * F;
  if (n) k_gr_glu_fwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(Z2, G, F, y);
}
void launch_gr_ln_fwd(hipStream_t s, const float* h, int64_t G, int F, const float* gsc, const float* gb, float* y,
                      float* nrm, float* inv) {
  if (G) k_gr_ln_fwd<<<(unsigned)G, 256, 0, s>>>(h, F, gsc, gb, y, nrm, inv);
}
void launch_gr_ln_bwd(hipStream_t s, const float* dy, const float* nrm, const float* inv, const float* gsc, int64_t G,
                      int F, float* dh) {
  if (G) k_gr_ln_bwd<<<(unsigned)G, 256, 0, s>>>(dy, nrm, inv, gsc, F, dh);
}
void launch_gr_glu_bwd(hipStream_t s, const float* dy, const float* Z2, int64_t G, int F, float* dZ2) {
  const int64_t n = G * F;
  if (n) k_gr_glu_bwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dy, Z2, G, F, dZ2);
}

}  // namespace fire
