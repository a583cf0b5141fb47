// FiRE embedding kernels: ElecInit e-n edges + LayerNorm, GLU, activations, e-e messages.
//
// References (sparse_wf): ElecInit new_sparse_model.py:190-215, ElecNucEdge :139-158,
// ElecElecEdges :161-187, contract_with_fwd_lap :95-131, ElementWise.apply_with_fwd_lap
// sparse_fwd_lap.py:184-206, GatedLinearUnit model/utils.py:55-62, flax LayerNorm (eps 1e-6).
//
// Every kernel here works on tangent-row groups (fire_common.hpp).  With `tangents`
// false only the value row is produced (groups of one row): the value path used by
// log|psi| and the MCMC sweep.
#include "fire_common.hpp"
#include "fire_kernels.hpp"

namespace fire {

// max |v| over the 64 lanes of the wavefront (all active), added to *dst (float bits, which
// order as integers for non-negative floats) by lane 0 with atomicMax: the row maxima that the
// FP16 planes of the next contraction are scaled by
__device__ __forceinline__ void wave_max_abs_to(float v, uint32_t* dst) {
  float m = fabsf(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(dst, __float_as_uint(m));
}

#define EN_CHUNK 64

// Forward-Laplacian e-n filter of every (electron, nucleus) pair, one thread per pair:
// ef[g][J][row][0 .. W1) = beta, [W1 .. W1 + 4) = g for the 5 rows; eok[g][J] = nucleus J
// within cutoff_1el of electron g.  Kept out of k_elec_init_en, whose register budget it
// would set.
template <int W1>
__global__ void __launch_bounds__(64) k_en_filter(const float* __restrict__ r, const float* __restrict__ R, int B,
                                                  int N, int n_up, int M, float cutoff_1el, EnParams P,
                                                  float* __restrict__ ef, uint8_t* __restrict__ eok) {
  constexpr int ST = 5 * (W1 + 4);
  const int g = blockIdx.x;
  int w, i;
  group_inv(g, B, n_up, N - n_up, w, i);
  const float* ri = r + ((size_t)w * N + i) * 3;
  const float rx = ri[0], ry = ri[1], rz = ri[2];
  for (int J = threadIdx.x; J < M; J += blockDim.x) {
    const float diff[3] = {__fsub_rn(rx, R[3 * J]), __fsub_rn(ry, R[3 * J + 1]), __fsub_rn(rz, R[3 * J + 2])};
    const float s2 = __fadd_rn(__fadd_rn(__fmul_rn(diff[0], diff[0]), __fmul_rn(diff[1], diff[1])),
                               __fmul_rn(diff[2], diff[2]));
    const bool ok = __fsqrt_rn(s2) < cutoff_1el;
    eok[(size_t)g * M + J] = ok;
    if (!ok) continue;
    FilterParams fp;
    fp.K = P.K + (size_t)J * 4 * P.w0;
    fp.b = P.b + (size_t)J * P.w0;
    fp.sp = P.sp + (size_t)J * P.E;
    fp.Wd = P.Wd;
    fp.bd = P.bd;
    fp.We = P.We;
    fp.w0 = P.w0;
    fp.w1 = P.w1;
    fp.E = P.E;
    float beta[5][W1], gg[5][4];
    pair_filter_fl<W1, 5>(diff, fp, cutoff_1el, beta, gg);
    float* o = ef + ((size_t)g * M + J) * ST;
#pragma unroll
    for (int rr = 0; rr < 5; ++rr) {
#pragma unroll
      for (int k = 0; k < W1; ++k) o[rr * (W1 + 4) + k] = beta[rr][k];
#pragma unroll
      for (int m = 0; m < 4; ++m) o[rr * (W1 + 4) + W1 + m] = gg[rr][m];
    }
  }
}

// One workgroup (blockDim = F) per electron: sum over nuclei within cutoff_1el of
// Gamma_J * e_J with forward-Laplacian rows, then LayerNorm with forward-Laplacian.  With
// tangent rows the filter values come from k_en_filter (ef, eok); value rows compute them here.
template <bool TAN, int W1>
__global__ void __launch_bounds__(256) k_elec_init_en(const float* __restrict__ r, const float* __restrict__ R, int B,
                                                       int N, int n_up, int M, float cutoff_1el, int F, EnParams P,
                                                       float* __restrict__ out, const int2* __restrict__ list,
                                                       const float* __restrict__ ef, const uint8_t* __restrict__ eok) {
  constexpr int NRF = TAN ? 5 : 1;  // forward-Laplacian rows staged per nucleus
  __shared__ float s_beta[EN_CHUNK][NRF][W1];
  __shared__ float s_g[EN_CHUNK][NRF][4];
  __shared__ int s_ok[EN_CHUNK];
  __shared__ float s_red[16 * 8];
  const int g = blockIdx.x;
  int w, i;
  if (list) {
    w = list[g].x;
    i = list[g].y;
  } else {
    group_inv(g, B, n_up, N - n_up, w, i);
  }
  const float* ri = r + ((size_t)w * N + i) * 3;
  const float rx = ri[0], ry = ri[1], rz = ri[2];
  const int f = threadIdx.x;
  float acc[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float wg[W1], wf[4];
#pragma unroll
  for (int k = 0; k < W1; ++k) wg[k] = P.WG[k * F + f];
#pragma unroll
  for (int m = 0; m < 4; ++m) wf[m] = P.WF[m * F + f];

  for (int J0 = 0; J0 < M; J0 += EN_CHUNK) {
    const int J = J0 + threadIdx.x;
    if constexpr (TAN) {
      constexpr int ST = 5 * (W1 + 4);
      const int nvc = min(EN_CHUNK, M - J0);
      const float* src = ef + ((size_t)g * M + J0) * ST;
      for (int e = threadIdx.x; e < nvc * ST; e += blockDim.x) {
        const int k = e / ST, rem = e - k * ST, rr = rem / (W1 + 4), q = rem - rr * (W1 + 4);
        const float v = src[e];
        if (q < W1)
          s_beta[k][rr][q] = v;
        else
          s_g[k][rr][q - W1] = v;
      }
      if (threadIdx.x < EN_CHUNK) s_ok[threadIdx.x] = J < M ? eok[(size_t)g * M + J] : 0;
    } else {
    if (threadIdx.x < EN_CHUNK) s_ok[threadIdx.x] = 0;
    if (threadIdx.x < EN_CHUNK && J < M) {
      const float diff[3] = {__fsub_rn(rx, R[3 * J]), __fsub_rn(ry, R[3 * J + 1]), __fsub_rn(rz, R[3 * J + 2])};
      const float s2 = __fadd_rn(__fadd_rn(__fmul_rn(diff[0], diff[0]), __fmul_rn(diff[1], diff[1])),
                                 __fmul_rn(diff[2], diff[2]));
      if (__fsqrt_rn(s2) < cutoff_1el) {
        FilterParams fp;
        fp.K = P.K + (size_t)J * 4 * P.w0;
        fp.b = P.b + (size_t)J * P.w0;
        fp.sp = P.sp + (size_t)J * P.E;
        fp.Wd = P.Wd;
        fp.bd = P.bd;
        fp.We = P.We;
        fp.w0 = P.w0;
        fp.w1 = P.w1;
        fp.E = P.E;
        float beta[NRF][W1], gg[NRF][4];
        pair_filter_fl<W1, NRF>(diff, fp, cutoff_1el, beta, gg);
        const int slot = threadIdx.x;
        s_ok[slot] = 1;
#pragma unroll
        for (int rr = 0; rr < NRF; ++rr) {
#pragma unroll
          for (int k = 0; k < W1; ++k) s_beta[slot][rr][k] = beta[rr][k];
#pragma unroll
          for (int m = 0; m < 4; ++m) s_g[slot][rr][m] = gg[rr][m];
        }
      }
    }
    }
    __syncthreads();
    // nuclei in ascending index order (the order of the reference's e-n list)
    const int nv = min(EN_CHUNK, M - J0);
    if (f < F) {
      for (int t = 0; t < nv; ++t) {
        if (!s_ok[t]) continue;
        const int bestJ = J0 + t;
        float G5[5], E5[5];
#pragma unroll
        for (int rr = 0; rr < NRF; ++rr) {
          float a = 0.f;
#pragma unroll
          for (int k = 0; k < W1; ++k) a += s_beta[t][rr][k] * wg[k];
          G5[rr] = a;
          float b = 0.f;
#pragma unroll
          for (int m = 0; m < 4; ++m) b += s_g[t][rr][m] * wf[m];
          E5[rr] = b;
        }
        E5[0] += P.hnuc[(size_t)bestJ * F + f];
        acc[0] += G5[0] * E5[0];
        if (TAN) {
          acc[1] += G5[1] * E5[0] + G5[0] * E5[1] + 2.f * (G5[2] * E5[2] + G5[3] * E5[3] + G5[4] * E5[4]);
#pragma unroll
          for (int t3 = 0; t3 < 3; ++t3) acc[2 + t3] += G5[2 + t3] * E5[0] + G5[0] * E5[2 + t3];
        }
      }
    }
    __syncthreads();
  }
  // LayerNorm with forward-Laplacian (mean/variance over the F features)
  const float invF = 1.0f / (float)F;
  float m5[5] = {acc[0], acc[1], acc[2], acc[3], acc[4]};
  block_sum<5>(m5, s_red);
  const float c = acc[0] - m5[0] * invF;
  const float cl = acc[1] - m5[1] * invF;
  const float ct[3] = {acc[2] - m5[2] * invF, acc[3] - m5[3] * invF, acc[4] - m5[4] * invF};
  float v6[6] = {c * c, c * ct[0], c * ct[1], c * ct[2], ct[0] * ct[0] + ct[1] * ct[1] + ct[2] * ct[2], c * cl};
  block_sum<6>(v6, s_red);
  const float var = v6[0] * invF;
  const float vt[3] = {2.f * v6[1] * invF, 2.f * v6[2] * invF, 2.f * v6[3] * invF};
  const float vl = 2.f * (v6[4] + v6[5]) * invF;
  const float ve = var + 1e-6f;
  const float sq = sqrtf(ve);
  const float s = 1.0f / sq;
  const float s32 = s / ve;        // ve^-3/2
  const float s52 = s32 / ve;      // ve^-5/2
  const float st[3] = {-0.5f * s32 * vt[0], -0.5f * s32 * vt[1], -0.5f * s32 * vt[2]};
  const float sl = 0.75f * s52 * (vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]) - 0.5f * s32 * vl;
  const float gam = P.ln_g[f], bet = P.ln_b[f];
  if (TAN) {
    float* o = out + (size_t)g * 5 * F + f;
    o[0] = gam * c * s + bet;
    o[F] = gam * (cl * s + 2.f * (ct[0] * st[0] + ct[1] * st[1] + ct[2] * st[2]) + c * sl);
    for (int t3 = 0; t3 < 3; ++t3) o[(2 + t3) * F] = gam * (ct[t3] * s + c * st[t3]);
  } else {
    out[(size_t)g * F + f] = gam * c * s + bet;
  }
}

#define W1_DISPATCH(w1, ...)          \
  switch (w1) {                       \
    case 4: { constexpr int W1 = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int W1 = 8; __VA_ARGS__; } break;   \
    default: { constexpr int W1 = 16; __VA_ARGS__; } break; \
  }

void launch_elec_init_en(hipStream_t s, const float* r, const float* R, int B, int N, int n_up, int M, float cutoff_1el,
                         int F, EnParams P, bool tangents, float* out, const int2* list, int n_list, float* ef,
                         uint8_t* eok) {
  const int G = list ? n_list : B * N;
  if (G == 0) return;
  if (tangents) {  // (no list: every electron of the batch)
    W1_DISPATCH(P.w1, {
      k_en_filter<W1><<<G, 64, 0, s>>>(r, R, B, N, n_up, M, cutoff_1el, P, ef, eok);
      k_elec_init_en<true, W1><<<G, F, 0, s>>>(r, R, B, N, n_up, M, cutoff_1el, F, P, out, nullptr, ef, eok);
    });
  } else {
    W1_DISPATCH(P.w1, k_elec_init_en<false, W1><<<G, F, 0, s>>>(r, R, B, N, n_up, M, cutoff_1el, F, P, out, list,
                                                                nullptr, nullptr));
  }
}

// GatedLinearUnit product y = a * silu(gate) on 5-row groups; Z = [a | gate] per row.
template <bool TAN>
__global__ void k_glu(const float* __restrict__ Z, int G, int F, float* __restrict__ out) {
  const int g = blockIdx.x;
  const int f = blockIdx.y * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int nr = TAN ? 5 : 1;
  const float* z = Z + (size_t)g * nr * 2 * F;
  float* o = out + (size_t)g * nr * F;
  const float ax = z[f], gx = z[F + f];
  float sf, s1, s2;
  silu3(gx, sf, s1, s2);
  o[f] = ax * sf;
  if (TAN) {
    const float al = z[2 * F + f], gl = z[3 * F + f];
    float gt[3], at[3], sum2 = 0.f, cross = 0.f;
    for (int t = 0; t < 3; ++t) {
      at[t] = z[(2 + t) * 2 * F + f];
      gt[t] = z[(2 + t) * 2 * F + F + f];
      sum2 += gt[t] * gt[t];
    }
    const float sl = s1 * gl + s2 * sum2;
    for (int t = 0; t < 3; ++t) {
      const float stt = s1 * gt[t];
      o[(2 + t) * F + f] = at[t] * sf + ax * stt;
      cross += at[t] * stt;
    }
    o[F + f] = al * sf + ax * sl + 2.f * cross;
  }
}

void launch_glu(hipStream_t s, const float* Z, int G, int F, bool tangents, float* out) {
  dim3 grid(G, (F + 255) / 256);
  if (tangents)
    k_glu<true><<<grid, 256, 0, s>>>(Z, G, F, out);
  else
    k_glu<false><<<grid, 256, 0, s>>>(Z, G, F, out);
}

// In-place elementwise activation with the forward-Laplacian rule on every group.
__global__ void k_act(float* __restrict__ Y, int64_t ld, int ncols, GroupDesc gd, int act, float sc,
                      const int32_t* __restrict__ dG) {
  const int g = blockIdx.x;
  if (dG && g >= *dG) return;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int64_t r0 = gd.row(g), r1 = gd.row(g + 1);
  float* y = Y + r0 * ld + c;
  const float x = y[0];
  float f, f1, f2;
  if (act == ACT_SILU) {
    silu3(x, f, f1, f2);
  } else {  // exp(-sc * x)
    f = expf(-sc * x);
    f1 = -sc * f;
    f2 = sc * sc * f;
  }
  y[0] = f;
  const int64_t nr = r1 - r0;
  if (nr == 1) return;
  // jacobian rows 8 at a time: all loads before the stores (the sum keeps the row order)
  float ss = 0.f;
  int64_t k = 2;
  for (; k + 8 <= nr; k += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = y[(k + u) * ld];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      ss += v[u] * v[u];
      y[(k + u) * ld] = v[u] * f1;
    }
  }
  for (; k < nr; ++k) {
    const float v = y[k * ld];
    ss += v * v;
    y[k * ld] = v * f1;
  }
  y[ld] = y[ld] * f1 + f2 * ss;
}

// amax[row] = bits of max_k |A[row][k]| (the row maxima the FP16-plane contractions scale by);
// one wavefront per row, float4 loads (K % 4 == 0, lda % 4 == 0).
__global__ void __launch_bounds__(256) k_rowmax(const float* __restrict__ A, int64_t lda, int64_t M, int K,
                                                uint32_t* __restrict__ amax, const int32_t* __restrict__ dcount) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= M || (dcount && row >= *dcount)) return;
  float m = 0.f;
  for (int k = 4 * lane; k < K; k += 256) {
    const float4 v = *reinterpret_cast<const float4*>(A + row * lda + k);
    m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if (lane == 0) amax[row] = __float_as_uint(m);
}

void launch_rowmax(hipStream_t s, const float* A, int64_t lda, int64_t M, int K, uint32_t* amax,
                   const int32_t* dcount) {
  if (M) k_rowmax<<<(unsigned)((M + 3) / 4), 256, 0, s>>>(A, lda, M, K, amax, dcount);
}

void launch_act(hipStream_t s, float* Y, int64_t ld, int ncols, int G, GroupDesc gd, int act, float act_scale,
                const int32_t* dG) {
  if (G == 0) return;
  dim3 grid(G, (ncols + 255) / 256);
  k_act<<<grid, 256, 0, s>>>(Y, ld, ncols, gd, act, act_scale, dG);
}

// Row codes of the fused silu epilogue: -1 value row, -2 Laplacian row, g on the jacobian rows
// of group g.  One wavefront per group.
__global__ void k_rowgrp(int G, GroupDesc gd, int32_t* __restrict__ code) {
  const int g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (g >= G) return;
  const int64_t r0 = gd.row(g), r1 = gd.row(g + 1);
  for (int64_t r = r0 + (threadIdx.x & 63); r < r1; r += 64) code[r] = r == r0 ? -1 : (r == r0 + 1 ? -2 : g);
}

void launch_rowgrp(hipStream_t s, int G, GroupDesc gd, int32_t* rowgrp) {
  if (G > 0) k_rowgrp<<<(G + 3) / 4, 256, 0, s>>>(G, gd, rowgrp);
}

// out[g][:] = A[row(g)][:K] (the value rows, packed)
__global__ void k_gather_rows(const float* __restrict__ A, int64_t lda, int K, int G, GroupDesc gd,
                              float* __restrict__ out) {
  const int g = blockIdx.x;
  const float* src = A + gd.row(g) * lda;
  for (int k = threadIdx.x; k < K; k += blockDim.x) out[(int64_t)g * K + k] = src[k];
}

void launch_gather_rows(hipStream_t s, const float* A, int64_t lda, int K, int G, GroupDesc gd, float* out) {
  if (G > 0) k_gather_rows<<<G, 256, 0, s>>>(A, lda, K, G, gd, out);
}

__global__ void k_silu_grad(const float* __restrict__ z0, float* __restrict__ f1, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float f, d1, d2;
    silu3(z0[i], f, d1, d2);
    f1[i] = d1;
  }
}

void launch_silu_grad(hipStream_t s, const float* z0, float* f1, int64_t n) {
  if (n > 0) k_silu_grad<<<2048, 256, 0, s>>>(z0, f1, n);
}

// Laplacian rows after the fused epilogue: f1 z + f2 * (sum of z^2 over the group's jacobian
// rows), the sum assembled from the slots the epilogue wrote (a slot exists for each 64-row
// block the jacobian rows touch and each lane half (row / 4 odd or even) that holds one of
// them there).
__global__ void k_act_lapfix(float* __restrict__ C, int64_t ldc, int N, GroupDesc gd, const float* __restrict__ z0,
                             const float* __restrict__ part, int ldf, uint32_t* __restrict__ omax) {
  const int g = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= N) return;  // whole wavefronts: N is a multiple of 128 here
  const int64_t r0 = gd.row(g), r1 = gd.row(g + 1);
  if (r1 - r0 < 2) return;
  float* y = C + (r0 + 1) * ldc + c;
  const float zl = *y;
  const float zv = z0[(int64_t)g * ldf + c];
  // which slots exist (uniform over the block): bit 2 (block - first block) + half
  unsigned mask = 0u;
  const int64_t js = r0 + 2, je = r1;
  if (je > js) {
    const int64_t b0 = js >> 6, b1 = (je - 1) >> 6;
    for (int64_t blk = b0; blk <= b1 && blk - b0 < ACT_EPI_SLOTS / 2; ++blk) {
      const int64_t s0 = js > blk * 64 ? js : blk * 64, s1 = je < blk * 64 + 64 ? je : blk * 64 + 64;
      for (int64_t r = s0; r < s1 && r < s0 + 8; ++r) mask |= 1u << (2 * (blk - b0) + ((r >> 2) & 1));
    }
  }
  // all slot loads issued before the sum (in slot order)
  const float* pg = part + (int64_t)g * ACT_EPI_SLOTS * ldf + c;
  float v[ACT_EPI_SLOTS];
#pragma unroll
  for (int q = 0; q < ACT_EPI_SLOTS; ++q) v[q] = (mask >> q) & 1u ? pg[(int64_t)q * ldf] : 0.f;
  float ss = 0.f;
#pragma unroll
  for (int q = 0; q < ACT_EPI_SLOTS; ++q) ss += v[q];
  float f, f1, f2;
  silu3(zv, f, f1, f2);
  const float out = zl * f1 + f2 * ss;
  *y = out;
  if (omax) wave_max_abs_to(out, omax + r0 + 1);
}

void launch_act_lapfix(hipStream_t s, float* C, int64_t ldc, int N, int G, GroupDesc gd, const float* z0,
                       const float* part, int ldf, uint32_t* omax) {
  if (G == 0) return;
  dim3 grid(G, (N + 255) / 256);
  k_act_lapfix<<<grid, 256, 0, s>>>(C, ldc, N, gd, z0, part, ldf, omax);
}

__global__ void k_rowflag(int G, GroupDesc gd, uint8_t* flag) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) flag[gd.row(g)] = 1;
}

__global__ void k_zero_f32(float* p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] = 0.f;
}

void launch_zero_f32(hipStream_t s, float* p, int64_t n) {
  if (n > 0) k_zero_f32<<<2048, 256, 0, s>>>(p, n);
}

__global__ void k_zero_u8(uint8_t* p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] = 0;
}

void launch_rowflag(hipStream_t s, int G, GroupDesc gd, int64_t n_rows, uint8_t* flag) {
  k_zero_u8<<<1024, 256, 0, s>>>(flag, n_rows);
  k_rowflag<<<(G + 255) / 256, 256, 0, s>>>(G, gd, flag);
}

__global__ void k_scale(float* p, int64_t n, float a) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] *= a;
}

void launch_scale(hipStream_t s, float* p, int64_t n, float a) { k_scale<<<1024, 256, 0, s>>>(p, n, a); }

#define MSG_CHUNK 64

// Forward-Laplacian pair filter of every pair (one thread per pair, workgroups per centre):
// pf[p][row][0 .. W1) = beta, pf[p][row][W1 .. W1 + 4) = g for the 5 rows of pair p (CSR
// order).  Kept out of the message kernel, whose register budget it would set.
template <int W1>
__global__ void __launch_bounds__(64) k_pair_filter(const float* __restrict__ r, int B, int N, int n_up,
                                                    const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                                                    EdgeParams Ps, EdgeParams Pd, float cut, float* __restrict__ pf) {
  constexpr int ST = 5 * (W1 + 4);
  const int g = blockIdx.x;
  int w, i;
  group_inv(g, B, n_up, N - n_up, w, i);
  const int64_t p0 = ptr[g];
  const int nb = (int)(ptr[g + 1] - p0);
  const float* rw = r + (size_t)w * N * 3;
  const bool up_i = i < n_up;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    const int j = nbr[p0 + k];
    const float diff[3] = {rw[3 * i] - rw[3 * j], rw[3 * i + 1] - rw[3 * j + 1], rw[3 * i + 2] - rw[3 * j + 2]};
    const bool same = up_i == (j < n_up);
    float beta[5][W1], gg[5][4];
    pair_filter_fl<W1, 5>(diff, same ? Ps.filt : Pd.filt, cut, beta, gg);
    float* o = pf + (p0 + k) * ST;
#pragma unroll
    for (int rr = 0; rr < 5; ++rr) {
#pragma unroll
      for (int q = 0; q < W1; ++q) o[rr * (W1 + 4) + q] = beta[rr][q];
#pragma unroll
      for (int m = 0; m < 4; ++m) o[rr * (W1 + 4) + W1 + m] = gg[rr][m];
    }
  }
}

// Message passing h_i = h0_i + sum_j Gamma_ij * silu(e_ij + m_ct[i] + m_nb[j]) with
// forward-Laplacian rows; one workgroup (blockDim = F) per centre electron.  The pair
// filter quantities of up to MSG_CHUNK neighbours are staged in LDS (with tangent rows: read
// from pf, written by k_pair_filter; value rows: computed here by one thread per pair); each
// thread then owns one feature column.
template <bool TAN, int W1>
__global__ void __launch_bounds__(256) k_messages(const float* __restrict__ r, int B, int N, int n_up,
                                                   const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                                                   const float* __restrict__ M5, EdgeParams Ps, EdgeParams Pd,
                                                   float cut, int F, float* __restrict__ out,
                                                   const int2* __restrict__ list, int kmove,
                                                   const float* __restrict__ M5k, const int32_t* __restrict__ dn,
                                                   const float* __restrict__ pf, uint32_t* __restrict__ omax) {
  // list mode (value rows only): centres are list[g] = (walker, electron); the h0/message
  // rows of electron `kmove` of every walker come from M5k[walker] instead of M5; dn
  // (optional, device) is the number of list entries in use.
  if (list && dn && (int)blockIdx.x >= *dn) return;
  constexpr int NRF = TAN ? 5 : 1;  // forward-Laplacian rows staged per neighbour
  __shared__ float s_beta[MSG_CHUNK][NRF][W1];
  __shared__ float s_g[MSG_CHUNK][NRF][4];
  __shared__ const float* s_src[MSG_CHUNK];
  __shared__ int s_same[MSG_CHUNK];
  __shared__ int s_nup;  // up-spin neighbours in the chunk (they come first)
  const int g = list ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
  const int n_dn = N - n_up;
  int w, i;
  if (list) {
    w = list[g].x;
    i = list[g].y;
  } else {
    group_inv(g, B, n_up, n_dn, w, i);
  }
  const int f = threadIdx.x;
  const int NR = TAN ? 5 : 1;
  const int64_t p0 = ptr[g];
  const int nb = (int)(ptr[g + 1] - p0);
  const int64_t row0 = TAN ? (5 * (int64_t)g + 3 * p0) : g;
  const float* rw = r + (size_t)w * N * 3;
  const bool up_i = i < n_up;
  auto m5_of = [&](int e) -> const float* {
    if (list && e == kmove) return M5k + (size_t)w * NR * 5 * F;
    return M5 + (size_t)group_of(w, e, B, n_up, n_dn) * NR * 5 * F;
  };

  // weights of this feature column for same- and opposite-spin edges
  float wgs[W1], wgd[W1], wfs[4], wfd[4];
#pragma unroll
  for (int k = 0; k < W1; ++k) { wgs[k] = Ps.WG[k * F + f]; wgd[k] = Pd.WG[k * F + f]; }
#pragma unroll
  for (int m = 0; m < 4; ++m) { wfs[m] = Ps.WF[m * F + f]; wfd[m] = Pd.WF[m * F + f]; }
  const float bfs = Ps.BF[f], bfd = Pd.BF[f];

  const float* m5i = m5_of(i);
  float acc[5], mcs[5], mcd[5];
  float gmx = 0.f;  // max |value| this thread wrote (tangent rows: row maxima of the group)
  for (int rr = 0; rr < NR; ++rr) {
    acc[rr] = m5i[rr * 5 * F + 0 * F + f];
    mcs[rr] = m5i[rr * 5 * F + 1 * F + f];
    mcd[rr] = m5i[rr * 5 * F + 2 * F + f];
  }
  for (int c0 = 0; c0 < nb; c0 += MSG_CHUNK) {
    const int cn = min(MSG_CHUNK, nb - c0);
    if constexpr (TAN) {
      constexpr int ST = 5 * (W1 + 4);
      const float* src = pf + (p0 + c0) * ST;
      for (int e = threadIdx.x; e < cn * ST; e += blockDim.x) {
        const int k = e / ST, rem = e - k * ST, rr = rem / (W1 + 4), q = rem - rr * (W1 + 4);
        const float v = src[e];
        if (q < W1)
          s_beta[k][rr][q] = v;
        else
          s_g[k][rr][q - W1] = v;
      }
      if (threadIdx.x < cn) {
        const int j = nbr[p0 + c0 + threadIdx.x];
        s_src[threadIdx.x] = m5_of(j);
        s_same[threadIdx.x] = up_i == (j < n_up);
      }
    } else if (threadIdx.x < cn) {
      const int k = c0 + threadIdx.x;
      const int j = nbr[p0 + k];
      const float diff[3] = {rw[3 * i] - rw[3 * j], rw[3 * i + 1] - rw[3 * j + 1], rw[3 * i + 2] - rw[3 * j + 2]};
      const bool same = up_i == (j < n_up);
      float beta[NRF][W1], gg[NRF][4];
      pair_filter_fl<W1, NRF>(diff, same ? Ps.filt : Pd.filt, cut, beta, gg);
#pragma unroll
      for (int rr = 0; rr < NRF; ++rr) {
#pragma unroll
        for (int q = 0; q < W1; ++q) s_beta[threadIdx.x][rr][q] = beta[rr][q];
#pragma unroll
        for (int m = 0; m < 4; ++m) s_g[threadIdx.x][rr][m] = gg[rr][m];
      }
      s_src[threadIdx.x] = m5_of(j);
      s_same[threadIdx.x] = same;
    }
    if (!TAN && threadIdx.x < 64) {
      const bool upj = threadIdx.x < cn && nbr[p0 + c0 + threadIdx.x] < n_up;
      const uint64_t bal = __ballot(upj);
      if (threadIdx.x == 0) s_nup = __popcll(bal);
    }
    __syncthreads();
    if constexpr (!TAN) {
      // Neighbours ascend in j and up-spin electrons come first, so the chunk is a run of
      // up-spin neighbours followed by a run of down-spin ones; each run uses one weight
      // set.  Message rows are loaded 8 neighbours ahead.  sigmoid from the hardware exp2
      // and reciprocal (a few ulp).
      const int nup = s_nup;
      for (int run = 0; run < 2; ++run) {
        const int lo = run ? nup : 0, hi = run ? cn : nup;
        const bool same = up_i == (run == 0);
        float wg[W1], wf[4];
#pragma unroll
        for (int q = 0; q < W1; ++q) wg[q] = same ? wgs[q] : wgd[q];
#pragma unroll
        for (int m = 0; m < 4; ++m) wf[m] = same ? wfs[m] : wfd[m];
        const float bf = same ? bfs : bfd, mc = same ? mcs[0] : mcd[0];
        const int off = (same ? 3 : 4) * F + f;
        for (int k0 = lo; k0 < hi; k0 += 8) {
          float mn[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) mn[u] = k0 + u < hi ? s_src[k0 + u][off] : 0.f;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int kk = k0 + u;
            if (kk < hi) {
              float g = 0.f, e = 0.f;
#pragma unroll
              for (int q = 0; q < W1; ++q) g += s_beta[kk][0][q] * wg[q];
#pragma unroll
              for (int m = 0; m < 4; ++m) e += s_g[kk][0][m] * wf[m];
              const float ux = e + bf + mc + mn[u];
              const float sg = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * ux));
              acc[0] += g * (ux * sg);
            }
          }
        }
      }
      __syncthreads();
      continue;
    }
    // neighbours in groups of U: the message rows of the whole group are loaded first
    constexpr int U = TAN ? 2 : 8;
    for (int k0 = 0; k0 < cn; k0 += U) {
      float mnu[U][TAN ? 5 : 1];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k0 + u < cn) {
          const float* m5j = s_src[k0 + u] + (s_same[k0 + u] ? 3 : 4) * F + f;
#pragma unroll
          for (int rr = 0; rr < (TAN ? 5 : 1); ++rr) mnu[u][rr] = m5j[rr * 5 * F];
        }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = k0 + u;
        if (kk >= cn) break;
        const bool same = s_same[kk];
        float G5[5], E5[5], mn[5], mc[5];
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) {
          float a = 0.f;
#pragma unroll
          for (int q = 0; q < W1; ++q) a += s_beta[kk][rr][q] * (same ? wgs[q] : wgd[q]);
          G5[rr] = a;
          float b = 0.f;
#pragma unroll
          for (int m = 0; m < 4; ++m) b += s_g[kk][rr][m] * (same ? wfs[m] : wfd[m]);
          E5[rr] = b;
          mn[rr] = mnu[u][rr];
          mc[rr] = same ? mcs[rr] : mcd[rr];
        }
        E5[0] += same ? bfs : bfd;
        const float ux = E5[0] + mc[0] + mn[0];
        float sv, s1, s2;
        silu3(ux, sv, s1, s2);
        acc[0] += G5[0] * sv;
        if (TAN) {
          float uc[3], un[3], q2 = 0.f;
          for (int t = 0; t < 3; ++t) {
            uc[t] = E5[2 + t] + mc[2 + t];
            un[t] = -E5[2 + t] + mn[2 + t];
            q2 += uc[t] * uc[t] + un[t] * un[t];
          }
          const float ul = 2.f * E5[1] + mc[1] + mn[1];
          const float vl = s1 * ul + s2 * q2;
          float cross = 0.f;
          const int64_t jrow = row0 + 5 + 3 * (int64_t)(c0 + kk);
          float* orow = out + jrow * F + f;
          for (int t = 0; t < 3; ++t) {
            const float vc = s1 * uc[t], vn = s1 * un[t];
            acc[2 + t] += G5[2 + t] * sv + G5[0] * vc;
            const float o = -G5[2 + t] * sv + G5[0] * vn;
            store_rows_out<TAN>(orow + t * F, o);
            gmx = fmaxf(gmx, fabsf(o));
            cross += G5[2 + t] * (vc - vn);
          }
          acc[1] += 2.f * G5[1] * sv + G5[0] * vl + 2.f * cross;
        }
      }
    }
    __syncthreads();
  }
  float* o = out + row0 * F + f;
  for (int rr = 0; rr < NR; ++rr) {
    store_rows_out<TAN>(o + rr * F, acc[rr]);
    gmx = fmaxf(gmx, fabsf(acc[rr]));
  }
  if (TAN && omax) {
    // one maximum for all rows of the group (block reduction), written to each of its rows:
    // an upper bound of every row's own maximum, which is all the FP16 scales need
    __shared__ float s_mx[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) gmx = fmaxf(gmx, __shfl_xor(gmx, off, 64));
    if ((threadIdx.x & 63) == 0) s_mx[threadIdx.x >> 6] = gmx;
    __syncthreads();
    float m = s_mx[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmaxf(m, s_mx[q]);
    const int nrow = 5 + 3 * nb;
    for (int rr = threadIdx.x; rr < nrow; rr += blockDim.x) omax[row0 + rr] = __float_as_uint(m);
  }
}

void launch_messages(hipStream_t s, const float* r, int B, int N, int n_up, const int64_t* ptr, const int32_t* nbr,
                     const float* M5, EdgeParams same, EdgeParams diff, float cut, int F, bool tangents, float* out,
                     const int2* list, int n_list, int kmove, const float* M5k, const int32_t* dn, float* pf,
                     uint32_t* omax) {
  const int G = list ? n_list : B * N;
  if (G == 0) return;
  if (tangents) {
    W1_DISPATCH(same.filt.w1, {
      k_pair_filter<W1><<<G, 64, 0, s>>>(r, B, N, n_up, ptr, nbr, same, diff, cut, pf);
      k_messages<true, W1><<<G, F, 0, s>>>(r, B, N, n_up, ptr, nbr, M5, same, diff, cut, F, out, list, kmove, M5k,
                                           dn, pf, omax);
    });
  } else {
    W1_DISPATCH(same.filt.w1, k_messages<false, W1><<<G, F, 0, s>>>(r, B, N, n_up, ptr, nbr, M5, same, diff, cut, F,
                                                                    out, list, kmove, M5k, dn, nullptr, nullptr));
  }
}

// floats per pair of the k_pair_filter output, for the width W1_DISPATCH picks
size_t pair_filter_floats(int w1) { return (size_t)5 * ((w1 == 4 ? 4 : w1 == 8 ? 8 : 16) + 4); }

}  // namespace fire
