// Single-electron low-rank Metropolis steps on resident walker state.
//
// References (sparse_wf): mcmc_steps_low_rank mcmc.py:196-275, proposal_single_electron
// mcmc.py:100-111, MoonLikeWaveFunction.log_psi_low_rank_update wave_function.py:251-287,
// get_low_rank_pair_indices new_sparse_model.py:439-474, NewSparseEmbedding.low_rank_update
// :553-629, Orbitals.low_rank_update orbitals.py:121-146, woodburry_update model/utils.py:216-244,
// Jastrow.low_rank_update jastrow.py:309-361.
//
// Every step moves electron k of every walker.  The proposal is evaluated into scratch
// buffers ("p_*", rows compacted over the affected electrons of all walkers, up-spin rows
// first) and committed to the state only for accepted walkers, so a rejected walker's
// state is never touched.  Determinant work is FP64.
#include <algorithm>

#include "fire_common.hpp"
#include "fire_kernels2.hpp"
#include "fire_mcmc.hpp"

namespace fire {

// Three N(0,1) variates and one U(0,1) variate for (seed, walker, step).
__device__ inline void step_randoms(uint64_t seed, int w, uint64_t step, double n3[3], double& u) {
  const uint2 key = make_uint2((uint32_t)seed ^ (uint32_t)(seed >> 32), (uint32_t)w);
  const uint4 a = philox4x32_10(make_uint4((uint32_t)step, (uint32_t)(step >> 32), 0u, 0u), key);
  const uint4 b = philox4x32_10(make_uint4((uint32_t)step, (uint32_t)(step >> 32), 1u, 0u), key);
  const double two_pi = 6.283185307179586;
  const double r1 = sqrt(-2.0 * log(u01(a.x))), r2 = sqrt(-2.0 * log(u01(a.z)));
  n3[0] = r1 * cos(two_pi * u01(a.y));
  n3[1] = r1 * sin(two_pi * u01(a.y));
  n3[2] = r2 * cos(two_pi * u01(a.w));
  u = u01(b.x);
}

// Proposal: p_r = state positions with electron k moved.  mode 0: Gaussian step of width
// `width` (proposal_single_electron); mode 1: new position given in rk_in [B,3].
__global__ void k_mc_propose(const float* __restrict__ st_r, int B, int N, int k, float width, uint64_t seed,
                             uint64_t step, const int64_t* __restrict__ dstep, int mode,
                             const float* __restrict__ rk_in, float* __restrict__ p_r, float* __restrict__ p_rk,
                             double* __restrict__ p_u) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  float rk[3];
  if (mode == 0) {
    double n3[3], u;
    // step index = device base (advanced by the sweep itself under graph replay) + offset
    step_randoms(seed, w, (dstep ? (uint64_t)*dstep : 0) + step, n3, u);
    for (int c = 0; c < 3; ++c) rk[c] = st_r[((size_t)w * N + k) * 3 + c] + (float)(n3[c] * (double)width);
    p_u[w] = u;
  } else {
    for (int c = 0; c < 3; ++c) rk[c] = rk_in[3 * w + c];
    p_u[w] = 0.5;
  }
  for (int c = 0; c < 3; ++c) {
    p_r[((size_t)w * N + k) * 3 + c] = rk[c];
    p_rk[3 * w + c] = rk[c];
  }
}

void launch_mc_propose(hipStream_t s, const float* st_r, int B, int N, int k, float width, uint64_t seed, uint64_t step,
                       const int64_t* dstep, int mode, const float* rk_in, float* p_r, float* p_rk, double* p_u) {
  k_mc_propose<<<(B + 63) / 64, 64, 0, s>>>(st_r, B, N, k, width, seed, step, dstep, mode, rk_in, p_r, p_rk, p_u);
}

__global__ void k_add_i64(int64_t* p, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] += v;
}
__global__ void k_set_i64(int64_t* p, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}
void launch_add_i64(hipStream_t s, int64_t* p, int64_t v) { k_add_i64<<<1, 64, 0, s>>>(p, v); }
void launch_set_i64(hipStream_t s, int64_t* p, int64_t v) { k_set_i64<<<1, 64, 0, s>>>(p, v); }

__device__ __forceinline__ bool near32(const float* a, const float* b, float cutoff) {
  const float dx = __fsub_rn(a[0], b[0]), dy = __fsub_rn(a[1], b[1]), dz = __fsub_rn(a[2], b[2]);
  const float s = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
  return __fsqrt_rn(s) < cutoff;
}

// Affected electrons of each walker (one wavefront per walker): k itself and every j that
// is a neighbour of k before or after the move (get_low_rank_pair_indices :449-453),
// ascending; counts split by spin.
__global__ void k_mc_affected(const float* __restrict__ st_r, const float* __restrict__ p_r, int B, int N, int n_up,
                              int k, float cutoff, int cap, int32_t* __restrict__ aff, int32_t* __restrict__ cnt_up,
                              int32_t* __restrict__ cnt_dn, int32_t* __restrict__ overflow) {
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (w >= B) return;
  const float* ro = st_r + (size_t)w * N * 3;
  const float* rn = p_r + (size_t)w * N * 3;
  int count = 0, cu = 0;
  for (int j0 = 0; j0 < N; j0 += 64) {
    const int j = j0 + lane;
    bool in = false;
    if (j < N) in = (j == k) || near32(ro + 3 * k, ro + 3 * j, cutoff) || near32(rn + 3 * k, rn + 3 * j, cutoff);
    const uint64_t m = __ballot(in);
    if (in) aff[(size_t)w * N + count + __popcll(m & ((1ull << lane) - 1ull))] = j;
    const uint64_t mu = __ballot(in && j < n_up);
    cu += __popcll(mu);
    count += __popcll(m);
  }
  if (lane == 0) {
    // static capacity: at most `cap` changed electrons per move (an excess is flagged and
    // truncated so that every buffer access stays in bounds; the step is then invalid);
    // overflow[1] keeps the largest count before truncation
    int cd = count - cu;
    if (count > cap) {
      overflow[0] = 1;
      atomicMax(overflow + 1, count);
      cu = min(cu, cap);
      cd = min(cd, cap - cu);
    }
    cnt_up[w] = cu;
    cnt_dn[w] = cd;
  }
}

void launch_mc_affected(hipStream_t s, const float* st_r, const float* p_r, int B, int N, int n_up, int k, float cutoff,
                        int cap, int32_t* aff, int32_t* cnt_up, int32_t* cnt_dn, int32_t* overflow) {
  k_mc_affected<<<(B * 64 + 255) / 256, 256, 0, s>>>(st_r, p_r, B, N, n_up, k, cutoff, cap, aff, cnt_up, cnt_dn,
                                                     overflow);
}

// Exclusive scans of the per-walker up/down counts; tail = {T_up, T_dn, max count, T};
// stats = {sum over steps of T, max over steps of the max count}.
__global__ void __launch_bounds__(1024) k_mc_scan(const int32_t* __restrict__ cu, const int32_t* __restrict__ cd, int B,
                                                  int32_t* __restrict__ off_up, int32_t* __restrict__ off_dn,
                                                  int32_t* __restrict__ tail, int64_t* __restrict__ stats) {
  __shared__ int32_t su[1024], sd[1024], sm[1024];
  const int t = threadIdx.x;
  const int per = (B + 1023) / 1024;
  const int lo = min(B, t * per), hi = min(B, lo + per);
  int32_t a = 0, b = 0, m = 0;
  for (int q = lo; q < hi; ++q) {
    a += cu[q];
    b += cd[q];
    m = max(m, cu[q] + cd[q]);
  }
  su[t] = a;
  sd[t] = b;
  sm[t] = m;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int32_t va = t >= o ? su[t - o] : 0, vb = t >= o ? sd[t - o] : 0, vm = t >= o ? sm[t - o] : 0;
    __syncthreads();
    su[t] += va;
    sd[t] += vb;
    sm[t] = max(sm[t], vm);
    __syncthreads();
  }
  int32_t ra = su[t] - a, rb = sd[t] - b;
  for (int q = lo; q < hi; ++q) {
    off_up[q] = ra;
    off_dn[q] = rb;
    ra += cu[q];
    rb += cd[q];
  }
  if (t == 1023) {
    tail[0] = su[1023];
    tail[1] = sd[1023];
    tail[2] = sm[1023];
    tail[3] = su[1023] + sd[1023];
    if (stats) {  // null for a proposal that is not counted (graph warm-up)
      stats[0] += su[1023] + sd[1023];
      stats[1] = max(stats[1], (int64_t)sm[1023]);
    }
  }
}

void launch_mc_scan(hipStream_t s, const int32_t* cu, const int32_t* cd, int B, int32_t* off_up, int32_t* off_dn,
                    int32_t* tail, int64_t* stats) {
  k_mc_scan<<<1, 1024, 0, s>>>(cu, cd, B, off_up, off_dn, tail, stats);
}

// Compact row list of the affected electrons: up-spin rows of all walkers first
// (walker-major), then down-spin rows.  rowof[w*N + i] = row or -1.
__global__ void k_mc_rowlist(const int32_t* __restrict__ aff, const int32_t* __restrict__ cu,
                             const int32_t* __restrict__ cd, const int32_t* __restrict__ off_up,
                             const int32_t* __restrict__ off_dn, const int32_t* __restrict__ tail, int N, int n_up,
                             int2* __restrict__ list, int32_t* __restrict__ rowof) {
  const int w = blockIdx.x;
  for (int i = threadIdx.x; i < N; i += blockDim.x) rowof[(size_t)w * N + i] = -1;
  __syncthreads();
  const int a = cu[w] + cd[w];
  for (int j = threadIdx.x; j < a; j += blockDim.x) {
    const int i = aff[(size_t)w * N + j];
    const int t = i < n_up ? off_up[w] + j : tail[0] + off_dn[w] + (j - cu[w]);
    list[t] = make_int2(w, i);
    rowof[(size_t)w * N + i] = t;
  }
}

void launch_mc_rowlist(hipStream_t s, const int32_t* aff, const int32_t* cu, const int32_t* cd, const int32_t* off_up,
                       const int32_t* off_dn, const int32_t* tail, int B, int N, int n_up, int2* list, int32_t* rowof) {
  k_mc_rowlist<<<B, 128, 0, s>>>(aff, cu, cd, off_up, off_dn, tail, N, n_up, list, rowof);
}

// (walker, k) list for the moved electron.
__global__ void k_mc_klist(int B, int k, int2* __restrict__ list) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < B) list[w] = make_int2(w, k);
}

void launch_mc_klist(hipStream_t s, int B, int k, int2* list) { k_mc_klist<<<(B + 255) / 256, 256, 0, s>>>(B, k, list); }

// ---------------------------------------------------------------------------------------
// FP64 Woodbury kernels on the matrix cores.
//
// v_mfma_f64_16x16x4f64 computes D = A B + C with A 16x4, B 4x16, C and D 16x16.  Lane l
// holds A[l%16][l/16] and B[l/16][l%16]; register i of lane l holds D[4i + l/16][l%16]
// (measured on gfx950).  At most kWbCap electrons change per move (the state's capacity),
// so every a x a matrix fits four 16-wide tiles.
//
// Per (walker, det) the update reads, in 16x4 operand tiles of 64 doubles (one coalesced
// access per tile):
//   vs[(t * KS + ks) * 64 + l] = V[16t + l%16][4ks + l/16],  V = Phi_new - A[aff] (a x N),
//     written by k_wb_s;
//   xs[(to * NKA + ks) * 64 + l] = X[16to + l%16][4ks + l/16],  X = Ainv[:, aff] S^-1 (N x a),
//     written by k_wb_x;
// with KS = ceil(N / 4), NKA = 4 ceil(a / 16) and zeros outside the matrices.
// ---------------------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma_f64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kWbCap = 64;        // changed electrons per move handled here
constexpr int kSld = kWbCap + 1;  // LDS row stride of S

// value of x in lane `src` (wave-uniform) for every lane
__device__ __forceinline__ double bcast_lane(double x, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
  return __hiloint2double(hi, lo);
}

// wave-wide maximum of a 32-bit key (all 64 lanes active): DPP butterflies within each row
// of 16 lanes, then the four row maxima by readlane; the result is wave-uniform
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));  // row_ror 4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));  // row_ror 8
  const uint32_t r0 = __builtin_amdgcn_readlane((int)v, 0), r1 = __builtin_amdgcn_readlane((int)v, 16);
  const uint32_t r2 = __builtin_amdgcn_readlane((int)v, 32), r3 = __builtin_amdgcn_readlane((int)v, 48);
  return max(max(r0, r1), max(r2, r3));
}

// Woodbury preparation, first kernel, per (walker, det): V = Phi_new - A[aff] of the
// affected electrons and S = I + V P, P = Ainv[:, aff], on the matrix cores.  One wave per
// 16x16 tile (t, u) of S with K over all electrons (the u = 0 waves also write V in operand
// tiles); S goes to `sbuf` (64 x 64 per (walker, det), row-major).
__global__ void __launch_bounds__(256) k_wb_s(const float* __restrict__ p_orb, const float* __restrict__ p_envk,
                                              const float* __restrict__ st_env, const double* __restrict__ st_A,
                                              const double* __restrict__ st_Ainv, const int32_t* __restrict__ aff,
                                              const int32_t* __restrict__ cu, const int32_t* __restrict__ cd,
                                              const int32_t* __restrict__ rowof, int B, int N, int n_up, int D, int k,
                                              double* __restrict__ vs, double* __restrict__ sbuf) {
  constexpr int CH = 8;  // k steps loaded ahead
  const int wd = blockIdx.x;
  const int w = wd / D, d = wd - w * D;
  const int a = cu[w] + cd[w];
  const int nT = (a + 15) >> 4;
  const int ND = N * D, KS = (N + 3) >> 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, g = lane >> 4;
  const double* Ai = st_Ainv + (size_t)wd * N * N;
  const double* Aw = st_A + (size_t)wd * N * N;
  double* VS = vs + (size_t)wd * wb_vs_doubles(N);
  double* SB = sbuf + (size_t)wd * kWbCap * kWbCap;
  const int32_t* af = aff + (size_t)w * N;
  for (int tt = wv; tt < nT * nT; tt += 4) {
    const int t = tt / nT, u = tt - t * nT;
    const int j = 16 * t + m, l = 16 * u + m;
    const bool vj = j < a, vl = l < a;
    const int ij = vj ? af[j] : 0;
    const float* orb = p_orb + (size_t)(vj ? rowof[(size_t)w * N + ij] : 0) * ND + (size_t)d * N;
    const float* env = (ij == k ? p_envk + (size_t)w * ND : st_env + (size_t)group_of(w, ij, B, n_up, N - n_up) * ND) +
                       (size_t)d * N;
    const double* arow = Aw + (size_t)ij * N;
    const int cl = vl ? af[l] : 0;
    dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int ks0 = 0; ks0 < KS; ks0 += CH) {
      float fo[CH], fe[CH];
      double fa[CH], fb[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int o = 4 * (ks0 + q) + g;  // o >= N for k steps past the end
        const bool oj = vj && o < N;
        fo[q] = oj ? orb[o] : 0.0f;
        fe[q] = oj ? env[o] : 0.0f;
        fa[q] = oj ? arow[o] : 0.0;
        fb[q] = (vl && o < N) ? Ai[(size_t)o * N + cl] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int ks = ks0 + q;
        const int o = 4 * ks + g;
        // product rounded in FP32 exactly as the full evaluation forms the Slater matrix
        const double v = (vj && o < N) ? (double)__fmul_rn(fo[q], fe[q]) - fa[q] : 0.0;
        if (u == 0 && ks < KS) VS[((size_t)t * KS + ks) * 64 + lane] = v;
        acc = mfma_f64(v, fb[q], acc);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * t + 4 * i + g;
      if (r < a && vl) SB[r * kWbCap + l] = acc[i] + (r == l ? 1.0 : 0.0);
    }
  }
}

// Woodbury preparation, second kernel: S^-1 (in place in `sbuf`), log|det S| and its sign
// by Gauss-Jordan elimination with partial pivoting, one block of NW wavefronts per
// (walker, det) with S in registers: element e = tid + 64 NW kk of the a x a matrix (row
// e / a, column e % a) is val[kk], its indices packed in idx[kk] = column | row << 8.
// Pivot rows are chosen without swaps: step c picks the unused row p with the largest
// |S[p][c]| (|S| rounded to FP32 with its 7 lowest mantissa bits replaced by the row, so
// the lowest row wins among near-ties; every wave finds the same p), scales it and
// eliminates column c from every other row; the pivot column and the scaled pivot row are
// exchanged through LDS.  The result in place is (P S)^-1 with row c of P S in row piv[c],
// so S^-1[i][piv[c]] = stored[piv[i]][c] and det S = sign(piv) prod(pivots).  NW = 1
// handles a <= 32, NW = 4 the larger a.
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_wb_gj(const int32_t* __restrict__ cu, const int32_t* __restrict__ cd,
                                                   int D, double* __restrict__ sbuf, double* __restrict__ ldS,
                                                   double* __restrict__ sgS) {
  constexpr int KM = 16;  // slots per lane: 64 NW KM >= a^2
  // entry kWbCap of s_f / s_pr is a sink for the lanes whose element is not exchanged, so
  // the passes below run without branches
  __shared__ double s_f[kWbCap + 1];
  __shared__ double s_pr[kWbCap + 1];
  __shared__ int s_piv[kWbCap];
  __shared__ int s_qinv[kWbCap];
  constexpr int kSink = kWbCap | (kWbCap << 8);
  const int wd = blockIdx.x;
  const int w = wd / D;
  const int a = cu[w] + cd[w];
  if (NW == 1 ? a > 32 : a <= 32) return;
  const int a2 = a * a, nk = (a2 + 64 * NW - 1) / (64 * NW);
  const int tid = threadIdx.x, lane = tid & 63;
  const uint32_t ma = ((1u << 20) + a - 1) / a;  // e / a == (e * ma) >> 20 for e < 4096
  double* SB = sbuf + (size_t)wd * kWbCap * kWbCap;
  double val[KM];
  int idx[KM];
#pragma unroll
  for (int kk = 0; kk < KM; ++kk) {
    const int e = tid + 64 * NW * kk;
    const int r = (int)(__umul24((uint32_t)e, ma) >> 20), cc = e - __mul24(r, a);
    const bool in = kk < nk && e < a2;
    idx[kk] = in ? (cc | (r << 8)) : kSink;
    val[kk] = in ? SB[r * kWbCap + cc] : 0.0;
  }
  bool used = false;  // lane r: row r has been a pivot row
  double mant = 1.0;
  int ex2 = 0, nneg = 0;
  for (int c = 0; c < a; ++c) {
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < nk) s_f[(idx[kk] & 255) == c ? idx[kk] >> 8 : kWbCap] = val[kk];
    if (NW > 1) __syncthreads();
    const double fr = lane < a ? s_f[lane] : 0.0;  // S[lane][c]
    const uint32_t fb = __float_as_uint((float)fabs(fr));
    const uint32_t key = (lane < a && !used) ? ((fb & ~127u) | (uint32_t)(127 - lane)) : 0u;
    const int p = 127 - (int)(wave_max_u32(key) & 127u);
    used = used || lane == p;
    if (tid == 0) s_piv[c] = p;
    const double pv = bcast_lane(fr, p);
    nneg += pv < 0.0;
    int ee;
    mant = frexp(mant * fabs(pv), &ee);  // log|det| accumulated as mantissa and exponent
    ex2 += ee;
    const double ip = 1.0 / pv;
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < nk) {
        const int cc = idx[kk] & 255;
        s_pr[(idx[kk] >> 8) == p ? cc : kWbCap] = cc == c ? ip : val[kk] * ip;
      }
    if (NW > 1) __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < nk) {
        const int r = idx[kk] >> 8, cc = idx[kk] & 255;
        const double f = s_f[r], pr = s_pr[cc];
        const double y = cc == c ? -f * ip : val[kk] - f * pr;
        val[kk] = r == p ? pr : y;
      }
    if (NW > 1) __syncthreads();
  }
  int inv = 0;
  if (tid < a) {
    const int pk = s_piv[tid];
    s_qinv[pk] = tid;
    for (int q = tid + 1; q < a; ++q) inv += s_piv[q] < pk;  // inversions of the permutation
  }
  if (NW > 1) __syncthreads();
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) inv += __shfl_xor(inv, off, 64);
#pragma unroll
  for (int kk = 0; kk < KM; ++kk)
    if (kk < nk && idx[kk] != kSink) SB[s_qinv[idx[kk] >> 8] * kWbCap + s_piv[idx[kk] & 255]] = val[kk];
  if (tid == 0) {
    ldS[wd] = log(mant) + (double)ex2 * 0.6931471805599453;
    sgS[wd] = ((inv + nneg) & 1) ? -1.0 : 1.0;
  }
}

void launch_wb_prep(hipStream_t s, const float* p_orb, const float* p_envk, const float* st_env, const double* st_A,
                    const double* st_Ainv, const int32_t* aff, const int32_t* cu, const int32_t* cd,
                    const int32_t* rowof, int B, int N, int n_up, int D, int k, double* vs, double* sbuf, double* ldS,
                    double* sgS) {
  k_wb_s<<<B * D, 256, 0, s>>>(p_orb, p_envk, st_env, st_A, st_Ainv, aff, cu, cd, rowof, B, N, n_up, D, k, vs, sbuf);
  k_wb_gj<1><<<B * D, 64, 0, s>>>(cu, cd, D, sbuf, ldS, sgS);
  if (N > 32) k_wb_gj<4><<<B * D, 256, 0, s>>>(cu, cd, D, sbuf, ldS, sgS);
}

// X^T = S^-T P^T (X = Ainv[:, aff] S^-1) for accepted walkers, in A-operand tiles of the
// update product; the four waves take the 16-row tiles of X in turn.
template <int NTA>
__device__ __forceinline__ void wb_x_body(int a, int N, const int32_t* __restrict__ af, const double* __restrict__ SB,
                                          const double* __restrict__ Ai, double* __restrict__ XS) {
  constexpr int NKA = 4 * NTA;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, g = lane >> 4;
  const int NT = (N + 15) >> 4;
  for (int to = wv; to < NT; to += 4) {
    const int o = 16 * to + m;
    dbl4 acc[NTA];
#pragma unroll
    for (int t = 0; t < NTA; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks0 = 0; ks0 < NKA; ks0 += 4) {
      double bb[4], sv[4][NTA];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kk = 4 * (ks0 + u) + g;
        const bool kv = kk < a;
        bb[u] = (kv && o < N) ? Ai[(size_t)o * N + af[kk]] : 0.0;
#pragma unroll
        for (int t = 0; t < NTA; ++t) sv[u][t] = (kv && 16 * t + m < a) ? SB[kk * kWbCap + 16 * t + m] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int t = 0; t < NTA; ++t) acc[t] = mfma_f64(sv[u][t], bb[u], acc[t]);
    }
    // register i of lane l: X[16to + m][16t + 4i + g], the A operand of k step 4t + i
#pragma unroll
    for (int t = 0; t < NTA; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) XS[((size_t)to * NKA + 4 * t + i) * 64 + lane] = acc[t][i];
  }
}

__global__ void __launch_bounds__(256) k_wb_x(const int32_t* __restrict__ accept, const int32_t* __restrict__ aff,
                                              const int32_t* __restrict__ cu, const int32_t* __restrict__ cd, int N,
                                              int D, const double* __restrict__ st_Ainv,
                                              const double* __restrict__ sbuf, double* __restrict__ xs) {
  const int wd = blockIdx.x;
  const int w = wd / D;
  if (!accept[w]) return;
  const int a = cu[w] + cd[w];
  const double* Ai = st_Ainv + (size_t)wd * N * N;
  const double* SB = sbuf + (size_t)wd * kWbCap * kWbCap;
  double* XS = xs + (size_t)wd * (wb_update_scratch(N) / sizeof(double));
  const int32_t* af = aff + (size_t)w * N;
  switch ((a + 15) >> 4) {
    case 1: wb_x_body<1>(a, N, af, SB, Ai, XS); break;
    case 2: wb_x_body<2>(a, N, af, SB, Ai, XS); break;
    case 3: wb_x_body<3>(a, N, af, SB, Ai, XS); break;
    default: wb_x_body<4>(a, N, af, SB, Ai, XS); break;
  }
}

// Woodbury inverse update for accepted walkers (woodburry_update, utils.py:243):
//   Ainv <- Ainv - X Q,  Q = V Ainv;  then A[aff] <- A[aff] + V,
// with V and X in operand tiles as left by k_wb_s and k_wb_x.
//
// Register-tile form (N <= 16 NT): one 128-thread block per 16-wide column tile tc of Ainv;
// wave h holds the row tiles [h NT0, min(NT, (h + 1) NT0)) of the column tile in registers,
// register kk = 4 tt + i being Ainv[16 (h NT0 + tt) + 4i + l/16][16tc + l%16].  These are
// the B operands of the wave's part of Q[:, tc] = V Ainv[:, tc] and, unchanged, register i
// of the accumulator of output tile (h NT0 + tt, tc).  The two partial Q are added in LDS in
// wave order; the D registers of Q are the B operands of X Q.  Ainv is read once and
// written once and Q never leaves the block.  Blocks are ordered so that the column tiles
// of one (walker, det) run on one XCD (blocks go to XCDs round-robin by index) and share its
// L2 for the V and X tiles.
template <int NT, int NTA>
__device__ __forceinline__ void wb_update_tile(int a, int N, int tc, const int32_t* __restrict__ aff_w,
                                               const double* __restrict__ VS, const double* __restrict__ XS,
                                               double* __restrict__ Ai, double* __restrict__ Aw, dbl4* sq) {
  constexpr int NKA = 4 * NTA, NT0 = (NT + 1) / 2, KH = 4 * NT0;
  constexpr int CH = 4 / NTA > 0 ? 4 / NTA : 1;  // k steps of V per chunk
  const int KS = (N + 3) >> 2;
  const int lane = threadIdx.x & 63, h = threadIdx.x >> 6, m = lane & 15, g = lane >> 4;
  const int t_lo = h * NT0, nt = h ? NT - NT0 : NT0;  // this wave's row tiles
  const int c = 16 * tc + m;
  const bool cv = c < N;
  double mreg[KH];
#pragma unroll
  for (int kk = 0; kk < KH; ++kk) {
    const int row = 16 * t_lo + 4 * kk + g;
    mreg[kk] = (kk < 4 * nt && row < N && cv) ? Ai[(size_t)row * N + c] : 0.0;
  }
  dbl4 q[NTA];
#pragma unroll
  for (int t = 0; t < NTA; ++t) q[t] = dbl4{0.0, 0.0, 0.0, 0.0};
  // partial Q over this wave's rows; V chunks double-buffered
  auto load_v = [&](int kk0, double (&vv)[CH][NTA]) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int ks = 4 * t_lo + kk0 + u;
#pragma unroll
      for (int t = 0; t < NTA; ++t)
        vv[u][t] = (kk0 + u < 4 * nt && ks < KS) ? VS[((size_t)t * KS + ks) * 64 + lane] : 0.0;
    }
  };
  double va[CH][NTA], vb[CH][NTA];
  load_v(0, va);
#pragma unroll
  for (int kk0 = 0; kk0 < KH; kk0 += 2 * CH) {
    if (kk0 + CH < KH) load_v(kk0 + CH, vb);
#pragma unroll
    for (int u = 0; u < CH; ++u)
#pragma unroll
      for (int t = 0; t < NTA; ++t) q[t] = mfma_f64(va[u][t], mreg[kk0 + u], q[t]);
    if (kk0 + CH < KH) {
      if (kk0 + 2 * CH < KH) load_v(kk0 + 2 * CH, va);
#pragma unroll
      for (int u = 0; u < CH; ++u)
#pragma unroll
        for (int t = 0; t < NTA; ++t) q[t] = mfma_f64(vb[u][t], mreg[kk0 + CH + u], q[t]);
    }
  }
  // Q = partial(wave 0) + partial(wave 1)
  if (h == 1)
#pragma unroll
    for (int t = 0; t < NTA; ++t) sq[t * 64 + lane] = q[t];
  __syncthreads();
  if (h == 0)
#pragma unroll
    for (int t = 0; t < NTA; ++t) {
      q[t] += sq[t * 64 + lane];
      sq[t * 64 + lane] = q[t];
    }
  __syncthreads();
  if (h == 1)
#pragma unroll
    for (int t = 0; t < NTA; ++t) q[t] = sq[t * 64 + lane];
  // Ainv[rows of this wave, tc] -= X Q; X operands of the next tile loaded ahead
  auto load_x = [&](int tt, double (&xa)[NKA]) {
    const int to = t_lo + tt;
#pragma unroll
    for (int ks = 0; ks < NKA; ++ks) xa[ks] = tt < nt ? -XS[((size_t)to * NKA + ks) * 64 + lane] : 0.0;
  };
  double xa[NKA], xb[NKA];
  load_x(0, xa);
#pragma unroll
  for (int tt = 0; tt < NT0; ++tt) {
    double(&xc)[NKA] = (tt & 1) ? xb : xa;
    double(&xn)[NKA] = (tt & 1) ? xa : xb;
    if (tt + 1 < NT0) load_x(tt + 1, xn);
    if (tt < nt) {
      dbl4 acc = dbl4{mreg[4 * tt], mreg[4 * tt + 1], mreg[4 * tt + 2], mreg[4 * tt + 3]};
#pragma unroll
      for (int t = 0; t < NTA; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc = mfma_f64(xc[4 * t + i], q[t][i], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * (t_lo + tt) + 4 * i + g;
        if (r < N && cv) Ai[(size_t)r * N + c] = acc[i];
      }
    }
  }
  // A[aff] += V on this tile's columns
  for (int e = threadIdx.x; e < a * 16; e += 128) {
    const int j = e >> 4, o = 16 * tc + (e & 15);
    if (o < N)
      Aw[(size_t)aff_w[j] * N + o] += VS[((size_t)(j >> 4) * KS + (o >> 2)) * 64 + 16 * (o & 3) + (j & 15)];
  }
}

// One launch per class NTA = ceil(a / 16): the block returns unless its walker is in the
// class, so each class gets its own register allocation.
template <int NT, int NTA>
__global__ void __launch_bounds__(128) k_wb_update_reg(const int32_t* __restrict__ accept,
                                                       const int32_t* __restrict__ aff, const int32_t* __restrict__ cu,
                                                       const int32_t* __restrict__ cd, int N, int D, int n_wd,
                                                       const double* __restrict__ vs, const double* __restrict__ xs,
                                                       double* __restrict__ st_A, double* __restrict__ st_Ainv) {
  __shared__ dbl4 sq[NTA * 64];
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int wd = (slot / NT) * 8 + xcd, tc = slot - (slot / NT) * NT;
  if (wd >= n_wd) return;
  const int w = wd / D;
  if (!accept[w]) return;
  const int a = cu[w] + cd[w];
  if (((a + 15) >> 4) != NTA && !(NTA == 4 && a > 64)) return;
  wb_update_tile<NT, NTA>(a, N, tc, aff + (size_t)w * N, vs + (size_t)wd * wb_vs_doubles(N),
                          xs + (size_t)wd * (wb_update_scratch(N) / sizeof(double)), st_Ainv + (size_t)wd * N * N,
                          st_A + (size_t)wd * N * N, sq);
}

// Generic form (any N): one block per (walker, det).  Q = V Ainv goes through global scratch
// (the second half of the per-(walker, det) scratch) in B-operand tiles; wavefronts own column
// tiles for Q and row tiles for the update.
__global__ void __launch_bounds__(256) k_wb_update(const int32_t* __restrict__ accept, const int32_t* __restrict__ aff,
                                                   const int32_t* __restrict__ cu, const int32_t* __restrict__ cd,
                                                   int N, int D, const double* __restrict__ vs,
                                                   double* __restrict__ xs, double* __restrict__ st_A,
                                                   double* __restrict__ st_Ainv) {
  const int wd = blockIdx.x;
  const int w = wd / D;
  if (!accept[w]) return;
  const int a = cu[w] + cd[w];
  const int nT = (a + 15) >> 4, nK = 4 * nT;
  const int NT = (N + 15) >> 4, KS = (N + 3) >> 2;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6, m = lane & 15, g = lane >> 4;
  double* Ai = st_Ainv + (size_t)wd * N * N;
  double* Aw = st_A + (size_t)wd * N * N;
  const double* VS = vs + (size_t)wd * wb_vs_doubles(N);
  const double* XS = xs + (size_t)wd * (wb_update_scratch(N) / sizeof(double));
  double* QS = xs + (size_t)wd * (wb_update_scratch(N) / sizeof(double)) + (size_t)NT * 16 * 64;
  // Q = V Ainv; register i of lane l is Q[16t + 4i + g][16tc + m], the B operand of k step
  // 4t + i in the same lane
  for (int tc = wv; tc < NT; tc += nw) {
    dbl4 acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int c = 16 * tc + m;
    for (int ks = 0; ks < KS; ++ks) {
      const int o = 4 * ks + g;
      const double b = (o < N && c < N) ? Ai[(size_t)o * N + c] : 0.0;
      for (int t = 0; t < nT; ++t) acc[t] = mfma_f64(VS[((size_t)t * KS + ks) * 64 + lane], b, acc[t]);
    }
    for (int t = 0; t < nT; ++t)
      for (int i = 0; i < 4; ++i) QS[((size_t)tc * nK + 4 * t + i) * 64 + lane] = acc[t][i];
  }
  __threadfence_block();
  __syncthreads();
  // Ainv <- Ainv - X Q, one 16x16 tile at a time
  for (int to = wv; to < NT; to += nw) {
    const int r0 = 16 * to + g;  // register i holds row r0 + 4i
    for (int tc = 0; tc < NT; ++tc) {
      const int c = 16 * tc + m;
      dbl4 acc;
      for (int i = 0; i < 4; ++i) acc[i] = (r0 + 4 * i < N && c < N) ? Ai[(size_t)(r0 + 4 * i) * N + c] : 0.0;
      for (int ks = 0; ks < nK; ++ks)
        acc = mfma_f64(-XS[((size_t)to * nK + ks) * 64 + lane], QS[((size_t)tc * nK + ks) * 64 + lane], acc);
      for (int i = 0; i < 4; ++i)
        if (r0 + 4 * i < N && c < N) Ai[(size_t)(r0 + 4 * i) * N + c] = acc[i];
    }
  }
  for (int e = tid; e < a * N; e += blockDim.x) {
    const int j = e / N, o = e - j * N;
    Aw[(size_t)aff[(size_t)w * N + j] * N + o] +=
        VS[((size_t)(j >> 4) * KS + (o >> 2)) * 64 + 16 * (o & 3) + (j & 15)];
  }
}

void launch_wb_update(hipStream_t s, const int32_t* accept, const int32_t* aff, const int32_t* cu, const int32_t* cd,
                      int B, int N, int D, const double* vs, const double* sbuf, double* scratch, double* st_A,
                      double* st_Ainv) {
  k_wb_x<<<B * D, 256, 0, s>>>(accept, aff, cu, cd, N, D, st_Ainv, sbuf, scratch);
  const int NT = (N + 15) / 16;
  const int G = (B * D + 7) / 8 * 8 * NT;
#define WB_REG(n)                                                                                         \
  case n:                                                                                                 \
    k_wb_update_reg<n, 1><<<G, 128, 0, s>>>(accept, aff, cu, cd, N, D, B * D, vs, scratch, st_A, st_Ainv); \
    k_wb_update_reg<n, 2><<<G, 128, 0, s>>>(accept, aff, cu, cd, N, D, B * D, vs, scratch, st_A, st_Ainv); \
    if (N > 32) {                                                                                         \
      k_wb_update_reg<n, 3><<<G, 128, 0, s>>>(accept, aff, cu, cd, N, D, B * D, vs, scratch, st_A, st_Ainv); \
      k_wb_update_reg<n, 4><<<G, 128, 0, s>>>(accept, aff, cu, cd, N, D, B * D, vs, scratch, st_A, st_Ainv); \
    }                                                                                                     \
    return;
  switch (NT) {
    WB_REG(1) WB_REG(2) WB_REG(3) WB_REG(4) WB_REG(5) WB_REG(6)
    WB_REG(7) WB_REG(8) WB_REG(9) WB_REG(10) WB_REG(11) WB_REG(12)
    default: break;
  }
#undef WB_REG
  k_wb_update<<<B * D, 256, 0, s>>>(accept, aff, cu, cd, N, D, vs, scratch, st_A, st_Ainv);
}

// Per walker: Jastrow sums with the affected rows replaced by the proposal rows
// (MLP Jastrow components, attention normaliser/values per spin) and the attention readout
// input vals/norm.  Rows are state rows (group order) unless rowof >= 0.
__global__ void __launch_bounds__(256) k_mc_jsum(const float* __restrict__ st_jr, const float* __restrict__ p_jr,
                                                 const float* __restrict__ st_av, const float* __restrict__ p_av,
                                                 const int32_t* __restrict__ rowof, int B, int N, int n_up, int H,
                                                 int Dk, int has_mlp, int has_att, double* __restrict__ jsum,
                                                 float* __restrict__ rin) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int HD = H * Dk, nc = H + HD;
  auto row_of = [&](int i, const float* st, const float* pp, int ld) {
    const int t = rowof[(size_t)w * N + i];
    return t >= 0 ? pp + (size_t)t * ld : st + (size_t)group_of(w, i, B, n_up, N - n_up) * ld;
  };
  if (has_mlp && wv == 0) {
    // per-spin sums, then their sum (the grouping of the full evaluation); lanes take the
    // electrons and the partial sums are added by a fixed shuffle tree
    double a[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i = lane; i < N; i += 64) {
      const float* row = row_of(i, st_jr, p_jr, 2);
      const int sp = i < n_up ? 0 : 1;
      a[sp][0] += row[0];
      a[sp][1] += row[1];
    }
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
#pragma unroll
      for (int c = 0; c < 2; ++c) a[sp][c] = warp_sum(a[sp][c]);
    if (lane < 2) jsum[(size_t)w * 2 + lane] = a[0][lane] + a[1][lane];
  }
  if (!has_att) return;
  // attention read-in sums: one thread per column, electrons in order; the rows of a chunk
  // of 128 electrons are looked up once into LDS
  __shared__ const float* s_row[128];
  const int ncol = 2 * HD;
  for (int col0 = 0; col0 < ncol; col0 += blockDim.x) {
    const int col = col0 + (int)threadIdx.x;
    const int sp = col < HD ? 0 : 1, hk = col - sp * HD, h = hk / Dk;
    double nsum = 0.0, vsum = 0.0;
    for (int i0 = 0; i0 < N; i0 += 128) {
      const int ni = min(128, N - i0);
      __syncthreads();
      if ((int)threadIdx.x < ni) s_row[threadIdx.x] = row_of(i0 + threadIdx.x, st_av, p_av, nc);
      __syncthreads();
      if (col < ncol) {
        const int lo = max(i0, sp ? n_up : 0) - i0, hi = min(i0 + ni, sp ? N : n_up) - i0;
#pragma unroll 8
        for (int ii = lo; ii < hi; ++ii) {
          const float* row = s_row[ii];
          nsum += row[h];
          vsum += row[H + hk];
        }
      }
    }
    if (col < ncol) rin[(size_t)w * ncol + col] = (float)(vsum * (1.0 / nsum));
  }
}

// Change of the e-e cusp Jastrow sum_{i<j} fa al^2 / (al + d_ij) (FP64) when electron k of
// walker w moves from st_r to p_r: only the pairs (k, j) change.  One wave per walker,
// lanes over j, fixed shuffle-tree reduction.
__global__ void __launch_bounds__(64) k_mc_cusp_delta(const float* __restrict__ st_r, const float* __restrict__ p_r,
                                                      int N, int n_up, int k, const float* __restrict__ cp,
                                                      double* __restrict__ cdelta) {
  const int w = blockIdx.x, lane = threadIdx.x;
  const float* ro = st_r + (size_t)w * N * 3;
  const float* rn = p_r + (size_t)w * N * 3;
  const double xo = ro[3 * k], yo = ro[3 * k + 1], zo = ro[3 * k + 2];
  const double xn = rn[3 * k], yn = rn[3 * k + 1], zn = rn[3 * k + 2];
  double acc = 0.0;
  for (int j = lane; j < N; j += 64) {
    if (j == k) continue;
    const double xj = ro[3 * j], yj = ro[3 * j + 1], zj = ro[3 * j + 2];
    const bool same = (k < n_up) == (j < n_up);
    const double al = same ? cp[0] : cp[1];
    const double fa = same ? cp[2] : cp[3];
    const double d_o = sqrt((xo - xj) * (xo - xj) + (yo - yj) * (yo - yj) + (zo - zj) * (zo - zj));
    const double d_n = sqrt((xn - xj) * (xn - xj) + (yn - yj) * (yn - yj) + (zn - zj) * (zn - zj));
    acc += fa * al * al * (1.0 / (al + d_n) - 1.0 / (al + d_o));
  }
  acc = warp_sum(acc);
  if (lane == 0) cdelta[w] = acc;
}

void launch_mc_cusp_delta(hipStream_t s, const float* st_r, const float* p_r, int B, int N, int n_up, int k,
                          const float* cusp_params, double* cdelta) {
  k_mc_cusp_delta<<<B, 64, 0, s>>>(st_r, p_r, N, n_up, k, cusp_params, cdelta);
}

// Per walker: cusp total sum_i cval[i] (the pair sums of the full evaluation).
__global__ void __launch_bounds__(64) k_mc_cusp_total(const double* __restrict__ cval, int N,
                                                      double* __restrict__ ccusp) {
  const int w = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < N; i += 64) acc += cval[(size_t)w * N + i];
  acc = warp_sum(acc);
  if (threadIdx.x == 0) ccusp[w] = acc;
}

void launch_mc_cusp_total(hipStream_t s, const double* cval, int B, int N, double* ccusp) {
  k_mc_cusp_total<<<B, 64, 0, s>>>(cval, N, ccusp);
}

void launch_mc_jsum(hipStream_t s, const float* st_jr, const float* p_jr, const float* st_av, const float* p_av,
                    const int32_t* rowof, int B, int N, int n_up, int H, int Dk, int has_mlp, int has_att,
                    double* jsum, float* rin) {
  k_mc_jsum<<<B, 256, 0, s>>>(st_jr, p_jr, st_av, p_av, rowof, B, N, n_up, H, Dk, has_mlp, has_att, jsum, rin);
}

// log psi of the proposal and the Metropolis decision (mcmc.py:233-238):
// accept if 2 (logpsi_new - logpsi_old) > log u.  mode 1 (caller-driven update) accepts
// nothing here; the caller commits explicitly.
__global__ void k_mc_accept(int B, int D, int use_mlp, int use_log, int use_att, int cusp,
                            const double* __restrict__ st_ld, const double* __restrict__ st_sg,
                            const double* __restrict__ ldS, const double* __restrict__ sgS,
                            const double* __restrict__ jsum, const float* __restrict__ att_out,
                            const float* __restrict__ att_scale, const float* __restrict__ att_zb,
                            const double* __restrict__ ccusp, const double* __restrict__ cdelta,
                            const double* __restrict__ st_logpsi,
                            const double* __restrict__ p_u, int mode, double* __restrict__ p_ld,
                            double* __restrict__ p_sg, double* __restrict__ p_logpsi, double* __restrict__ p_sign,
                            int32_t* __restrict__ accept, int32_t* __restrict__ n_acc) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  double m = -1e300;
  for (int d = 0; d < D; ++d) {
    const double l = st_ld[w * D + d] + ldS[w * D + d];
    p_ld[w * D + d] = l;
    p_sg[w * D + d] = st_sg[w * D + d] * sgS[w * D + d];
    m = fmax(m, l);
  }
  double sum = 0.0;
  for (int d = 0; d < D; ++d) sum += p_sg[w * D + d] * exp(p_ld[w * D + d] - m);
  double logpsi = m + log(fabs(sum));
  double sign = sum < 0 ? -1.0 : 1.0;
  double J[2] = {0.0, 0.0};
  if (use_mlp || use_log) {
    J[0] += jsum[2 * w];
    J[1] += jsum[2 * w + 1];
  }
  if (use_att)
    for (int c = 0; c < 2; ++c) J[c] += (double)att_out[2 * w + c] * att_scale[c] + att_zb[c];
  if (use_mlp) logpsi += J[0];
  if (use_log) {
    logpsi += log(fabs(J[1]));
    sign *= J[1] < 0 ? -1.0 : 1.0;
  }
  if (cusp) logpsi += ccusp[w] + cdelta[w];
  p_logpsi[w] = logpsi;
  p_sign[w] = sign;
  int acc = 0;
  if (mode == 0) acc = 2.0 * (logpsi - st_logpsi[w]) > log(p_u[w]);
  accept[w] = acc;
  if (n_acc) n_acc[w] += acc;
}

void launch_mc_accept(hipStream_t s, int B, int D, int use_mlp, int use_log, int use_att, int cusp, const double* st_ld,
                      const double* st_sg, const double* ldS, const double* sgS, const double* jsum,
                      const float* att_out, const float* att_scale, const float* att_zb, const double* ccusp,
                      const double* cdelta, const double* st_logpsi, const double* p_u, int mode, double* p_ld,
                      double* p_sg, double* p_logpsi, double* p_sign, int32_t* accept, int32_t* n_acc) {
  k_mc_accept<<<(B + 63) / 64, 64, 0, s>>>(B, D, use_mlp, use_log, use_att, cusp, st_ld, st_sg, ldS, sgS, jsum, att_out,
                                           att_scale, att_zb, ccusp, cdelta, st_logpsi, p_u, mode, p_ld, p_sg,
                                           p_logpsi, p_sign, accept, n_acc);
}

// Commit the compact per-row proposal buffers of accepted walkers into the state.
__global__ void k_mc_commit_rows(const int32_t* __restrict__ accept, const int2* __restrict__ list,
                                 const int32_t* __restrict__ dT, int B, int N, int n_up, int F, int ND, int nav,
                                 const float* __restrict__ p_h, const float* __restrict__ p_orb,
                                 const float* __restrict__ p_jr, const float* __restrict__ p_av, float* __restrict__ st_h,
                                 float* __restrict__ st_orb, float* __restrict__ st_jr, float* __restrict__ st_av) {
  const int t = blockIdx.x;
  if (dT && t >= *dT) return;
  const int w = list[t].x, i = list[t].y;
  if (!accept[w]) return;
  const size_t g = group_of(w, i, B, n_up, N - n_up);
  for (int c = threadIdx.x; c < F; c += blockDim.x) st_h[g * F + c] = p_h[(size_t)t * F + c];
  for (int c = threadIdx.x; c < ND; c += blockDim.x) st_orb[g * ND + c] = p_orb[(size_t)t * ND + c];
  if (st_jr && threadIdx.x < 2) st_jr[g * 2 + threadIdx.x] = p_jr[(size_t)t * 2 + threadIdx.x];
  if (st_av)
    for (int c = threadIdx.x; c < nav; c += blockDim.x) st_av[g * nav + c] = p_av[(size_t)t * nav + c];
}

// Commit the per-walker proposal values (moved electron and scalars) of accepted walkers,
// then make the proposal positions equal to the state positions again for every walker.
__global__ void k_mc_commit_walker(const int32_t* __restrict__ accept, int B, int N, int n_up, int D, int F5, int ND,
                                   int k, const float* __restrict__ p_rk, const float* __restrict__ p_M5k,
                                   const float* __restrict__ p_envk, const double* __restrict__ p_ld,
                                   const double* __restrict__ p_sg, const double* __restrict__ p_logpsi,
                                   const double* __restrict__ p_sign, float* __restrict__ st_r,
                                   float* __restrict__ st_M5, float* __restrict__ st_env, double* __restrict__ st_ld,
                                   double* __restrict__ st_sg, double* __restrict__ st_logpsi,
                                   double* __restrict__ st_sign, float* __restrict__ p_r,
                                   const double* __restrict__ cdelta, double* __restrict__ ccusp) {
  const int w = blockIdx.x;
  float* rk = st_r + ((size_t)w * N + k) * 3;
  if (accept[w]) {
    const size_t g = group_of(w, k, B, n_up, N - n_up);
    for (int c = threadIdx.x; c < F5; c += blockDim.x) st_M5[g * F5 + c] = p_M5k[(size_t)w * F5 + c];
    for (int c = threadIdx.x; c < ND; c += blockDim.x) st_env[g * ND + c] = p_envk[(size_t)w * ND + c];
    if (threadIdx.x < 3) rk[threadIdx.x] = p_rk[3 * w + threadIdx.x];
    if (threadIdx.x < D) {
      st_ld[w * D + threadIdx.x] = p_ld[w * D + threadIdx.x];
      st_sg[w * D + threadIdx.x] = p_sg[w * D + threadIdx.x];
    }
    if (threadIdx.x == 0) {
      st_logpsi[w] = p_logpsi[w];
      st_sign[w] = p_sign[w];
      if (ccusp) ccusp[w] += cdelta[w];
    }
  }
  if (threadIdx.x < 3) p_r[((size_t)w * N + k) * 3 + threadIdx.x] = rk[threadIdx.x];
}

void launch_mc_commit(hipStream_t s, const int32_t* accept, const int2* list, int T, const int32_t* dT, int B, int N,
                      int n_up, int D,
                      int F, int ND, int nav, int k, const float* p_h, const float* p_orb, const float* p_jr,
                      const float* p_av, const float* p_rk, const float* p_M5k, const float* p_envk, const double* p_ld,
                      const double* p_sg, const double* p_logpsi, const double* p_sign, float* st_h, float* st_orb,
                      float* st_jr, float* st_av, float* st_r, float* st_M5, float* st_env, double* st_ld,
                      double* st_sg, double* st_logpsi, double* st_sign, float* p_r, const double* cdelta,
                      double* ccusp) {
  if (T > 0)
    k_mc_commit_rows<<<T, 256, 0, s>>>(accept, list, dT, B, N, n_up, F, ND, nav, p_h, p_orb, p_jr, p_av, st_h,
                                       st_orb, st_jr, st_av);
  k_mc_commit_walker<<<B, 256, 0, s>>>(accept, B, N, n_up, D, 5 * F, ND, k, p_rk, p_M5k, p_envk, p_ld, p_sg, p_logpsi,
                                       p_sign, st_r, st_M5, st_env, st_ld, st_sg, st_logpsi, st_sign, p_r, cdelta,
                                       ccusp);
}

// fill kernel for counters (usable under stream capture, unlike hipMemsetAsync)
__global__ void k_fill_i32(int32_t* p, int n, int v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

void launch_fill_i32(hipStream_t s, int32_t* p, int n, int v) {
  if (n > 0) k_fill_i32<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
}

__global__ void k_fill_f64(double* p, int n, double v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

void launch_fill_f64(hipStream_t s, double* p, int n, double v) {
  if (n > 0) k_fill_f64<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
}

}  // namespace fire
