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

// ---------------------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon et al., SC'11): key (seed, walker),
// counter (step, lane, 0, 0).
// ---------------------------------------------------------------------------------------
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

__device__ __forceinline__ double u01(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

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
    // truncated so that every buffer access stays in bounds; the step is then invalid)
    int cd = count - cu;
    if (count > cap) {
      overflow[0] = 1;
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
    stats[0] += su[1023] + sd[1023];
    stats[1] = max(stats[1], (int64_t)sm[1023]);
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

// Woodbury preparation per (walker, det): new Slater rows Phi_new of the affected
// electrons, S = I + V Ainv[:, aff] with V = Phi_new - A[aff], log|det S|, sign, S^-1.
// Writes V = Phi_new - A[aff] (not Phi_new) to `phin`.  Dynamic LDS: S [amax][amax], a pivot
// column, pivots and (PLDS) P = Ainv[:, aff] [N][amax]; without PLDS, P is read from Ainv.
template <bool PLDS>
__global__ void __launch_bounds__(256) k_wb_prep(const float* __restrict__ p_orb, const float* __restrict__ p_envk,
                                                 const float* __restrict__ st_env, const double* __restrict__ st_A,
                                                 const double* __restrict__ st_Ainv, const int32_t* __restrict__ aff,
                                                 const int32_t* __restrict__ cu, const int32_t* __restrict__ cd,
                                                 const int32_t* __restrict__ rowof, int B, int N, int n_up, int D,
                                                 int k, int amax, double* __restrict__ phin,
                                                 double* __restrict__ sinv, double* __restrict__ ldS,
                                                 double* __restrict__ sgS) {
  extern __shared__ unsigned char smem_raw[];
  double* S = reinterpret_cast<double*>(smem_raw);  // [amax][amax]
  double* colc = S + (size_t)amax * amax;            // [amax]
  int* piv = reinterpret_cast<int*>(colc + amax);    // [amax]
  double* P = reinterpret_cast<double*>(smem_raw + (size_t)amax * amax * 8 + (size_t)amax * 8 +
                                        (((size_t)amax * 4 + 7) & ~(size_t)7));  // [N][amax] (PLDS)
  __shared__ double s_val[4];
  __shared__ int s_idx[4];
  __shared__ int s_p;
  const int wd = blockIdx.x;
  const int w = wd / D, d = wd - w * D;
  const int a = cu[w] + cd[w];
  const int ND = N * D;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* Ai = st_Ainv + (size_t)wd * N * N;
  const double* Aw = st_A + (size_t)wd * N * N;
  double* PH = phin + (size_t)wd * amax * N;
  // Phi_new rows (FP64) and P = Ainv[:, aff]
  for (int e = tid; e < a * N; e += blockDim.x) {
    const int j = e / N, o = e - j * N;
    const int i = aff[(size_t)w * N + j];
    const int t = rowof[(size_t)w * N + i];
    const float env = (i == k) ? p_envk[(size_t)w * ND + d * N + o]
                               : st_env[(size_t)group_of(w, i, B, n_up, N - n_up) * ND + d * N + o];
    // product rounded in FP32 exactly as the full evaluation forms the Slater matrix
    const double ph = (double)__fmul_rn(p_orb[(size_t)t * ND + d * N + o], env);
    PH[(size_t)j * N + o] = ph - Aw[(size_t)i * N + o];  // V = Phi_new - A[aff] (A rows restored by the update)
    if (PLDS) P[(size_t)o * amax + j] = Ai[(size_t)o * N + i];
  }
  __syncthreads();
  // S = I + V P: one row j of V per wavefront pass, P from LDS
  for (int e = tid; e < a * a; e += blockDim.x) {
    const int j = e / a, l = e - j * a;
    const double* vr = PH + (size_t)j * N;
    const int cl = aff[(size_t)w * N + l];
    auto pv = [&](int o) { return PLDS ? P[(size_t)o * amax + l] : Ai[(size_t)o * N + cl]; };
    double acc0 = 0.0, acc1 = 0.0;
    int o = 0;
    for (; o + 2 <= N; o += 2) {
      acc0 += vr[o] * pv(o);
      acc1 += vr[o + 1] * pv(o + 1);
    }
    if (o < N) acc0 += vr[o] * pv(o);
    S[j * amax + l] = (j == l ? 1.0 : 0.0) + (acc0 + acc1);
  }
  __syncthreads();
  // Gauss-Jordan inverse of S with partial pivoting; log|det S| and sign
  double ld = 0.0;
  int flips = 0;
  for (int c = 0; c < a; ++c) {
    double best = -1.0;
    int bi = a;
    for (int rr = c + tid; rr < a; rr += blockDim.x) {
      const double v = fabs(S[rr * amax + c]);
      if (v > best || (v == best && rr < bi)) { best = v; bi = rr; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { s_val[wid] = best; s_idx[wid] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b = s_val[0];
      int ii = s_idx[0];
      for (int q = 1; q < 4; ++q)
        if (s_val[q] > b || (s_val[q] == b && s_idx[q] < ii)) { b = s_val[q]; ii = s_idx[q]; }
      s_p = ii;
      piv[c] = ii;
    }
    __syncthreads();
    const int p = s_p;
    if (p != c)
      for (int cc = tid; cc < a; cc += blockDim.x) {
        const double t = S[c * amax + cc];
        S[c * amax + cc] = S[p * amax + cc];
        S[p * amax + cc] = t;
      }
    __syncthreads();
    const double pv = S[c * amax + c];
    if (tid == 0) {
      ld += log(fabs(pv));
      flips += (p != c) + (pv < 0.0);
    }
    const double ip = 1.0 / pv;
    __syncthreads();
    for (int cc = tid; cc < a; cc += blockDim.x) S[c * amax + cc] = (cc == c) ? ip : S[c * amax + cc] * ip;
    __syncthreads();
    for (int rr = tid; rr < a; rr += blockDim.x) colc[rr] = S[rr * amax + c];
    __syncthreads();
    for (int i = wid; i < a; i += blockDim.x >> 6) {
      if (i == c) continue;
      const double fk = colc[i];
      for (int j = lane; j < a; j += 64) S[i * amax + j] = (j == c) ? -fk * ip : S[i * amax + j] - fk * S[c * amax + j];
    }
    __syncthreads();
  }
  for (int c = a - 1; c >= 0; --c) {
    const int p = piv[c];
    if (p != c)
      for (int rr = tid; rr < a; rr += blockDim.x) {
        const double t = S[rr * amax + c];
        S[rr * amax + c] = S[rr * amax + p];
        S[rr * amax + p] = t;
      }
    __syncthreads();
  }
  double* SI = sinv + (size_t)wd * amax * amax;
  for (int e = tid; e < a * a; e += blockDim.x) {
    const int j = e / a, l = e - j * a;
    SI[j * amax + l] = S[j * amax + l];
  }
  if (tid == 0) {
    ldS[wd] = ld;
    sgS[wd] = (flips & 1) ? -1.0 : 1.0;
  }
}

static size_t wb_prep_fixed(int amax) {
  return (size_t)amax * amax * 8 + (size_t)amax * 8 + (((size_t)amax * 4 + 7) & ~(size_t)7);
}
// Dynamic LDS of the preparation: P staged when that keeps the block within 64 KB (several
// blocks per CU), else P is read from Ainv.
size_t wb_prep_smem(int N, int amax) {
  const size_t with_p = wb_prep_fixed(amax) + (size_t)N * amax * 8;
  return with_p <= 64 * 1024 ? with_p : wb_prep_fixed(amax);
}

void launch_wb_prep(hipStream_t s, const float* p_orb, const float* p_envk, const float* st_env, const double* st_A,
                    const double* st_Ainv, const int32_t* aff, const int32_t* cu, const int32_t* cd,
                    const int32_t* rowof, int B, int N, int n_up, int D, int k, int amax, double* phin, double* sinv,
                    double* ldS, double* sgS) {
  const size_t sm = wb_prep_smem(N, amax);
  if (sm > wb_prep_fixed(amax))
    k_wb_prep<true><<<B * D, 256, sm, s>>>(p_orb, p_envk, st_env, st_A, st_Ainv, aff, cu, cd, rowof, B, N, n_up, D, k,
                                           amax, phin, sinv, ldS, sgS);
  else
    k_wb_prep<false><<<B * D, 256, sm, s>>>(p_orb, p_envk, st_env, st_A, st_Ainv, aff, cu, cd, rowof, B, N, n_up, D,
                                            k, amax, phin, sinv, ldS, sgS);
}

hipError_t wb_set_smem_limit(size_t prep, size_t upd);

// Woodbury inverse update for accepted walkers (woodburry_update, utils.py:243):
//   Ainv <- Ainv - (Ainv[:, aff] S^-1) (V Ainv);  then A[aff] <- A[aff] + V,
// with V = Phi_new - A[aff] as left in `phin` by k_wb_prep.
//
// Thread c owns column c of Ainv; q = (V Ainv)[:, c] is held in AMAX registers.  V^T and
// X = Ainv[:, aff] S^-1 are staged per (walker, det) in global scratch (row o of both is
// read by every thread of the block: broadcast loads served by the caches), so the kernel
// needs only S^-1 in LDS and several blocks share a CU.  Ainv is streamed twice, coalesced.
template <int AMAX>
__global__ void __launch_bounds__(256) k_wb_update_cols(const int32_t* __restrict__ accept,
                                                        const int32_t* __restrict__ aff, const int32_t* __restrict__ cu,
                                                        const int32_t* __restrict__ cd, int N, int D, int amax,
                                                        const double* __restrict__ vmat,
                                                        const double* __restrict__ sinv, double* __restrict__ st_A,
                                                        double* __restrict__ st_Ainv, double* __restrict__ scr) {
  __shared__ double SL[AMAX * AMAX];
  __shared__ int s_af[AMAX];
  const int wd = blockIdx.x;
  const int w = wd / D;
  if (!accept[w]) return;
  const int a = cu[w] + cd[w];
  const int tid = threadIdx.x;
  double* Ai = st_Ainv + (size_t)wd * N * N;
  double* Aw = st_A + (size_t)wd * N * N;
  const double* V = vmat + (size_t)wd * amax * N;
  const double* SI = sinv + (size_t)wd * amax * amax;
  double* VT = scr + (size_t)wd * 2 * N * AMAX;  // [N][AMAX]
  double* X = VT + (size_t)N * AMAX;             // [N][AMAX]
  for (int j = tid; j < a; j += blockDim.x) s_af[j] = aff[(size_t)w * N + j];
  for (int e = tid; e < AMAX * AMAX; e += blockDim.x) {
    const int l = e / AMAX, j = e - l * AMAX;
    SL[e] = (l < a && j < a) ? SI[l * amax + j] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < N * AMAX; e += blockDim.x) {
    const int o = e / AMAX, j = e - o * AMAX;
    VT[e] = j < a ? V[(size_t)j * N + o] : 0.0;
  }
  // the j loops below run in chunks of 16 and skip chunks beyond this walker's count a
  // (uniform in the block), so the cost follows a, not the capacity AMAX
  for (int o = tid; o < N; o += blockDim.x) {
    double acc[AMAX];
#pragma unroll
    for (int j = 0; j < AMAX; ++j) acc[j] = 0.0;
    const double* ar = Ai + (size_t)o * N;
    for (int l = 0; l < a; ++l) {
      const double p = ar[s_af[l]];
#pragma unroll
      for (int ch = 0; ch < AMAX / 16; ++ch)
        if (ch * 16 < a) {
#pragma unroll
          for (int u = 0; u < 16; ++u) acc[ch * 16 + u] += p * SL[l * AMAX + ch * 16 + u];
        }
    }
#pragma unroll
    for (int j = 0; j < AMAX; ++j) X[(size_t)o * AMAX + j] = j < a ? acc[j] : 0.0;
  }
  __threadfence_block();
  __syncthreads();
  for (int c = tid; c < N; c += blockDim.x) {
    double q[AMAX];
#pragma unroll
    for (int j = 0; j < AMAX; ++j) q[j] = 0.0;
    for (int o = 0; o < N; ++o) {
      const double x = Ai[(size_t)o * N + c];
      const double* v = VT + (size_t)o * AMAX;
#pragma unroll
      for (int ch = 0; ch < AMAX / 16; ++ch)
        if (ch * 16 < a) {
#pragma unroll
          for (int u = 0; u < 16; ++u) q[ch * 16 + u] += v[ch * 16 + u] * x;
        }
    }
    for (int o = 0; o < N; ++o) {
      const double* xr = X + (size_t)o * AMAX;
      double acc = 0.0;
#pragma unroll
      for (int ch = 0; ch < AMAX / 16; ++ch)
        if (ch * 16 < a) {
#pragma unroll
          for (int u = 0; u < 16; ++u) acc += xr[ch * 16 + u] * q[ch * 16 + u];
        }
      Ai[(size_t)o * N + c] -= acc;
    }
  }
  for (int e = tid; e < a * N; e += blockDim.x) {
    const int j = e / N, o = e - j * N;
    Aw[(size_t)s_af[j] * N + o] += V[(size_t)j * N + o];
  }
}

// Generic form (any number of changed rows): X and Q = V Ainv in dynamic LDS.
__global__ void __launch_bounds__(256) k_wb_update(const int32_t* __restrict__ accept, const int32_t* __restrict__ aff,
                                                   const int32_t* __restrict__ cu, const int32_t* __restrict__ cd,
                                                   int N, int D, int amax, const double* __restrict__ vmat,
                                                   const double* __restrict__ sinv, double* __restrict__ st_A,
                                                   double* __restrict__ st_Ainv) {
  extern __shared__ unsigned char smem_raw[];
  double* X = reinterpret_cast<double*>(smem_raw);  // [N][amax]
  double* Q = X + (size_t)N * amax;                  // [amax][N]
  const int wd = blockIdx.x;
  const int w = wd / D;
  if (!accept[w]) return;
  const int a = cu[w] + cd[w];
  const int tid = threadIdx.x;
  double* Ai = st_Ainv + (size_t)wd * N * N;
  double* Aw = st_A + (size_t)wd * N * N;
  const double* V = vmat + (size_t)wd * amax * N;
  const double* SI = sinv + (size_t)wd * amax * amax;
  const int32_t* af = aff + (size_t)w * N;
  for (int e = tid; e < N * a; e += blockDim.x) {
    const int o = e / a, j = e - o * a;
    double acc = 0.0;
    for (int l = 0; l < a; ++l) acc += Ai[(size_t)o * N + af[l]] * SI[l * amax + j];
    X[(size_t)o * amax + j] = acc;
  }
  for (int e = tid; e < a * N; e += blockDim.x) {
    const int j = e / N, c = e - j * N;
    double acc = 0.0;
    for (int o = 0; o < N; ++o) acc += V[(size_t)j * N + o] * Ai[(size_t)o * N + c];
    Q[(size_t)j * N + c] = acc;
  }
  __syncthreads();
  for (int e = tid; e < N * N; e += blockDim.x) {
    const int o = e / N, c = e - o * N;
    double acc = 0.0;
    for (int j = 0; j < a; ++j) acc += X[(size_t)o * amax + j] * Q[(size_t)j * N + c];
    Ai[e] -= acc;
  }
  for (int e = tid; e < a * N; e += blockDim.x) {
    const int j = e / N, o = e - j * N;
    Aw[(size_t)af[j] * N + o] += V[(size_t)j * N + o];
  }
}

size_t wb_update_smem(int N, int amax) { return amax <= 64 ? 0 : (size_t)2 * N * amax * 8; }
size_t wb_update_scratch(int N, int amax) { return amax <= 64 ? (size_t)2 * N * 64 * 8 : 0; }

void launch_wb_update(hipStream_t s, const int32_t* accept, const int32_t* aff, const int32_t* cu, const int32_t* cd,
                      int B, int N, int D, int amax, const double* vmat, const double* sinv, double* st_A,
                      double* st_Ainv, double* scratch) {
  const int threads = std::min(256, ((N + 63) / 64) * 64);
  if (amax <= 16)
    k_wb_update_cols<16><<<B * D, threads, 0, s>>>(accept, aff, cu, cd, N, D, amax, vmat, sinv, st_A, st_Ainv, scratch);
  else if (amax <= 32)
    k_wb_update_cols<32><<<B * D, threads, 0, s>>>(accept, aff, cu, cd, N, D, amax, vmat, sinv, st_A, st_Ainv, scratch);
  else if (amax <= 48)
    k_wb_update_cols<48><<<B * D, threads, 0, s>>>(accept, aff, cu, cd, N, D, amax, vmat, sinv, st_A, st_Ainv, scratch);
  else if (amax <= 64)
    k_wb_update_cols<64><<<B * D, threads, 0, s>>>(accept, aff, cu, cd, N, D, amax, vmat, sinv, st_A, st_Ainv, scratch);
  else
    k_wb_update<<<B * D, 256, wb_update_smem(N, amax), s>>>(accept, aff, cu, cd, N, D, amax, vmat, sinv, st_A,
                                                            st_Ainv);
}

hipError_t wb_set_smem_limit(size_t prep, size_t upd) {
  hipError_t e = hipFuncSetAttribute((const void*)k_wb_prep<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)prep);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_wb_prep<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)prep);
  if (e != hipSuccess || upd == 0) return e;
  return hipFuncSetAttribute((const void*)k_wb_update, hipFuncAttributeMaxDynamicSharedMemorySize, (int)upd);
}

// Per walker: Jastrow sums with the affected rows replaced by the proposal rows
// (MLP Jastrow components, attention normaliser/values per spin) and the attention readout
// input vals/norm.  Rows are state rows (group order) unless rowof >= 0.
__global__ void k_mc_jsum(const float* __restrict__ st_jr, const float* __restrict__ p_jr,
                          const float* __restrict__ st_av, const float* __restrict__ p_av,
                          const int32_t* __restrict__ rowof, int B, int N, int n_up, int H, int Dk, int has_mlp,
                          int has_att, double* __restrict__ jsum, float* __restrict__ rin) {
  const int w = blockIdx.x;
  const int HD = H * Dk, nc = H + HD;
  if (has_mlp && threadIdx.x < 2) {
    // per-spin sums in electron order, then their sum (the order of the full evaluation)
    const int c = threadIdx.x;
    double acc[2] = {0.0, 0.0};
    for (int i = 0; i < N; ++i) {
      const int t = rowof[(size_t)w * N + i];
      acc[i < n_up ? 0 : 1] +=
          t >= 0 ? p_jr[(size_t)t * 2 + c] : st_jr[(size_t)group_of(w, i, B, n_up, N - n_up) * 2 + c];
    }
    jsum[(size_t)w * 2 + c] = acc[0] + acc[1];
  }
  if (!has_att) return;
  for (int col = threadIdx.x; col < 2 * HD; col += blockDim.x) {
    const int sp = col / HD, hk = col - sp * HD, h = hk / Dk;
    const int i0 = sp == 0 ? 0 : n_up, i1 = sp == 0 ? n_up : N;
    double nsum = 0.0, vsum = 0.0;
    for (int i = i0; i < i1; ++i) {
      const int t = rowof[(size_t)w * N + i];
      const float* row = t >= 0 ? p_av + (size_t)t * nc : st_av + (size_t)group_of(w, i, B, n_up, N - n_up) * nc;
      nsum += row[h];
      vsum += row[H + hk];
    }
    rin[(size_t)w * 2 * HD + col] = (float)(vsum * (1.0 / nsum));
  }
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
                            const double* __restrict__ cval, int N, const double* __restrict__ st_logpsi,
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
  if (cusp) {
    double cv = 0.0;
    for (int i = 0; i < N; ++i) cv += cval[(size_t)w * N + i];
    logpsi += cv;
  }
  p_logpsi[w] = logpsi;
  p_sign[w] = sign;
  int acc = 0;
  if (mode == 0) acc = 2.0 * (logpsi - st_logpsi[w]) > log(p_u[w]);
  accept[w] = acc;
  if (n_acc) n_acc[w] += acc;
}

void launch_mc_accept(hipStream_t s, int B, int D, int use_mlp, int use_log, int use_att, int cusp, const double* st_ld,
                      const double* st_sg, const double* ldS, const double* sgS, const double* jsum,
                      const float* att_out, const float* att_scale, const float* att_zb, const double* cval, int N,
                      const double* st_logpsi, const double* p_u, int mode, double* p_ld, double* p_sg,
                      double* p_logpsi, double* p_sign, int32_t* accept, int32_t* n_acc) {
  k_mc_accept<<<(B + 63) / 64, 64, 0, s>>>(B, D, use_mlp, use_log, use_att, cusp, st_ld, st_sg, ldS, sgS, jsum, att_out,
                                           att_scale, att_zb, cval, N, st_logpsi, p_u, mode, p_ld, p_sg, p_logpsi,
                                           p_sign, accept, n_acc);
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
                                   double* __restrict__ st_sign, float* __restrict__ p_r) {
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
                      double* st_sg, double* st_logpsi, double* st_sign, float* p_r) {
  if (T > 0)
    k_mc_commit_rows<<<T, 256, 0, s>>>(accept, list, dT, B, N, n_up, F, ND, nav, p_h, p_orb, p_jr, p_av, st_h,
                                       st_orb, st_jr, st_av);
  k_mc_commit_walker<<<B, 256, 0, s>>>(accept, B, N, n_up, D, 5 * F, ND, k, p_rk, p_M5k, p_envk, p_ld, p_sg, p_logpsi,
                                       p_sign, st_r, st_M5, st_env, st_ld, st_sg, st_logpsi, st_sign, p_r);
}

// fill kernel for counters (usable under stream capture, unlike hipMemsetAsync)
__global__ void k_fill_i32(int32_t* p, int n, int v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

void launch_fill_i32(hipStream_t s, int32_t* p, int n, int v) {
  if (n > 0) k_fill_i32<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
}

}  // namespace fire
