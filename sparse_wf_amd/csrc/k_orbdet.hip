// Orbitals x envelopes, Slater LU / inverse / log-determinant (FP64) and the sparse
// determinant Laplacian.
//
// References (sparse_wf): EfficientIsotropicEnvelopes envelopes.py:55-82,
// Orbitals.fwd_lap orbitals.py:148-174, merge_up_down / multiply_with_1el_fn
// sparse_fwd_lap.py:129-137,322-335, slog_and_inverse model/utils.py:144-175,
// sparse_slogdet_with_fwd_lap sparse_fwd_lap.py:294-319 (triplets :228-291).
#include "fire_common.hpp"
#include "fire_kernels.hpp"
#include "fire_kernels2.hpp"

namespace fire {

// Envelope coefficients of one electron for EfficientIsotropicEnvelopes:
//   env[d,o] = sum_{n,e} pi[n,d,e,o] q_{n,d,e},  q = exp(-sp[n,d,e] |r - R_n|),
//   d/dr_t q = -sp q n_t,  Laplacian q = q (sp^2 - 2 sp / |r - R_n|).
// Written as coef[d][g*NR + row][n*E + e] so that the contraction with pi (stored as
// [d][n*E + e][o]) is one FP32 GEMM per (spin block, determinant).
template <bool TAN>
__global__ void k_env_coef(const float* __restrict__ r, const float* __restrict__ R, int B, int N, int n_up, int M,
                           int D, int E, const float* __restrict__ sp_up, const float* __restrict__ sp_dn,
                           float* __restrict__ coef, const int2* __restrict__ list, int n_list) {
  const int g = blockIdx.x;
  int w, i;
  if (list) {
    w = list[g].x;
    i = list[g].y;
  } else {
    group_inv(g, B, n_up, N - n_up, w, i);
  }
  const float* ri = r + ((size_t)w * N + i) * 3;
  const float* sp = i < n_up ? sp_up : sp_dn;
  const int NR = TAN ? 5 : 1;
  const int ME = M * E;
  const size_t G = list ? (size_t)n_list : (size_t)B * N;
  for (int idx = threadIdx.x; idx < M * D * E; idx += blockDim.x) {
    const int m = idx / (D * E), de = idx - m * (D * E), d = de / E, e = de - d * E;
    const float dx = ri[0] - R[3 * m], dy = ri[1] - R[3 * m + 1], dz = ri[2] - R[3 * m + 2];
    const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
    const float s = sp[idx];
    const float q = expf(-s * dn);
    float* c = coef + (size_t)d * G * NR * ME + (size_t)g * NR * ME + m * E + e;
    c[0] = q;
    if (TAN) {
      const float inv = 1.0f / dn;
      c[ME] = q * (s * s - 2.0f * s * inv);
      c[2 * ME] = -s * q * dx * inv;
      c[3 * ME] = -s * q * dy * inv;
      c[4 * ME] = -s * q * dz * inv;
    }
  }
}

void launch_env_coef(hipStream_t s, const float* r, const float* R, int B, int N, int n_up, int M, int D, int E,
                     const float* sp_up, const float* sp_dn, bool tangents, float* coef, const int2* list,
                     int n_list) {
  const int G = list ? n_list : B * N;
  if (G == 0) return;
  if (tangents)
    k_env_coef<true><<<G, 256, 0, s>>>(r, R, B, N, n_up, M, D, E, sp_up, sp_dn, coef, list, n_list);
  else
    k_env_coef<false><<<G, 256, 0, s>>>(r, R, B, N, n_up, M, D, E, sp_up, sp_dn, coef, list, n_list);
}

// Phi = orbital rows x envelope rows (product with a one-electron function), in place, for
// the value, Laplacian and self-gradient rows of each group.  A neighbour row d O_i / d r_j
// (j != i) becomes E_i d O_i / d r_j; that scale is applied where the rows are consumed
// (k_det_lap stages them through LDS), so the neighbour rows are not rewritten here.
template <bool TAN>
__global__ void k_orb_product(float* __restrict__ O, int ldo, int ND, GroupDesc gd, const float* __restrict__ env) {
  const int g = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= ND) return;
  const int NR = TAN ? 5 : 1;
  const float* e = env + (size_t)g * NR * ND + c;
  float* o = O + gd.row(g) * ldo + c;
  const float ox = o[0], ex = e[0];
  o[0] = ox * ex;
  if (!TAN) return;
  const float ol = o[ldo], el = e[ND];
  float cross = 0.f;
  for (int t = 0; t < 3; ++t) {
    const float ot = o[(2 + t) * (int64_t)ldo], et = e[(2 + t) * ND];
    cross += ot * et;
    o[(2 + t) * (int64_t)ldo] = ox * et + ex * ot;
  }
  o[ldo] = ol * ex + el * ox + 2.f * cross;
}

void launch_orb_product(hipStream_t s, float* O, int ldo, int ND, int G, GroupDesc gd, const float* env,
                        bool tangents) {
  dim3 grid(G, (ND + 255) / 256);
  if (tangents)
    k_orb_product<true><<<grid, 256, 0, s>>>(O, ldo, ND, gd, env);
  else
    k_orb_product<false><<<grid, 256, 0, s>>>(O, ldo, ND, gd, env);
}

// A[w,d][i][o] (FP64) from the value rows of the orbital groups.
__global__ void k_gather_A(const float* __restrict__ O, int ldo, int ND, GroupDesc gd, int B, int N, int n_up,
                           int D, double* __restrict__ A) {
  const int g = blockIdx.x;
  int w, i;
  group_inv(g, B, n_up, N - n_up, w, i);
  const float* o = O + gd.row(g) * ldo;
  for (int c = threadIdx.x; c < ND; c += blockDim.x) {
    const int d = c / N, col = c - d * N;
    A[(((size_t)w * D + d) * N + i) * N + col] = (double)o[c];
  }
}

void launch_gather_A(hipStream_t s, const float* O, int ldo, int ND, GroupDesc gd, int B, int N, int n_up, int D,
                     double* A) {
  k_gather_A<<<B * N, 256, 0, s>>>(O, ldo, ND, gd, B, N, n_up, D, A);
}

// LU with partial pivoting (first maximum wins), log|det|, sign and the inverse, FP64.
// One workgroup per matrix; A is overwritten by its LU factors; Ainv row-major.
__global__ void __launch_bounds__(256) k_lu_inv(double* __restrict__ Aall, int N, double* __restrict__ Ainv_all,
                                                double* __restrict__ logdet, double* __restrict__ sign, int want_inv) {
  extern __shared__ unsigned char smem_raw[];
  int* perm = reinterpret_cast<int*>(smem_raw);
  __shared__ double s_val[4];
  __shared__ int s_idx[4];
  __shared__ int s_piv;
  const int mat = blockIdx.x;
  double* A = Aall + (size_t)mat * N * N;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int k = tid; k < N; k += blockDim.x) perm[k] = k;
  __syncthreads();
  int swaps = 0;
  for (int k = 0; k < N; ++k) {
    // pivot search over rows k..N-1 of column k
    double best = -1.0;
    int bi = N;
    for (int rr = k + tid; rr < N; rr += blockDim.x) {
      const double v = fabs(A[(size_t)rr * N + k]);
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
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
        if (s_val[q] > b || (s_val[q] == b && s_idx[q] < ii)) { b = s_val[q]; ii = s_idx[q]; }
      s_piv = ii;
      if (ii != k) {
        const int t = perm[k];
        perm[k] = perm[ii];
        perm[ii] = t;
      }
    }
    __syncthreads();
    const int p = s_piv;
    if (p != k) {
      ++swaps;
      for (int c = tid; c < N; c += blockDim.x) {
        const double t = A[(size_t)k * N + c];
        A[(size_t)k * N + c] = A[(size_t)p * N + c];
        A[(size_t)p * N + c] = t;
      }
    }
    __syncthreads();
    const double piv = A[(size_t)k * N + k];
    const double ipiv = 1.0 / piv;
    for (int rr = k + 1 + tid; rr < N; rr += blockDim.x) A[(size_t)rr * N + k] *= ipiv;
    __syncthreads();
    // trailing update: one row per wavefront pass, columns across lanes
    for (int rr = k + 1 + wid; rr < N; rr += blockDim.x >> 6) {
      const double l = A[(size_t)rr * N + k];
      for (int c = k + 1 + lane; c < N; c += 64) A[(size_t)rr * N + c] -= l * A[(size_t)k * N + c];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double ld = 0.0;
    int neg = 0;
    for (int k = 0; k < N; ++k) {
      const double dk = A[(size_t)k * N + k];
      ld += log(fabs(dk));
      neg += dk < 0.0;
    }
    logdet[mat] = ld;
    sign[mat] = ((swaps + neg) & 1) ? -1.0 : 1.0;
  }
  if (!want_inv) return;
  // inverse: column c solves L U x = P e_c (thread c; Ainv row-major so stores coalesce)
  double* X = Ainv_all + (size_t)mat * N * N;
  for (int c = tid; c < N; c += blockDim.x) {
    for (int k = 0; k < N; ++k) {
      double y = (perm[k] == c) ? 1.0 : 0.0;
      for (int j = 0; j < k; ++j) y -= A[(size_t)k * N + j] * X[(size_t)j * N + c];
      X[(size_t)k * N + c] = y;
    }
    for (int k = N - 1; k >= 0; --k) {
      double y = X[(size_t)k * N + c];
      for (int j = k + 1; j < N; ++j) y -= A[(size_t)k * N + j] * X[(size_t)j * N + c];
      X[(size_t)k * N + c] = y / A[(size_t)k * N + k];
    }
  }
}

// Gauss-Jordan inversion with partial pivoting (first maximum wins), in place, FP64, with
// the matrix resident in LDS (N*N doubles + pivots; N <= 141 with gfx950's 160 KB).
// Step k: pivot search in column k, row swap, scale the pivot row, eliminate column k from
// every other row (all (i, j) updates in parallel).  log|det| = sum log|pivot|,
// sign = (-1)^(#swaps + #negative pivots) -- the same as LU with partial pivoting
// (slog_and_inverse, model/utils.py:144-175).  The column swaps at the end undo the row
// permutation.  Writes Ainv (row-major FP64) and AinvT32 (transposed, FP32) when requested.
__global__ void __launch_bounds__(512) k_gj_inv_lds(const double* __restrict__ Aall, int N,
                                                    double* __restrict__ Ainv_all, float* __restrict__ AinvT32,
                                                    double* __restrict__ logdet, double* __restrict__ sign,
                                                    int want_inv) {
  extern __shared__ unsigned char smem_raw[];
  double* A = reinterpret_cast<double*>(smem_raw);
  double* colk = A + (size_t)N * N;  // column k before elimination
  int* piv = reinterpret_cast<int*>(colk + N);
  __shared__ double s_val[8];
  __shared__ int s_idx[8];
  __shared__ int s_p;
  __shared__ double s_pivot;
  const int mat = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nthr = blockDim.x;
  const double* Ag = Aall + (size_t)mat * N * N;
  for (int e = tid; e < N * N; e += nthr) A[e] = Ag[e];
  __syncthreads();
  double ld = 0.0;
  int flips = 0;
  for (int k = 0; k < N; ++k) {
    double best = -1.0;
    int bi = N;
    for (int rr = k + tid; rr < N; rr += nthr) {
      const double v = fabs(A[rr * N + k]);
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
      for (int q = 1; q < nthr >> 6; ++q)
        if (s_val[q] > b || (s_val[q] == b && s_idx[q] < ii)) { b = s_val[q]; ii = s_idx[q]; }
      s_p = ii;
      piv[k] = ii;
    }
    __syncthreads();
    const int p = s_p;
    if (p != k) {
      for (int cc = tid; cc < N; cc += nthr) {
        const double t = A[k * N + cc];
        A[k * N + cc] = A[p * N + cc];
        A[p * N + cc] = t;
      }
      __syncthreads();
    }
    if (tid == 0) s_pivot = A[k * N + k];
    __syncthreads();
    const double pv = s_pivot;
    if (tid == 0) {
      ld += log(fabs(pv));
      flips += (p != k) + (pv < 0.0);
    }
    const double ip = 1.0 / pv;
    // scale the pivot row (column k becomes 1/pivot); keep column k for the elimination
    for (int cc = tid; cc < N; cc += nthr) A[k * N + cc] = (cc == k) ? ip : A[k * N + cc] * ip;
    for (int rr = tid; rr < N; rr += nthr) colk[rr] = A[rr * N + k];
    __syncthreads();
    // eliminate column k from the other rows: row i -= A[i][k] * row k; column k -> -A[i][k]/pivot
    // (one row per wavefront pass, columns across lanes: no index division in the loop)
    for (int i = wid; i < N; i += nthr >> 6) {
      if (i == k) continue;
      const double fk = colk[i];
      double* ri = A + (size_t)i * N;
      const double* rk = A + (size_t)k * N;
      for (int j = lane; j < N; j += 64) ri[j] = (j == k) ? -fk * ip : ri[j] - fk * rk[j];
    }
    __syncthreads();
  }
  if (tid == 0) {
    logdet[mat] = ld;
    sign[mat] = (flips & 1) ? -1.0 : 1.0;
  }
  if (!want_inv) return;
  // undo the row permutation as column swaps, last pivot first
  for (int k = N - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int rr = tid; rr < N; rr += nthr) {
        const double t = A[rr * N + k];
        A[rr * N + k] = A[rr * N + p];
        A[rr * N + p] = t;
      }
    __syncthreads();
  }
  double* X = Ainv_all + (size_t)mat * N * N;
  for (int e = tid; e < N * N; e += nthr) X[e] = A[e];
  if (!AinvT32) return;
  float* XT = AinvT32 + (size_t)mat * N * N;
  for (int j = wid; j < N; j += nthr >> 6)
    for (int i = lane; i < N; i += 64) XT[(size_t)j * N + i] = (float)A[(size_t)i * N + j];
}

// Register-resident Gauss-Jordan inverse with implicit partial pivoting, FP64.  One block of
// NW wavefronts per matrix; element (i, j) lives in a register of thread (i % NW, j % 64) at
// [i / NW][j / 64], so the matrix never touches LDS.  Step k picks the unused row p with the
// largest |A[p][k]| (ties: lowest row), broadcasts row p and column k through LDS and updates
// every element in registers (two barriers per step).  Rows are not swapped: the stored result
// S satisfies Ainv[q(r)][piv(c)] = S[r][c] with piv(k) = pivot row of step k and q its
// inverse; det(A) = sign(piv) * prod(pivots).
template <int NW, int RPT, int CPT>
__global__ void __launch_bounds__(NW * 64) k_gj_inv_reg(const double* __restrict__ Aall, int N,
                                                        double* __restrict__ Ainv_all, float* __restrict__ AinvT32,
                                                        double* __restrict__ logdet, double* __restrict__ sign,
                                                        int want_inv) {
  __shared__ double s_row[CPT * 64];
  __shared__ double s_col[RPT * NW];
  __shared__ double s_val[NW];
  __shared__ int s_idx[NW];
  __shared__ int s_piv[RPT * NW];
  __shared__ int s_inv[RPT * NW];
  const int mat = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* Ag = Aall + (size_t)mat * N * N;
  double S[RPT][CPT];
#pragma unroll
  for (int m = 0; m < RPT; ++m)
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int i = wid + NW * m, j = lane + 64 * c;
      S[m][c] = (i < N && j < N) ? Ag[(size_t)i * N + j] : 0.0;
    }
  uint32_t used = 0;  // bit m: row wid + NW * m was a pivot (uniform in the wavefront)
  double ld = 0.0;
  int neg = 0;
  for (int k = 0; k < N; ++k) {
    const int lk = k & 63, ck = k >> 6;
    if (lane == lk) {
      double best = -1.0;
      int bi = N;
#pragma unroll
      for (int m = 0; m < RPT; ++m) {
        const int i = wid + NW * m;
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < CPT; ++c)
          if (c == ck) x = S[m][c];
        const double v = fabs(x);
        if (i < N && !((used >> m) & 1u) && (v > best || (v == best && i < bi))) {
          best = v;
          bi = i;
        }
      }
      s_val[wid] = best;
      s_idx[wid] = bi;
    }
    __syncthreads();
    double best = s_val[0];
    int p = s_idx[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) {
      const double v = s_val[q];
      const int iq = s_idx[q];
      if (v > best || (v == best && iq < p)) {
        best = v;
        p = iq;
      }
    }
    const int mp = p / NW;
    if (wid == p - NW * mp) {
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        double x = 0.0;
#pragma unroll
        for (int m = 0; m < RPT; ++m)
          if (m == mp) x = S[m][c];
        s_row[lane + 64 * c] = x;
      }
      used |= 1u << mp;
    }
    if (lane == lk) {
#pragma unroll
      for (int m = 0; m < RPT; ++m) {
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < CPT; ++c)
          if (c == ck) x = S[m][c];
        s_col[wid + NW * m] = x;
      }
    }
    if (tid == 0) s_piv[k] = p;
    __syncthreads();
    const double pv = s_row[k];
    const double ip = 1.0 / pv;
    if (tid == 0) {
      ld += log(fabs(pv));
      neg += pv < 0.0;
    }
    double rs[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) rs[c] = s_row[lane + 64 * c] * ip;
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
      const int i = wid + NW * m;
      if (i == p) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) S[m][c] = (lane + 64 * c == k) ? ip : S[m][c] * ip;
      } else {
        const double fk = s_col[i < N ? i : 0];
#pragma unroll
        for (int c = 0; c < CPT; ++c) S[m][c] = (lane + 64 * c == k) ? -fk * ip : S[m][c] - fk * rs[c];
      }
    }
  }
  __syncthreads();
  for (int k = tid; k < N; k += NW * 64) s_inv[s_piv[k]] = k;
  __syncthreads();
  if (tid == 0) {
    // parity of k -> piv(k): N minus the number of cycles (visited marks kept in s_inv's sign bit)
    int cycles = 0;
    for (int s0 = 0; s0 < N; ++s0) {
      if (s_inv[s0] < 0) continue;
      ++cycles;
      for (int t = s0; s_inv[t] >= 0; t = s_piv[t]) s_inv[t] = -1 - s_inv[t];
    }
    for (int t = 0; t < N; ++t) s_inv[t] = -1 - s_inv[t];
    logdet[mat] = ld;
    sign[mat] = (((N - cycles) + neg) & 1) ? -1.0 : 1.0;
  }
  if (!want_inv) return;
  __syncthreads();
  double* X = Ainv_all + (size_t)mat * N * N;
  float* XT = AinvT32 ? AinvT32 + (size_t)mat * N * N : nullptr;
#pragma unroll
  for (int m = 0; m < RPT; ++m)
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int r = wid + NW * m, cc = lane + 64 * c;
      if (r < N && cc < N) {
        const int i = s_inv[r], j = s_piv[cc];
        X[(size_t)i * N + j] = S[m][c];
        if (XT) XT[(size_t)j * N + i] = (float)S[m][c];
      }
    }
}

__device__ __forceinline__ double readlane_f64(double x, int l) {  // l uniform
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l), hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// (best, bi) <- the better of itself and the lane CTRL (a DPP lane pattern) names: larger
// |value|, ties to the lower row
template <int CTRL>
__device__ __forceinline__ void argmax_dpp_step(double& best, int& bi) {
  const uint64_t u = __builtin_bit_cast(uint64_t, best);
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)u, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(u >> 32), (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(bi, bi, CTRL, 0xF, 0xF, false);
  const double ob = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
  const bool take = ob > best || (ob == best && oi < bi);
  best = take ? ob : best;
  bi = take ? oi : bi;
}

// Wavefront arg-max of (best, bi) (larger value, ties to the lower index; the order is total,
// so any reduction tree gives the same winner): DPP within rows of 16 lanes (quad swaps, half-
// row and row mirrors), then the four row winners through readlane.  Returns the winner's
// index, uniform over the wavefront.
__device__ __forceinline__ int argmax_wave(double best, int bi) {
  argmax_dpp_step<0xB1>(best, bi);   // quad_perm [1,0,3,2]
  argmax_dpp_step<0x4E>(best, bi);   // quad_perm [2,3,0,1]
  argmax_dpp_step<0x141>(best, bi);  // row_half_mirror
  argmax_dpp_step<0x140>(best, bi);  // row_mirror
  double b = readlane_f64(best, 0);
  int i = __builtin_amdgcn_readlane(bi, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    const double ob = readlane_f64(best, 16 * r);
    const int oi = __builtin_amdgcn_readlane(bi, 16 * r);
    if (ob > b || (ob == b && oi < i)) {
      b = ob;
      i = oi;
    }
  }
  return i;
}

// Column-distributed register Gauss-Jordan inverse with implicit partial pivoting, FP64: the
// same pivot sequence and result layout as k_gj_inv_reg.  Column j lives in wavefront j % NW
// (register c = j / NW), row i in lane i % 64 (register m = i / 64).  The wavefront holding
// column k picks the unused row p with the largest |A[p][k]| (ties: lowest row) by a wavefront
// reduction and publishes p, the pivot and column k through LDS (double-buffered by step
// parity, so one barrier per step suffices); every wavefront then reads row p of its own
// columns with readlane and updates its elements in registers.
template <int NW, int RPT, int CPT>
__global__ void __launch_bounds__(NW * 64) k_gj_inv_col(const double* __restrict__ Aall, int N,
                                                        double* __restrict__ Ainv_all, float* __restrict__ AinvT32,
                                                        double* __restrict__ logdet, double* __restrict__ sign,
                                                        int want_inv) {
  __shared__ double s_col[2][RPT * 64];
  __shared__ double s_pv[2];
  __shared__ int s_p[2];
  __shared__ int s_piv[RPT * 64];
  __shared__ int s_inv[RPT * 64];
  __shared__ double s_pvk[RPT * 64];  // pivot of step k (log |det| summed after the loop)
  const int mat = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* Ag = Aall + (size_t)mat * N * N;
  double S[RPT][CPT];
#pragma unroll
  for (int m = 0; m < RPT; ++m)
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int i = lane + 64 * m, j = wid + NW * c;
      S[m][c] = (i < N && j < N) ? Ag[(size_t)i * N + j] : 0.0;
    }
  uint32_t used = 0;  // bit m: row lane + 64 m was a pivot
  double ld = 0.0;
  int neg = 0;
  for (int k = 0; k < N; ++k) {
    // buffer k & 1: a wavefront still reading it for step k - 2 has passed the barrier of k - 1
    const int buf = k & 1;
    if (wid == k % NW) {
      const int ck = k / NW;
      double v[RPT];
#pragma unroll
      for (int m = 0; m < RPT; ++m) {
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < CPT; ++c)
          if (c == ck) x = S[m][c];
        v[m] = x;
        s_col[buf][lane + 64 * m] = x;
      }
      double best = -1.0;
      int bi = 1 << 30;
#pragma unroll
      for (int m = 0; m < RPT; ++m) {
        const int i = lane + 64 * m;
        const double a = fabs(v[m]);
        if (i < N && !((used >> m) & 1u) && (a > best || (a == best && i < bi))) {
          best = a;
          bi = i;
        }
      }
      bi = argmax_wave(best, bi);
      if (lane == (bi & 63)) {
        double x = 0.0;
#pragma unroll
        for (int m = 0; m < RPT; ++m)
          if (m == (bi >> 6)) x = v[m];
        s_pv[buf] = x;
        s_p[buf] = bi;
      }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(s_p[buf]);
    const double pv = s_pv[buf];
    const double ip = 1.0 / pv;
    if (tid == 0) {
      neg += pv < 0.0;
      s_pvk[k] = pv;
      s_piv[k] = p;
    }
    const int mp = p >> 6, lp = p & 63;
    if (lane == lp) used |= 1u << mp;
    double fk[RPT];
#pragma unroll
    for (int m = 0; m < RPT; ++m) fk[m] = s_col[buf][lane + 64 * m];
    // every row takes S - fk rs (the pivot row is overwritten below); column k becomes -fk ip
    double rs[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      double x = S[0][c];
#pragma unroll
      for (int m = 1; m < RPT; ++m) x = m == mp ? S[m][c] : x;
      if (wid + NW * c == k) {  // uniform: column k
        rs[c] = ip;
#pragma unroll
        for (int m = 0; m < RPT; ++m) S[m][c] = -fk[m] * ip;
      } else {
        rs[c] = readlane_f64(x, lp) * ip;  // row p of this column, scaled
#pragma unroll
        for (int m = 0; m < RPT; ++m) S[m][c] = S[m][c] - fk[m] * rs[c];
      }
    }
    if (lane == lp) {
#pragma unroll
      for (int m = 0; m < RPT; ++m)
        if (m == mp) {
#pragma unroll
          for (int c = 0; c < CPT; ++c) S[m][c] = rs[c];
        }
    }
  }
  __syncthreads();
  for (int k = tid; k < N; k += NW * 64) {
    s_inv[s_piv[k]] = k;
    s_pvk[k] = log(fabs(s_pvk[k]));
  }
  __syncthreads();
  if (tid == 0) {
    for (int k = 0; k < N; ++k) ld += s_pvk[k];  // summed in step order
    // parity of k -> piv(k): N minus the number of cycles (visited marks kept in s_inv's sign bit)
    int cycles = 0;
    for (int s0 = 0; s0 < N; ++s0) {
      if (s_inv[s0] < 0) continue;
      ++cycles;
      for (int t = s0; s_inv[t] >= 0; t = s_piv[t]) s_inv[t] = -1 - s_inv[t];
    }
    for (int t = 0; t < N; ++t) s_inv[t] = -1 - s_inv[t];
    logdet[mat] = ld;
    sign[mat] = (((N - cycles) + neg) & 1) ? -1.0 : 1.0;
  }
  if (!want_inv) return;
  __syncthreads();
  double* X = Ainv_all + (size_t)mat * N * N;
  float* XT = AinvT32 ? AinvT32 + (size_t)mat * N * N : nullptr;
#pragma unroll
  for (int m = 0; m < RPT; ++m)
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int r = lane + 64 * m, cc = wid + NW * c;
      if (r < N && cc < N) {
        const int i = s_inv[r], j = s_piv[cc];
        X[(size_t)i * N + j] = S[m][c];
        if (XT) XT[(size_t)j * N + i] = (float)S[m][c];
      }
    }
}

size_t lu_lds_bytes(int N) { return (size_t)N * N * sizeof(double) + (size_t)N * (sizeof(double) + sizeof(int)); }

hipError_t lu_set_smem_limit(size_t bytes) {
  return hipFuncSetAttribute((const void*)k_gj_inv_lds, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

__global__ void k_transpose_f32(const double* __restrict__ X, int N, float* __restrict__ XT) {
  const size_t base = (size_t)blockIdx.x * N * N;
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int i = e / N, j = e - i * N;
    XT[base + (size_t)j * N + i] = (float)X[base + e];
  }
}

// Inverse, log|det| and sign of n_mat matrices.  LDS Gauss-Jordan when the matrix fits
// (use_lds), else LU in global memory followed by an FP32 transposed copy.
void launch_lu_inv(hipStream_t s, double* A, int n_mat, int N, double* Ainv, float* AinvT32, double* logdet,
                   double* sign, bool want_inv, bool use_lds) {
  if (N <= 144) {
    k_gj_inv_col<16, 3, 9><<<n_mat, 1024, 0, s>>>(A, N, Ainv, AinvT32, logdet, sign, want_inv ? 1 : 0);
  } else if (N <= 192) {
    k_gj_inv_col<16, 3, 12><<<n_mat, 1024, 0, s>>>(A, N, Ainv, AinvT32, logdet, sign, want_inv ? 1 : 0);
  } else if (use_lds) {
    k_gj_inv_lds<<<n_mat, 512, lu_lds_bytes(N), s>>>(A, N, Ainv, AinvT32, logdet, sign, want_inv ? 1 : 0);
  } else {
    k_lu_inv<<<n_mat, 256, N * sizeof(int), s>>>(A, N, Ainv, logdet, sign, want_inv ? 1 : 0);
    if (want_inv && AinvT32) k_transpose_f32<<<n_mat, 256, 0, s>>>(Ainv, N, AinvT32);
  }
}

// tr(Ainv * Lap A) for every (walker, det): sum_i sum_o lap_i[d,o] Ainv[o,i].
__global__ void k_det_trace(const float* __restrict__ O, int ldo, GroupDesc gd, int B, int N, int n_up, int D,
                            const double* __restrict__ Ainv, double* __restrict__ tr) {
  __shared__ double red[16];
  const int wd = blockIdx.x;
  const int w = wd / D, d = wd - w * D;
  const double* X = Ainv + (size_t)wd * N * N;
  double acc = 0.0;
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int i = e / N, o = e - i * N;
    const int g = group_of(w, i, B, n_up, N - n_up);
    acc += (double)O[(gd.row(g) + 1) * ldo + d * N + o] * X[(size_t)o * N + i];
  }
  double v[1] = {acc};
  block_sum<1>(v, red);
  if (threadIdx.x == 0) tr[wd] = v[0];
}

void launch_det_trace(hipStream_t s, const float* O, int ldo, GroupDesc gd, int B, int N, int n_up, int D,
                      const double* Ainv, double* tr) {
  k_det_trace<<<B * D, 256, 0, s>>>(O, ldo, gd, B, N, n_up, D, Ainv, tr);
}

// Triplet part of the determinant Laplacian and the determinant gradient.  For electron n with
// dependency set S = {n} U N(n) (s members) and the jacobian rows J_a[x] = d Phi_a / d r_n,x
// (3 per member a), per determinant d:
//   B_x[a,b] = sum_o J_a[x, d N + o] Ainv_d[o, S_b]   (FP32 MFMA 16x16x4)
//   trip_d = sum_x sum_{a,b} B_x[a,b] B_x[b,a],  grad_d,x = sum_a B_x[a,a]   (FP64 sums).
// Each jacobian row (i, n) belongs to exactly one set S_n, so the orbital jacobian is read once.
// For a != n the stored row is d O_a / d r_n before the envelope product; it is scaled by the
// envelope value row E_a (env, [group][5][ND], value row first) as it is loaded, which is the
// product rule since E_a depends on r_a only.
//
// One wavefront per (walker, electron, determinant); the four wavefronts of a workgroup take
// consecutive units, so they read the same jacobian rows (different determinant slices) on one
// CU.  Members are padded to NAT tiles of 16 (padding rows read member 0's rows, padding
// columns are zero), so B_x is an NAT x NAT grid of 16x16 tiles; XG values of x are computed
// per pass over the columns.  The jacobian and the inverse rows (AinvT, the transposed
// inverse) stream from memory straight into MFMA operand registers: a chunk covers 16 U
// columns starting at c0; lane l holds row l%16 and the columns c0 + 16 u + 4 (l/16) + t
// (u < U, t < 4), and the MFMA step (u, t) contracts exactly those columns.  That permutes the
// contraction index identically for both operands, so the product is unchanged while every
// lane issues contiguous vector loads.  The chunk frame starts at the slice start rounded down
// to 16 U columns (to 4 for the scalar form); the off = d N - start columns of the previous
// slice it takes in meet zero inverse columns.  With a row stride ldo that is a multiple of 32
// floats, a chunk is one aligned half (U = 1) or all (U = 2) of a 128-byte line of each row.
// trip and grad come from the accumulators: tile (j,i) is transposed through 1.25 KB of LDS per
// wavefront into the layout of tile (i,j), so no workgroup barrier is needed.
// VEC: float4 jacobian/envelope loads and float2 inverse loads (N even, ldo and ND multiples
// of 4); otherwise scalar loads.  Vector loads may run past the slice (into the next slice, or
// for the last slice into the row padding, which the caller keeps finite) onto zero columns.
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef float float2_t __attribute__((ext_vector_type(2)));

// vector form: 16-column runs per lane and chunk (two runs, whole 128-byte lines, measured
// slower: the registers cut the 2-tile class to 3 waves per SIMD); the 2-tile class: x values
// per column pass and waves per SIMD (its register budget)
#define DL_U 1
#define DL_XG2 3
#define DL_WPE2 4

template <int NAT, int XG, int P, bool VEC, int U>
__device__ __forceinline__ void det_lap_unit(const float* __restrict__ O, int ldo, int ND, GroupDesc gd,
                                             const float* __restrict__ env, const int32_t* __restrict__ nbr,
                                             const int32_t* __restrict__ rev, int B, int N, int n_up, int D,
                                             const float* __restrict__ AinvT, double* __restrict__ trip,
                                             double* __restrict__ grad, int w, int n, int d, int gn, int64_t p0,
                                             int s, float* tb) {
  constexpr int CW = 16 * U;
  const int lane = threadIdx.x & 63, li = lane & 15, kg = lane >> 4;
  const int cbase = VEC ? (d * N) & ~(CW - 1) : (d * N) & ~3, off = d * N - cbase;
  const int ncol = N + off, nch = (ncol + CW - 1) / CW;
  const float* jb[NAT];  // row x = 0 of member t*16 + li, at the chunk frame
  const float* eb[NAT];
  const float* xb[NAT];  // inverse row of member t*16 + li, at the chunk frame
  bool sc[NAT], bok[NAT];
#pragma unroll
  for (int t = 0; t < NAT; ++t) {
    const int m = t * 16 + li;
    const bool ok = m < s;
    int i = n;
    int64_t row = gd.row(gn) + 2, ep = -1;
    if (ok && m > 0) {
      i = nbr[p0 + m - 1];
      const int g = group_of(w, i, B, n_up, N - n_up);
      row = gd.row(g) + 5 + 3 * (int64_t)rev[p0 + m - 1];
      ep = (int64_t)g * 5 * ND;
    }
    jb[t] = O + row * ldo + cbase + kg * 4;
    sc[t] = ep >= 0;
    eb[t] = env + (ep >= 0 ? ep : 0) + cbase + kg * 4;
    bok[t] = ok;
    xb[t] = AinvT + ((size_t)(w * D + d) * N + i) * N + kg * 4 - off;
  }
  struct Chunk {
    float4_t j[XG][NAT][U], e[NAT][U], x[NAT][U];
  };
  double tr = 0.0, gx[3] = {0.0, 0.0, 0.0};
  for (int x0 = 0; x0 < 3; x0 += XG) {
    auto load = [&](int k, Chunk& c) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int o = k * CW + 16 * u;  // this lane's columns: o + 4 kg + t in the chunk frame
#pragma unroll
        for (int t = 0; t < NAT; ++t) {
          if (VEC) {
            const bool ok = o + kg * 4 < ncol;
#pragma unroll
            for (int xi = 0; xi < XG; ++xi)
              c.j[xi][t][u] = ok ? *reinterpret_cast<const float4_t*>(jb[t] + (int64_t)(x0 + xi) * ldo + o)
                                 : (float4_t)(0.f);
            c.e[t][u] = ok && sc[t] ? *reinterpret_cast<const float4_t*>(eb[t] + o) : (float4_t)(1.f);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int cs = o + kg * 4 - off + 2 * h;  // slice column (even: N and off are even)
              const float2_t v = bok[t] && cs >= 0 && cs < N
                                     ? *reinterpret_cast<const float2_t*>(xb[t] + o + 2 * h)
                                     : (float2_t)(0.f);
              c.x[t][u][2 * h] = v[0];
              c.x[t][u][2 * h + 1] = v[1];
            }
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int cf = o + kg * 4 + q, cs = cf - off;
              const bool ok = cf < ncol;
#pragma unroll
              for (int xi = 0; xi < XG; ++xi)
                c.j[xi][t][u][q] = ok ? jb[t][(int64_t)(x0 + xi) * ldo + o + q] : 0.f;
              c.e[t][u][q] = ok && sc[t] ? eb[t][o + q] : 1.f;
              c.x[t][u][q] = bok[t] && cs >= 0 && cs < N ? xb[t][o + q] : 0.f;
            }
          }
        }
      }
    };
    float4_t acc[XG][NAT][NAT];
#pragma unroll
    for (int xi = 0; xi < XG; ++xi)
#pragma unroll
      for (int i = 0; i < NAT; ++i)
#pragma unroll
        for (int j = 0; j < NAT; ++j) acc[xi][i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    auto mul = [&](const Chunk& c) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int xi = 0; xi < XG; ++xi)
#pragma unroll
            for (int i = 0; i < NAT; ++i) {
              const float jv = c.j[xi][i][u][t] * c.e[i][u][t];
#pragma unroll
              for (int j = 0; j < NAT; ++j)
                acc[xi][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(jv, c.x[j][u][t], acc[xi][i][j], 0, 0, 0);
            }
    };
    Chunk c[P];
#pragma unroll
    for (int q = 0; q < P; ++q)
      if (q < nch) load(q, c[q]);
    for (int k = 0; k < nch; k += P) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (k + q < nch) {
          mul(c[q]);
          if (k + q + P < nch) load(k + q + P, c[q]);
        }
      }
    }
    // trip_x = sum_{i,j} T_ij o T_ji^T over tiles, grad_x = trace; padding rows meet zero columns
#pragma unroll
    for (int xi = 0; xi < XG; ++xi) {
      double g = 0.0;
#pragma unroll
      for (int i = 0; i < NAT; ++i) {
#pragma unroll
        for (int j = i; j < NAT; ++j) {
#pragma unroll
          for (int v = 0; v < 4; ++v) tb[(4 * kg + v) * 20 + li] = acc[xi][j][i][v];
          __builtin_amdgcn_wave_barrier();
          const float4_t tt = *reinterpret_cast<const float4_t*>(tb + li * 20 + 4 * kg);
          __builtin_amdgcn_wave_barrier();
          double p = 0.0;
#pragma unroll
          for (int v = 0; v < 4; ++v) p += (double)acc[xi][i][j][v] * (double)tt[v];
          tr += (i == j ? 1.0 : 2.0) * p;
          if (i == j) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (li == 4 * kg + v) g += (double)acc[xi][i][i][v];
          }
        }
      }
      const int x = x0 + xi;
      if (x == 0) gx[0] += g;
      else if (x == 1) gx[1] += g;
      else gx[2] += g;
    }
  }
  tr = warp_sum(tr);
#pragma unroll
  for (int x = 0; x < 3; ++x) gx[x] = warp_sum(gx[x]);
  if (lane == 0) {
    const int64_t o = ((int64_t)w * D + d) * N + n;
    trip[o] = tr;
#pragma unroll
    for (int x = 0; x < 3; ++x) grad[o * 3 + x] = gx[x];
  }
}

// Units of NAT = ceil(s / 16) tiles only (one launch per tile count, so each has its own
// register budget); the other wavefronts return at once.  Workgroups are mapped so that each
// XCD takes a contiguous range of units and keeps its walkers' inverses in its own L2.
template <int NAT, int XG, int P, bool VEC, int U, int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
k_det_lap(const float* __restrict__ O, int ldo, int ND, GroupDesc gd, const float* __restrict__ env,
          const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr, const int32_t* __restrict__ rev, int B,
          int N, int n_up, int D, const float* __restrict__ AinvT, double* __restrict__ trip,
          double* __restrict__ grad) {
  __shared__ float tbs[4][16 * 20];
  const int wid = threadIdx.x >> 6;
  const int64_t units = (int64_t)B * N * D;
  const int64_t nblk = (units + 3) / 4, G8 = (nblk + 7) >> 3;
  const int64_t lb = (int64_t)(blockIdx.x & 7) * G8 + (blockIdx.x >> 3);
  const int64_t u = lb * 4 + wid;
  if (lb >= nblk || u >= units) return;
  const int d = (int)(u % D);
  const int wn = (int)(u / D);
  const int w = wn / N, n = wn - w * N;
  const int gn = group_of(w, n, B, n_up, N - n_up);
  const int64_t p0 = ptr[gn];
  const int s = 1 + (int)(ptr[gn + 1] - p0);
  if (((s + 15) >> 4) != NAT) return;
  det_lap_unit<NAT, XG, P, VEC, U>(O, ldo, ND, gd, env, nbr, rev, B, N, n_up, D, AinvT, trip, grad, w, n, d, gn, p0,
                                   s, tbs[wid]);
}

template <bool VEC, int U>
static void det_lap_launch(hipStream_t st, const float* O, int ldo, int ND, GroupDesc gd, const float* env,
                           const int64_t* ptr, const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, int D,
                           int smax, const float* AinvT, double* trip, double* grad) {
  const int64_t nblk = ((int64_t)B * N * D + 3) / 4;
  const dim3 grid((unsigned)(8 * ((nblk + 7) / 8)));
  const int natm = (smax + 15) / 16;
  k_det_lap<1, 3, 2, VEC, U, 1><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT, trip,
                                                      grad);
  if (natm >= 2)
    k_det_lap<2, DL_XG2, 1, VEC, U, DL_WPE2><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D,
                                                                  AinvT, trip, grad);
  // the 3- and 4-tile classes (rare) keep 16-column runs: their accumulators fill the registers
  if (natm >= 3)
    k_det_lap<3, 3, 1, VEC, 1, 1><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT,
                                                        trip, grad);
  if (natm >= 4)
    k_det_lap<4, 1, 1, VEC, 1, 1><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT,
                                                        trip, grad);
}

void launch_det_lap(hipStream_t s, const float* O, int ldo, int ND, GroupDesc gd, const float* env,
                    const int64_t* ptr, const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, int D,
                    int smax, const float* AinvT, double* trip, double* grad) {
  if (B * N == 0) return;
  if (N % 2 == 0 && ND % 4 == 0 && ldo % 32 == 0)
    det_lap_launch<true, DL_U>(s, O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, smax, AinvT, trip, grad);
  else
    det_lap_launch<false, 1>(s, O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, smax, AinvT, trip, grad);
}

// FP64 form of the triplet kernel: the same units, chunking and column permutation, with the
// jacobian rows (FP32, times the envelope scale) widened to FP64 and multiplied with the FP64
// transposed inverse on v_mfma_f64_16x16x4f64, FP64 accumulators.  Lane l holds A[l%16][l/16]
// and B[l/16][l%16]; register v of lane l holds D[4v + l/16][l%16], so the tile transposes
// through LDS in that layout.  Vector loads only (N even, ldo and ND multiples of 4); the
// inverse rows are read as double2.
typedef double dbl4_t __attribute__((ext_vector_type(4)));
typedef double dbl2_t __attribute__((ext_vector_type(2)));

template <int NAT, int XG>
__device__ __forceinline__ void det_lap_unit64(const float* __restrict__ O, int ldo, int ND, GroupDesc gd,
                                               const float* __restrict__ env, const int32_t* __restrict__ nbr,
                                               const int32_t* __restrict__ rev, int B, int N, int n_up, int D,
                                               const double* __restrict__ AinvT, double* __restrict__ trip,
                                               double* __restrict__ grad, int w, int n, int d, int gn, int64_t p0,
                                               int s, double* tb) {
  const int lane = threadIdx.x & 63, li = lane & 15, kg = lane >> 4;
  const int cbase = (d * N) & ~15, off = d * N - cbase;
  const int ncol = N + off, nch = (ncol + 15) / 16;
  const float* jb[NAT];
  const float* eb[NAT];
  const double* xb[NAT];
  bool sc[NAT], bok[NAT];
#pragma unroll
  for (int t = 0; t < NAT; ++t) {
    const int m = t * 16 + li;
    const bool ok = m < s;
    int i = n;
    int64_t row = gd.row(gn) + 2, ep = -1;
    if (ok && m > 0) {
      i = nbr[p0 + m - 1];
      const int g = group_of(w, i, B, n_up, N - n_up);
      row = gd.row(g) + 5 + 3 * (int64_t)rev[p0 + m - 1];
      ep = (int64_t)g * 5 * ND;
    }
    jb[t] = O + row * ldo + cbase + kg * 4;
    sc[t] = ep >= 0;
    eb[t] = env + (ep >= 0 ? ep : 0) + cbase + kg * 4;
    bok[t] = ok;
    xb[t] = AinvT + ((size_t)(w * D + d) * N + i) * N + kg * 4 - off;
  }
  double tr = 0.0, gx[3] = {0.0, 0.0, 0.0};
  for (int x0 = 0; x0 < 3; x0 += XG) {
    dbl4_t acc[XG][NAT][NAT];
#pragma unroll
    for (int xi = 0; xi < XG; ++xi)
#pragma unroll
      for (int i = 0; i < NAT; ++i)
#pragma unroll
        for (int j = 0; j < NAT; ++j) acc[xi][i][j] = (dbl4_t){0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < nch; ++k) {
      const int o = k * 16;
      const bool okc = o + kg * 4 < ncol;
      float4_t jv[XG][NAT], ev[NAT];
      double xv[NAT][4];
#pragma unroll
      for (int t = 0; t < NAT; ++t) {
#pragma unroll
        for (int xi = 0; xi < XG; ++xi)
          jv[xi][t] = okc ? *reinterpret_cast<const float4_t*>(jb[t] + (int64_t)(x0 + xi) * ldo + o) : (float4_t)(0.f);
        ev[t] = okc && sc[t] ? *reinterpret_cast<const float4_t*>(eb[t] + o) : (float4_t)(1.f);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int cs = o + kg * 4 - off + 2 * h;
          const dbl2_t v = bok[t] && cs >= 0 && cs < N ? *reinterpret_cast<const dbl2_t*>(xb[t] + o + 2 * h)
                                                       : (dbl2_t)(0.0);
          xv[t][2 * h] = v[0];
          xv[t][2 * h + 1] = v[1];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int xi = 0; xi < XG; ++xi)
#pragma unroll
          for (int i = 0; i < NAT; ++i) {
            const double a = (double)jv[xi][i][q] * (double)ev[i][q];
#pragma unroll
            for (int j = 0; j < NAT; ++j)
              acc[xi][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xv[j][q], acc[xi][i][j], 0, 0, 0);
          }
    }
#pragma unroll
    for (int xi = 0; xi < XG; ++xi) {
      double g = 0.0;
#pragma unroll
      for (int i = 0; i < NAT; ++i) {
#pragma unroll
        for (int j = i; j < NAT; ++j) {
          // tile (j,i) into LDS as [row][col], then this lane's element (4v + kg, li) of tile
          // (i,j) meets element (li, 4v + kg) of tile (j,i)
#pragma unroll
          for (int v = 0; v < 4; ++v) tb[(4 * v + kg) * 17 + li] = acc[xi][j][i][v];
          __builtin_amdgcn_wave_barrier();
          double p = 0.0;
#pragma unroll
          for (int v = 0; v < 4; ++v) p += acc[xi][i][j][v] * tb[li * 17 + 4 * v + kg];
          __builtin_amdgcn_wave_barrier();
          tr += (i == j ? 1.0 : 2.0) * p;
          if (i == j) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (li == 4 * v + kg) g += acc[xi][i][i][v];
          }
        }
      }
      const int x = x0 + xi;
      if (x == 0) gx[0] += g;
      else if (x == 1) gx[1] += g;
      else gx[2] += g;
    }
  }
  tr = warp_sum(tr);
#pragma unroll
  for (int x = 0; x < 3; ++x) gx[x] = warp_sum(gx[x]);
  if (lane == 0) {
    const int64_t o = ((int64_t)w * D + d) * N + n;
    trip[o] = tr;
#pragma unroll
    for (int x = 0; x < 3; ++x) grad[o * 3 + x] = gx[x];
  }
}

template <int NAT, int XG, int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
k_det_lap64(const float* __restrict__ O, int ldo, int ND, GroupDesc gd, const float* __restrict__ env,
            const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr, const int32_t* __restrict__ rev, int B,
            int N, int n_up, int D, const double* __restrict__ AinvT, double* __restrict__ trip,
            double* __restrict__ grad) {
  __shared__ double tbs[4][16 * 17];
  const int wid = threadIdx.x >> 6;
  const int64_t units = (int64_t)B * N * D;
  const int64_t nblk = (units + 3) / 4, G8 = (nblk + 7) >> 3;
  const int64_t lb = (int64_t)(blockIdx.x & 7) * G8 + (blockIdx.x >> 3);
  const int64_t u = lb * 4 + wid;
  if (lb >= nblk || u >= units) return;
  const int d = (int)(u % D);
  const int wn = (int)(u / D);
  const int w = wn / N, n = wn - w * N;
  const int gn = group_of(w, n, B, n_up, N - n_up);
  const int64_t p0 = ptr[gn];
  const int s = 1 + (int)(ptr[gn + 1] - p0);
  if (((s + 15) >> 4) != NAT) return;
  det_lap_unit64<NAT, XG>(O, ldo, ND, gd, env, nbr, rev, B, N, n_up, D, AinvT, trip, grad, w, n, d, gn, p0, s,
                          tbs[wid]);
}

// FP64 transposed copy of n_mat N x N matrices (32 x 32 tiles through LDS).
__global__ void k_transpose_f64(const double* __restrict__ X, int N, double* __restrict__ XT) {
  __shared__ double t[32][33];
  const size_t base = (size_t)blockIdx.z * N * N;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int r = by + j, c = bx + threadIdx.x;
    if (r < N && c < N) t[j][threadIdx.x] = X[base + (size_t)r * N + c];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int r = bx + j, c = by + threadIdx.x;
    if (r < N && c < N) XT[base + (size_t)r * N + c] = t[threadIdx.x][j];
  }
}

void launch_transpose_f64(hipStream_t s, const double* X, int n_mat, int N, double* XT) {
  if (n_mat == 0) return;
  const dim3 grid((N + 31) / 32, (N + 31) / 32, n_mat);
  k_transpose_f64<<<grid, dim3(32, 8), 0, s>>>(X, N, XT);
}

bool det_lap64_supported(int N, int ND, int ldo) { return N % 2 == 0 && ND % 4 == 0 && ldo % 32 == 0; }

void launch_det_lap64(hipStream_t st, const float* O, int ldo, int ND, GroupDesc gd, const float* env,
                      const int64_t* ptr, const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, int D,
                      int smax, const double* AinvT, double* trip, double* grad) {
  if (B * N == 0) return;
  const int64_t nblk = ((int64_t)B * N * D + 3) / 4;
  const dim3 grid((unsigned)(8 * ((nblk + 7) / 8)));
  const int natm = (smax + 15) / 16;
  k_det_lap64<1, 3, 2><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT, trip, grad);
  if (natm >= 2)
    k_det_lap64<2, 1, 2><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT, trip, grad);
  if (natm >= 3)
    k_det_lap64<3, 1, 1><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT, trip, grad);
  if (natm >= 4)
    k_det_lap64<4, 1, 1><<<grid, 256, 0, st>>>(O, ldo, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, AinvT, trip, grad);
}

// Sum of the determinants: log|psi_det|, sign, gradient and Laplacian per walker (FP64).
//   w_d = s_d exp(L_d - m) / sum;  grad = sum_d w_d g_d;
//   lap = sum_d w_d (lap_d + |g_d|^2) - |grad|^2,  lap_d = tr_d - sum_n trip_{d,n}.
__global__ void k_det_combine(int N, int D, const double* __restrict__ logdet, const double* __restrict__ sgn,
                              const double* __restrict__ tr, const double* __restrict__ trip,
                              const double* __restrict__ gd_, double* __restrict__ out_logpsi,
                              double* __restrict__ out_sign, double* __restrict__ out_grad, double* __restrict__ out_lap,
                              int tangents) {
  __shared__ double s_w[16];
  __shared__ double red[16 * 2];
  const int w = blockIdx.x;
  if (threadIdx.x == 0) {
    double m = -1e300;
    for (int d = 0; d < D; ++d) m = fmax(m, logdet[w * D + d]);
    double sum = 0.0;
    for (int d = 0; d < D; ++d) sum += sgn[w * D + d] * exp(logdet[w * D + d] - m);
    for (int d = 0; d < D; ++d) s_w[d] = sgn[w * D + d] * exp(logdet[w * D + d] - m) / sum;
    out_logpsi[w] = m + log(fabs(sum));
    out_sign[w] = sum < 0 ? -1.0 : 1.0;
  }
  __syncthreads();
  if (!tangents) return;
  double v[2] = {0.0, 0.0};  // sum_d w_d (tr_d - trip_d + |g_d|^2),  |grad|^2
  for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) {
    double gsum = 0.0;
    for (int d = 0; d < D; ++d) {
      const double gv = gd_[((size_t)(w * D + d) * N) * 3 + e];
      gsum += s_w[d] * gv;
      v[0] += s_w[d] * gv * gv;
    }
    out_grad[(size_t)w * 3 * N + e] = gsum;
    v[1] += gsum * gsum;
  }
  for (int e = threadIdx.x; e < D * N; e += blockDim.x) {
    const int d = e / N;
    v[0] -= s_w[d] * trip[(size_t)(w * D + d) * N + (e - d * N)];
  }
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int d = 0; d < D; ++d) t += s_w[d] * tr[w * D + d];
    out_lap[w] = t + v[0] - v[1];
  }
}

void launch_det_combine(hipStream_t s, int B, int N, int D, const double* logdet, const double* sgn, const double* tr,
                        const double* trip, const double* grad, double* out_logpsi, double* out_sign,
                        double* out_grad, double* out_lap, bool tangents) {
  k_det_combine<<<B, 256, 0, s>>>(N, D, logdet, sgn, tr, trip, grad, out_logpsi, out_sign, out_grad, out_lap,
                                  tangents ? 1 : 0);
}

}  // namespace fire
