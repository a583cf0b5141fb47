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
__global__ void k_orb_product(float* __restrict__ O, int ND, GroupDesc gd, const float* __restrict__ env) {
  const int g = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= ND) return;
  const int NR = TAN ? 5 : 1;
  const float* e = env + (size_t)g * NR * ND + c;
  float* o = O + gd.row(g) * ND + c;
  const float ox = o[0], ex = e[0];
  o[0] = ox * ex;
  if (!TAN) return;
  const float ol = o[ND], el = e[ND];
  float cross = 0.f;
  for (int t = 0; t < 3; ++t) {
    const float ot = o[(2 + t) * ND], et = e[(2 + t) * ND];
    cross += ot * et;
    o[(2 + t) * (int64_t)ND] = ox * et + ex * ot;
  }
  o[ND] = ol * ex + el * ox + 2.f * cross;
}

void launch_orb_product(hipStream_t s, float* O, int ND, int G, GroupDesc gd, const float* env, bool tangents) {
  dim3 grid(G, (ND + 255) / 256);
  if (tangents)
    k_orb_product<true><<<grid, 256, 0, s>>>(O, ND, gd, env);
  else
    k_orb_product<false><<<grid, 256, 0, s>>>(O, ND, gd, env);
}

// A[w,d][i][o] (FP64) from the value rows of the orbital groups.
__global__ void k_gather_A(const float* __restrict__ O, int ND, GroupDesc gd, int B, int N, int n_up, int D,
                           double* __restrict__ A) {
  const int g = blockIdx.x;
  int w, i;
  group_inv(g, B, n_up, N - n_up, w, i);
  const float* o = O + gd.row(g) * ND;
  for (int c = threadIdx.x; c < ND; c += blockDim.x) {
    const int d = c / N, col = c - d * N;
    A[(((size_t)w * D + d) * N + i) * N + col] = (double)o[c];
  }
}

void launch_gather_A(hipStream_t s, const float* O, int ND, GroupDesc gd, int B, int N, int n_up, int D, double* A) {
  k_gather_A<<<B * N, 256, 0, s>>>(O, ND, gd, B, N, n_up, D, A);
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
    k_gj_inv_reg<16, 9, 3><<<n_mat, 1024, 0, s>>>(A, N, Ainv, AinvT32, logdet, sign, want_inv ? 1 : 0);
  } else if (N <= 192) {
    k_gj_inv_reg<16, 12, 3><<<n_mat, 1024, 0, s>>>(A, N, Ainv, AinvT32, logdet, sign, want_inv ? 1 : 0);
  } else if (use_lds) {
    k_gj_inv_lds<<<n_mat, 512, lu_lds_bytes(N), s>>>(A, N, Ainv, AinvT32, logdet, sign, want_inv ? 1 : 0);
  } else {
    k_lu_inv<<<n_mat, 256, N * sizeof(int), s>>>(A, N, Ainv, logdet, sign, want_inv ? 1 : 0);
    if (want_inv && AinvT32) k_transpose_f32<<<n_mat, 256, 0, s>>>(Ainv, N, AinvT32);
  }
}

// tr(Ainv * Lap A) for every (walker, det): sum_i sum_o lap_i[d,o] Ainv[o,i].
__global__ void k_det_trace(const float* __restrict__ O, int ND, GroupDesc gd, int B, int N, int n_up, int D,
                            const double* __restrict__ Ainv, double* __restrict__ tr) {
  __shared__ double red[16];
  const int wd = blockIdx.x;
  const int w = wd / D, d = wd - w * D;
  const double* X = Ainv + (size_t)wd * N * N;
  double acc = 0.0;
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int i = e / N, o = e - i * N;
    const int g = group_of(w, i, B, n_up, N - n_up);
    acc += (double)O[(gd.row(g) + 1) * ND + d * N + o] * X[(size_t)o * N + i];
  }
  double v[1] = {acc};
  block_sum<1>(v, red);
  if (threadIdx.x == 0) tr[wd] = v[0];
}

void launch_det_trace(hipStream_t s, const float* O, int ND, GroupDesc gd, int B, int N, int n_up, int D,
                      const double* Ainv, double* tr) {
  k_det_trace<<<B * D, 256, 0, s>>>(O, ND, gd, B, N, n_up, D, Ainv, tr);
}

#define DL_KC 32  // contraction chunk staged in LDS
#define DL_TPW 12  // max 16x16 output tiles per wavefront (3*64 x 64 outputs / 4 waves)

typedef float float4_t __attribute__((ext_vector_type(4)));

// Triplet part of the determinant Laplacian and the determinant gradient for one
// (walker, det, electron n): with S = {n} U N(n) and J_a the jacobian rows d Phi_a / d r_n,
//   B[a,x,b] = sum_o J_a[x,o] Ainv[o, S_b]       (FP32 MFMA 16x16x4, Ainv rounded from FP64)
//   trip = sum B[a,x,b] B[b,x,a],  grad_x = sum_a B[a,x,a]   (FP64 sums).
// Each jacobian row (i, n) belongs to exactly one set S_n, so the orbital jacobian is read
// once.  For a != n the stored row is d O_a / d r_n before the envelope product; it is
// scaled by the envelope value row E_a (env, [group][5][ND], value row first) while it is
// staged, which is the product rule since E_a depends on r_a only.
// Dynamic LDS: J chunk [RT*16][KC+4], X chunk [CT*16][KC+4], B [3s][s], indices.
// VW: width of the vector loads of jacobian and inverse rows (4, 2 or 1 floats; N % VW == 0
// keeps every row start aligned).  The next orbital chunk is prefetched into registers while
// the current one is multiplied.
template <int VW>
__global__ void __launch_bounds__(256) k_det_lap(const float* __restrict__ O, int ND, GroupDesc gd,
                                                 const float* __restrict__ env, const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                                                 const int32_t* __restrict__ rev, int B, int N, int n_up, int D,
                                                 int smax, const float* __restrict__ AinvT, double* __restrict__ trip,
                                                 double* __restrict__ grad) {
  extern __shared__ unsigned char smem_raw[];
  const int RTm = (3 * smax + 15) / 16, CTm = (smax + 15) / 16;
  const int LDK = DL_KC + 4;
  float* Js = reinterpret_cast<float*>(smem_raw);
  float* Xs = Js + RTm * 16 * LDK;
  float* Bs = Xs + CTm * 16 * LDK;
  int* s_idx = reinterpret_cast<int*>(Bs + ((3 * smax * smax + 1) & ~1));
  int64_t* s_row = reinterpret_cast<int64_t*>(s_idx + smax + (smax & 1));  // first jacobian row of a
  int64_t* s_rp = s_row + smax;                                             // element offset of row (a, x)
  int64_t* s_ep = s_rp + 3 * smax;  // envelope value row offset of a (-1 for a = 0)
  __shared__ double red[16 * 4];
  const int blk = blockIdx.x;
  const int n = blk % N;
  const int wd = blk / N;
  const int w = wd / D, d = wd - w * D;
  const int gn = group_of(w, n, B, n_up, N - n_up);
  const int64_t p0 = ptr[gn];
  const int s = 1 + (int)(ptr[gn + 1] - p0);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int a = tid; a < s; a += blockDim.x) {
    if (a == 0) {
      s_idx[0] = n;
      s_row[0] = gd.row(gn) + 2;
      s_ep[0] = -1;
    } else {
      const int i = nbr[p0 + a - 1];
      const int gi = group_of(w, i, B, n_up, N - n_up);
      s_idx[a] = i;
      s_row[a] = gd.row(gi) + 5 + 3 * (int64_t)rev[p0 + a - 1];
      s_ep[a] = (int64_t)gi * 5 * ND + (int64_t)d * N;
    }
  }
  const int RT = (3 * s + 15) / 16, CT = (s + 15) / 16, NT = RT * CT;
  float4_t acc[DL_TPW];
  for (int t = 0; t < DL_TPW; ++t) acc[t] = (float4_t){0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  for (int ax = tid; ax < 3 * s; ax += blockDim.x) {
    const int a = ax / 3, x = ax - 3 * a;
    s_rp[ax] = (s_row[a] + x) * ND + (int64_t)d * N;
  }
  __syncthreads();
  const float* XT = AinvT + (size_t)wd * N * N;
  typedef float vec_t __attribute__((ext_vector_type(VW)));
  constexpr int UPR = DL_KC / VW;                                   // vector units per chunk row
  constexpr int JPT = ((FIRE_SMAX * 3 + 15) / 16 * 16 * UPR + 255) / 256;  // J units per thread
  constexpr int XPT = ((FIRE_SMAX + 15) / 16 * 16 * UPR + 255) / 256;      // X units per thread
  const int nJ = RT * 16 * UPR, nX = CT * 16 * UPR;
  vec_t rj[JPT], re[JPT], rx[XPT];
  auto load = [&](int o0) {
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
      const int e = tid + q * 256;
      const int ax = e / UPR, k = (e % UPR) * VW;
      vec_t v = (vec_t)(0.f), ev = (vec_t)(1.f);
      if (e < nJ && ax < 3 * s && o0 + k < N) {
        v = *reinterpret_cast<const vec_t*>(O + s_rp[ax] + o0 + k);
        const int64_t ep = s_ep[ax / 3];
        if (ep >= 0) ev = *reinterpret_cast<const vec_t*>(env + ep + o0 + k);
      }
      rj[q] = v;
      re[q] = ev;
    }
#pragma unroll
    for (int q = 0; q < XPT; ++q) {
      const int e = tid + q * 256;
      const int b = e / UPR, k = (e % UPR) * VW;
      vec_t v = (vec_t)(0.f);
      if (e < nX && b < s && o0 + k < N) v = *reinterpret_cast<const vec_t*>(XT + (size_t)s_idx[b] * N + o0 + k);
      rx[q] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
      const int e = tid + q * 256;
      if (e < nJ) *reinterpret_cast<vec_t*>(Js + (e / UPR) * LDK + (e % UPR) * VW) = rj[q] * re[q];
    }
#pragma unroll
    for (int q = 0; q < XPT; ++q) {
      const int e = tid + q * 256;
      if (e < nX) *reinterpret_cast<vec_t*>(Xs + (e / UPR) * LDK + (e % UPR) * VW) = rx[q];
    }
  };
  load(0);
  for (int o0 = 0; o0 < N; o0 += DL_KC) {
    store();
    __syncthreads();
    if (o0 + DL_KC < N) load(o0 + DL_KC);
#pragma unroll
    for (int q = 0; q < DL_TPW; ++q) {
      const int t = wid + 4 * q;
      if (t < NT) {
        const int rt = t / CT, ct = t - rt * CT;
        const float* ja = Js + (rt * 16 + (lane & 15)) * LDK + (lane >> 4);
        const float* xb = Xs + (ct * 16 + (lane & 15)) * LDK + (lane >> 4);
#pragma unroll
        for (int kk = 0; kk < DL_KC; kk += 4)
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ja[kk], xb[kk], acc[q], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // scatter the 16x16 tiles into B[ax][b]
  {
#pragma unroll
    for (int q = 0; q < DL_TPW; ++q) {
      const int t = wid + 4 * q;
      if (t < NT) {
        const int rt = t / CT, ct = t - rt * CT;
        const int b = ct * 16 + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ax = rt * 16 + 4 * (lane >> 4) + v;
          if (ax < 3 * s && b < s) Bs[ax * s + b] = acc[q][v];
        }
      }
    }
  }
  __syncthreads();
  double v4[4] = {0.0, 0.0, 0.0, 0.0};
  for (int e = tid; e < 3 * s * s; e += blockDim.x) {
    const int ax = e / s, b = e - ax * s;
    const int a = ax / 3, x = ax - 3 * a;
    const double bv = Bs[e];
    v4[0] += bv * (double)Bs[(3 * b + x) * s + a];
    if (a == b) v4[1 + x] += bv;
  }
  block_sum<4>(v4, red);
  if (tid == 0) {
    trip[blk] = v4[0];
    for (int x = 0; x < 3; ++x) grad[(size_t)blk * 3 + x] = v4[1 + x];
  }
}

size_t det_lap_smem(int smax) {
  const int RTm = (3 * smax + 15) / 16, CTm = (smax + 15) / 16;
  return (size_t)(RTm * 16 + CTm * 16) * (DL_KC + 4) * 4 + (size_t)((3 * smax * smax + 1) & ~1) * 4 +
         (size_t)(smax + (smax & 1)) * 4 + (size_t)smax * 8 + (size_t)3 * smax * 8 + (size_t)smax * 8;
}

void launch_det_lap(hipStream_t s, const float* O, int ND, GroupDesc gd, const float* env, const int64_t* ptr,
                    const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, int D, int smax,
                    const float* AinvT, double* trip, double* grad) {
  const size_t sm = det_lap_smem(smax);
  if (N % 4 == 0)
    k_det_lap<4><<<B * D * N, 256, sm, s>>>(O, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, smax, AinvT, trip, grad);
  else if (N % 2 == 0)
    k_det_lap<2><<<B * D * N, 256, sm, s>>>(O, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, smax, AinvT, trip, grad);
  else
    k_det_lap<1><<<B * D * N, 256, sm, s>>>(O, ND, gd, env, ptr, nbr, rev, B, N, n_up, D, smax, AinvT, trip, grad);
}

hipError_t det_lap_set_smem_limit(size_t bytes) {
  hipError_t e = hipFuncSetAttribute((const void*)k_det_lap<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_det_lap<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_det_lap<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  return e;
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
