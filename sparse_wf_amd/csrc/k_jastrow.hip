// Jastrow factor, Coulomb potential and local-energy assembly.
//
// References (sparse_wf): Jastrow.apply_with_fwd_lap jastrow.py:387-459 (MLP Jastrow
// SparseMLP :216 with silu after the last layer, _mlp_to_logpsi :372-382), GlobalAttentionJastrow
// :67-155, ElElCusp :48-64, NodeWithFwdLap.sum_over_nodes / sum_from_to sparse_fwd_lap.py:38-56,
// potential_energy hamiltonian.py:15-26, kinetic energy wave_function.py:215-221.
#include "fire_common.hpp"
#include "fire_kernels2.hpp"

namespace fire {

// Last MLP-Jastrow layer (W -> 2), silu with forward-Laplacian, * mlp_scale + [0, log_bias].
// Z holds the pre-activation rows of the layer before: its silu (forward-Laplacian rule) is
// applied here rather than in a separate pass.  With f, f1, f2 = silu and its derivatives at
// the group's value row (per column k), the activated rows are f (value), f1 z_jac (each
// jacobian row) and f1 z_lap + f2 sum_jac z_jac^2 (Laplacian row), so the row dot products
// with W3 are
//   value: sum f W3,  jacobian: sum z (f1 W3),  Laplacian: sum (f1 z_lap + f2 ss) W3,
// with ss[k] the column sums of squares of the jacobian rows (partials per row slot in LDS,
// added in a fixed order).  JR gets two columns per tangent row of the group.  The row dot
// products go to LDS (rows past 5 + 3 (FIRE_SMAX - 1), which the neighbour limit excludes, to
// JR itself); the activation rule of the last layer then runs over the rows in parallel (128
// threads per column).  Dynamic LDS: 20 W floats.
#define JM_MAXR (5 + 3 * (FIRE_SMAX - 1))
__global__ void __launch_bounds__(256) k_jmlp_last(const float* __restrict__ Z, int W, GroupDesc gd,
                                                   const float* __restrict__ W3, const float* __restrict__ b3,
                                                   const float* __restrict__ scale, const float* __restrict__ zb,
                                                   float* __restrict__ JR, const int32_t* __restrict__ dG) {
  extern __shared__ float jm_lds[];
  float* s_f1 = jm_lds;         // [W]
  float* s_f2 = s_f1 + W;       // [W]
  float* s_v0 = s_f2 + W;       // f1 * W3[:, 0]
  float* s_v1 = s_v0 + W;       // f1 * W3[:, 1]
  float* s_ss = s_v1 + W;       // [16][W] partial sums of squares
  __shared__ float s_a[2][JM_MAXR];
  __shared__ float s_red[2][4];
  __shared__ float s_ss2[2][2];
  const int g = blockIdx.x;
  if (dG && g >= *dG) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t r0 = gd.row(g);
  const int nr = (int)(gd.row(g + 1) - r0);
  auto put = [&](int rr, float a0, float a1) {
    if (rr == 0) { a0 += b3[0]; a1 += b3[1]; }
    if (rr < JM_MAXR) {
      s_a[0][rr] = a0;
      s_a[1][rr] = a1;
    } else {
      JR[(r0 + rr) * 2] = a0;
      JR[(r0 + rr) * 2 + 1] = a1;
    }
  };
  // block sum of two values; every thread gets them
  auto bsum2 = [&](float& a0, float& a1) {
    a0 = warp_sum(a0);
    a1 = warp_sum(a1);
    if (lane == 0) { s_red[0][wid] = a0; s_red[1][wid] = a1; }
    __syncthreads();
    a0 = s_red[0][0] + s_red[0][1] + s_red[0][2] + s_red[0][3];
    a1 = s_red[1][0] + s_red[1][1] + s_red[1][2] + s_red[1][3];
    __syncthreads();
  };
  // value row: activation coefficients per column and the value-row dot products
  {
    float a0 = 0.f, a1 = 0.f;
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
      float f, f1, f2;
      silu3(Z[r0 * W + k], f, f1, f2);
      const float w0 = W3[2 * k], w1 = W3[2 * k + 1];
      s_f1[k] = f1;
      s_f2[k] = f2;
      s_v0[k] = f1 * w0;
      s_v1[k] = f1 * w1;
      a0 += f * w0;
      a1 += f * w1;
    }
    bsum2(a0, a1);  // (its barriers also publish the coefficients)
    if (threadIdx.x == 0) put(0, a0, a1);
  }
  if (nr > 1) {
    // jacobian rows: dots with f1 W3, and per-column sums of squares in 16 partial slots
    if (W % 64 == 0) {
      // four rows per wavefront, 16 lanes per row with float4 loads
      const int sub = lane >> 4, l16 = lane & 15, slot = wid * 4 + sub;
      for (int k = 4 * l16; k < W; k += 64)
        *reinterpret_cast<float4*>(s_ss + slot * W + k) = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int rb = 2 + wid * 4; rb < nr; rb += (blockDim.x >> 6) * 4) {
        const int rr = rb + sub;
        float a0 = 0.f, a1 = 0.f;
        if (rr < nr) {
          const float* z = Z + (r0 + rr) * W;
          for (int k = 4 * l16; k < W; k += 64) {
            const float4 v = *reinterpret_cast<const float4*>(z + k);
            const float4 c0 = *reinterpret_cast<const float4*>(s_v0 + k);
            const float4 c1 = *reinterpret_cast<const float4*>(s_v1 + k);
            a0 += v.x * c0.x + v.y * c0.y + v.z * c0.z + v.w * c0.w;
            a1 += v.x * c1.x + v.y * c1.y + v.z * c1.z + v.w * c1.w;
            float4* sp = reinterpret_cast<float4*>(s_ss + slot * W + k);
            float4 q = *sp;
            q.x += v.x * v.x;
            q.y += v.y * v.y;
            q.z += v.z * v.z;
            q.w += v.w * v.w;
            *sp = q;
          }
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
          a0 += __shfl_xor(a0, o, 16);
          a1 += __shfl_xor(a1, o, 16);
        }
        if (l16 == 0 && rr < nr) put(rr, a0, a1);
      }
    } else {
      for (int k = lane; k < W; k += 64) s_ss[wid * W + k] = 0.f;
      for (int rr = 2 + wid; rr < nr; rr += blockDim.x >> 6) {
        float a0 = 0.f, a1 = 0.f;
        for (int k = lane; k < W; k += 64) {
          const float v = Z[(r0 + rr) * W + k];
          a0 += v * s_v0[k];
          a1 += v * s_v1[k];
          s_ss[wid * W + k] += v * v;
        }
        a0 = warp_sum(a0);
        a1 = warp_sum(a1);
        if (lane == 0) put(rr, a0, a1);
      }
    }
    __syncthreads();
    // Laplacian row
    const int nslot = W % 64 == 0 ? 16 : 4;
    float a0 = 0.f, a1 = 0.f;
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
      float ss = 0.f;
      for (int q = 0; q < nslot; ++q) ss += s_ss[q * W + k];
      const float y = s_f1[k] * Z[(r0 + 1) * W + k] + s_f2[k] * ss;
      a0 += y * W3[2 * k];
      a1 += y * W3[2 * k + 1];
    }
    bsum2(a0, a1);
    if (threadIdx.x == 0) put(1, a0, a1);
  }
  __syncthreads();
  const int c = threadIdx.x >> 7, t = threadIdx.x & 127;  // column, thread within the column
  float f, f1, f2;
  silu3(s_a[c][0], f, f1, f2);
  const float sc = scale[c];
  float ss = 0.f;
  for (int rr = 2 + t; rr < nr; rr += 128) {
    const float v = rr < JM_MAXR ? s_a[c][rr] : JR[(r0 + rr) * 2 + c];
    ss += v * v;
    JR[(r0 + rr) * 2 + c] = v * f1 * sc;
  }
  ss = warp_sum(ss);
  if (lane == 0) s_ss2[c][wid & 1] = ss;
  __syncthreads();
  if (t == 0) {
    JR[r0 * 2 + c] = f * sc + zb[c];
    if (nr > 1) JR[(r0 + 1) * 2 + c] = (s_a[c][1] * f1 + f2 * (s_ss2[c][0] + s_ss2[c][1])) * sc;
  }
}

void launch_jmlp_last(hipStream_t s, const float* Y, int W, int G, GroupDesc gd, const float* W3, const float* b3,
                      const float* scale, const float* zb, float* JR, const int32_t* dG) {
  if (G == 0) return;
  const size_t lds = sizeof(float) * 20 * (size_t)W;
  static size_t lds_set = 64 * 1024;
  if (lds > lds_set) {
    (void)hipFuncSetAttribute((const void*)k_jmlp_last, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    lds_set = lds;
  }
  k_jmlp_last<<<G, 256, lds, s>>>(Y, W, gd, W3, b3, scale, zb, JR, dG);
}

// Per walker and spin: sum over centre electrons of the value row and the Laplacian row.
__global__ void k_node_sums(const float* __restrict__ X, int64_t ld, int ncols, GroupDesc gd, int B, int N, int n_up,
                            int tangents, double* __restrict__ out) {
  const int w = blockIdx.x;
  {
    const int c = blockIdx.y * blockDim.x + threadIdx.x;  // one column per thread
    if (c >= ncols) return;
    double s[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i = 0; i < N; ++i) {
      const int g = group_of(w, i, B, n_up, N - n_up);
      const int64_t r0 = gd.row(g);
      const int sp = i < n_up ? 0 : 1;
      s[sp][0] += X[r0 * ld + c];
      if (tangents) s[sp][1] += X[(r0 + 1) * ld + c];
    }
    for (int sp = 0; sp < 2; ++sp)
      for (int k = 0; k < 2; ++k) out[(((size_t)w * 2 + sp) * 2 + k) * ncols + c] = s[sp][k];
  }
}

void launch_node_sums(hipStream_t s, const float* X, int64_t ld, int ncols, GroupDesc gd, int B, int N, int n_up,
                      double* out) {
  const int tangents = gd.a != 1 || gd.b != 0;
  k_node_sums<<<dim3(B, (ncols + 63) / 64), 64, 0, s>>>(X, ld, ncols, gd, B, N, n_up, tangents, out);
}

// Dense jacobian of a sum over electrons: for dependency n, sum the jacobian rows d X_i/d r_n
// over i in {n} U N(n), split by the spin of the centre i.
__global__ void k_dep_gather(const float* __restrict__ X, int64_t ld, int ncols, GroupDesc gd,
                             const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                             const int32_t* __restrict__ rev, int B, int N, int n_up, double* __restrict__ out) {
  const int blk = blockIdx.x;
  const int w = blk / N, n = blk - w * N;
  const int gn = group_of(w, n, B, n_up, N - n_up);
  const int64_t p0 = ptr[gn];
  const int nb = (int)(ptr[gn + 1] - p0);
  for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
    double s[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    {
      const int64_t rr = gd.row(gn) + 2;
      const int sp = n < n_up ? 0 : 1;
      for (int x = 0; x < 3; ++x) s[sp][x] += X[(rr + x) * ld + c];
    }
    // dependencies 8 at a time: all their rows are loaded before they are summed (in order)
    for (int a0 = 0; a0 < nb; a0 += 8) {
      float v[8][3];
      int spu[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        spu[u] = -1;
        if (a0 + u < nb) {
          const int i = nbr[p0 + a0 + u];
          const int gi = group_of(w, i, B, n_up, N - n_up);
          const int64_t rr = gd.row(gi) + 5 + 3 * (int64_t)rev[p0 + a0 + u];
          spu[u] = i < n_up ? 0 : 1;
#pragma unroll
          for (int x = 0; x < 3; ++x) v[u][x] = X[(rr + x) * ld + c];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          if (spu[u] == 0) s[0][x] += v[u][x];
          if (spu[u] == 1) s[1][x] += v[u][x];
        }
    }
    for (int sp = 0; sp < 2; ++sp)
      for (int x = 0; x < 3; ++x) out[((((size_t)w * 2 + sp) * N + n) * 3 + x) * ncols + c] = s[sp][x];
  }
}

void launch_dep_gather(hipStream_t s, const float* X, int64_t ld, int ncols, GroupDesc gd, const int64_t* ptr,
                       const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, double* out) {
  // every column in one pass: a narrow second pass would repeat the whole gather's latency
  const int threads = ncols > 1024 ? 1024 : (ncols + 63) / 64 * 64;
  k_dep_gather<<<B * N, threads, 0, s>>>(X, ld, ncols, gd, ptr, nbr, rev, B, N, n_up, out);
}

// Attention Jastrow per group: att = exp(-a / sqrt(Dk)) and vals = att * v with
// forward-Laplacian rows, where a = sum_k q[h,k] keys[h,k] arrives precomputed per row and
// head (the key-weighted head sums of the query projection are a linear map of the features,
// contracted as such).  V rows [H*Dk] at row stride ldv, QA rows [H] at row stride ldq;
// AV rows = [att (H) | vals (H*Dk)].
// Thread (h, k) owns column h*Dk + k, so every row is read once with coalesced loads.
template <bool TAN>
__global__ void __launch_bounds__(256) k_att_group(const float* __restrict__ V, int ldv, const float* __restrict__ QA,
                                                   int ldq, int H, int Dk, GroupDesc gd, float* __restrict__ AV,
                                                   const int32_t* __restrict__ dG) {
  const int g = blockIdx.x;
  if (dG && g >= *dG) return;
  const int HD = H * Dk;
  const int col = threadIdx.x;  // (h, k)
  if (col >= HD) return;
  const int h = col / Dk, k = col - h * Dk;
  const float sc = rsqrtf((float)Dk);
  const int64_t r0 = gd.row(g), r1 = TAN ? gd.row(g + 1) : r0 + 1;
  const int lda = H + HD;
  const float ax = QA[r0 * ldq + h];
  const float f = expf(-sc * ax), f1 = -sc * f, f2 = sc * sc * f;
  const float vx = V[r0 * ldv + col];
  AV[r0 * lda + H + col] = f * vx;
  if (k == 0) AV[r0 * lda + h] = f;
  if (!TAN) return;
  const float a_l = QA[(r0 + 1) * ldq + h];
  float ss = 0.f, cross = 0.f;
  auto row = [&](int64_t rr, float a_t, float vt) {
    const float at = a_t * f1;
    ss += a_t * a_t;
    cross += at * vt;
    AV[rr * lda + H + col] = at * vx + f * vt;
    if (k == 0) AV[rr * lda + h] = at;
  };
  // jacobian rows 8 at a time: the rows' a and V values are loaded before the stores
  int64_t rr = r0 + 2;
  for (; rr + 8 <= r1; rr += 8) {
    float q[8], v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      q[u] = QA[(rr + u) * ldq + h];
      v[u] = V[(rr + u) * ldv + col];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) row(rr + u, q[u], v[u]);
  }
  for (; rr < r1; ++rr) row(rr, QA[rr * ldq + h], V[rr * ldv + col]);
  const float al = a_l * f1 + f2 * ss;
  const float vl = V[(r0 + 1) * ldv + col];
  AV[(r0 + 1) * lda + H + col] = al * vx + f * vl + 2.f * cross;
  if (k == 0) AV[(r0 + 1) * lda + h] = al;
}

void launch_att_group(hipStream_t s, const float* V, int ldv, const float* QA, int ldq, int H, int Dk, int G,
                      GroupDesc gd, bool tangents, float* AV, const int32_t* dG) {
  if (G == 0) return;
  const int threads = ((H * Dk + 63) / 64) * 64;
  if (tangents)
    k_att_group<true><<<G, threads, 0, s>>>(V, ldv, QA, ldq, H, Dk, gd, AV, dG);
  else
    k_att_group<false><<<G, threads, 0, s>>>(V, ldv, QA, ldq, H, Dk, gd, AV, dG);
}

// Readout input u[s,h,k] = vals[s,h,k] / norm[s,h] with dense forward-Laplacian over the
// 3N coordinates (readout, jastrow.py:93-100).  Rows: value, Laplacian, then (n, xyz).
template <bool TAN>
__global__ void k_att_readin(const double* __restrict__ NS, const double* __restrict__ DG, int N, int H, int Dk,
                             float* __restrict__ RIN) {
  const int w = blockIdx.x;
  const int HD = H * Dk, ncol = H + HD;
  const int nrow = TAN ? 2 + 3 * N : 1;
  {
    const int col = blockIdx.y * blockDim.x + threadIdx.x;  // one column per thread
    if (col >= 2 * HD) return;
    const int sp = col / HD, hk = col - sp * HD, h = hk / Dk;
    const double* ns = NS + ((size_t)w * 2 + sp) * 2 * ncol;
    const double nx = ns[h], vx = ns[H + hk];
    const double rinv = 1.0 / nx;
    float* out = RIN + (size_t)w * nrow * 2 * HD + col;
    out[0] = (float)(vx * rinv);
    if (!TAN) return;
    const double nl = ns[ncol + h], vl = ns[ncol + H + hk];
    double sn2 = 0.0, cross = 0.0;
    const double* dg = DG + ((size_t)w * 2 + sp) * N * 3 * ncol;
    for (int t = 0; t < 3 * N; ++t) {
      const double nt = dg[(size_t)t * ncol + h], vt = dg[(size_t)t * ncol + H + hk];
      const double rt = -nt * rinv * rinv;
      sn2 += nt * nt;
      cross += vt * rt;
      out[(size_t)(2 + t) * 2 * HD] = (float)(vt * rinv + vx * rt);
    }
    const double rl = -nl * rinv * rinv + 2.0 * rinv * rinv * rinv * sn2;
    out[2 * HD] = (float)(vl * rinv + vx * rl + 2.0 * cross);
  }
}

void launch_att_readin(hipStream_t s, const double* NS, const double* DG, int B, int N, int H, int Dk, bool tangents,
                       float* RIN) {
  const dim3 grid(B, (2 * H * Dk + 63) / 64);
  if (tangents)
    k_att_readin<true><<<grid, 64, 0, s>>>(NS, DG, N, H, Dk, RIN);
  else
    k_att_readin<false><<<grid, 64, 0, s>>>(NS, DG, N, H, Dk, RIN);
}

// Per (walker, electron i): e-e cusp value/Laplacian (pairs j > i), cusp gradient on i
// (all j), Coulomb e-e (j > i) and e-n terms, all in FP64.
__global__ void k_cusp_coulomb(const float* __restrict__ r, const float* __restrict__ R, const float* __restrict__ Z,
                               int N, int n_up, int M, int cusp, const float* __restrict__ cp, double* __restrict__ cval,
                               double* __restrict__ clap, double* __restrict__ cgrad, double* __restrict__ epot) {
  __shared__ double red[16 * 6];
  const int blk = blockIdx.x;
  const int w = blk / N, i = blk - w * N;
  const float* rw = r + (size_t)w * N * 3;
  const double xi = rw[3 * i], yi = rw[3 * i + 1], zi = rw[3 * i + 2];
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // val, lap, gx, gy, gz, epot
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    if (j == i) continue;
    const double dx = xi - rw[3 * j], dy = yi - rw[3 * j + 1], dz = zi - rw[3 * j + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (j > i) v[5] += 1.0 / d;
    if (cusp) {
      const bool same = (i < n_up) == (j < n_up);
      const double al = same ? cp[0] : cp[1];
      const double fa = same ? cp[2] : cp[3];
      const double q = 1.0 / (al + d);
      const double c0 = fa * al * al * q;
      const double c1 = -c0 * q;
      const double c2 = -2.0 * c1 * q;
      v[2] += c1 * dx / d;
      v[3] += c1 * dy / d;
      v[4] += c1 * dz / d;
      if (j > i) {
        v[0] += c0;
        v[1] += 2.0 * (c2 + 2.0 * c1 / d);
      }
    }
  }
  for (int J = threadIdx.x; J < M; J += blockDim.x) {
    const double dx = xi - R[3 * J], dy = yi - R[3 * J + 1], dz = zi - R[3 * J + 2];
    v[5] -= Z[J] / sqrt(dx * dx + dy * dy + dz * dz);
  }
  block_sum<6>(v, red);
  if (threadIdx.x == 0) {
    cval[blk] = v[0];
    clap[blk] = v[1];
    cgrad[3 * (size_t)blk] = v[2];
    cgrad[3 * (size_t)blk + 1] = v[3];
    cgrad[3 * (size_t)blk + 2] = v[4];
    epot[blk] = v[5];
  }
}

void launch_cusp_coulomb(hipStream_t s, const float* r, const float* R, const float* Z, int B, int N, int n_up, int M,
                         int cusp, const float* cusp_params, double* cval, double* clap, double* cgrad, double* epot) {
  k_cusp_coulomb<<<B * N, 64, 0, s>>>(r, R, Z, N, n_up, M, cusp, cusp_params, cval, clap, cgrad, epot);
}

// log psi = det part + Jastrow; E_kin = -1/2 (lap + |grad|^2); E_loc = E_kin + E_pot.
__global__ void k_assemble(AssembleArgs A) {
  __shared__ double red[16 * 4];
  __shared__ double s_J[2][2];  // [component][value, lap]
  const int w = blockIdx.x, N = A.N;
  const int T = A.tangents;
  if (threadIdx.x == 0) {
    double Jx[2] = {0.0, 0.0}, Jl[2] = {0.0, 0.0};
    if (A.use_mlp || A.use_log)
      for (int sp = 0; sp < 2; ++sp)
        for (int c = 0; c < 2; ++c) {
          Jx[c] += A.mlp_ns[(((size_t)w * 2 + sp) * 2 + 0) * 2 + c];
          Jl[c] += A.mlp_ns[(((size_t)w * 2 + sp) * 2 + 1) * 2 + c];
        }
    if (A.use_att) {
      const size_t nrow = T ? 2 + 3 * N : 1;
      const float* ja = A.att_out + (size_t)w * nrow * 2;
      for (int c = 0; c < 2; ++c) {
        Jx[c] += (double)ja[c] * A.att_scale[c] + A.att_zb[c];
        if (T) Jl[c] += (double)ja[2 + c] * A.att_scale[c];
      }
    }
    for (int c = 0; c < 2; ++c) { s_J[c][0] = Jx[c]; s_J[c][1] = Jl[c]; }
  }
  __syncthreads();
  const double J0 = s_J[0][0], J1 = s_J[1][0];
  double v[4] = {0.0, 0.0, 0.0, 0.0};  // |grad|^2, |g1|^2 (log Jastrow), cusp val, cusp lap + epot...
  double cv = 0.0, cl = 0.0, ep = 0.0;
  for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) {
    if (!T) break;
    const int n = e / 3, x = e - 3 * n;
    double g0 = 0.0, g1 = 0.0;
    if (A.use_mlp || A.use_log)
      for (int sp = 0; sp < 2; ++sp) {
        g0 += A.mlp_dg[((((size_t)w * 2 + sp) * N + n) * 3 + x) * 2 + 0];
        g1 += A.mlp_dg[((((size_t)w * 2 + sp) * N + n) * 3 + x) * 2 + 1];
      }
    if (A.use_att) {
      const float* ja = A.att_out + ((size_t)w * (2 + 3 * N) + 2 + e) * 2;
      g0 += (double)ja[0] * A.att_scale[0];
      g1 += (double)ja[1] * A.att_scale[1];
    }
    double gJ = 0.0;
    if (A.use_mlp) gJ += g0;
    if (A.use_log) { gJ += g1 / J1; v[1] += g1 * g1; }
    if (A.cusp) gJ += A.cgrad[(size_t)w * 3 * N + e];
    const double gt = A.det_grad[(size_t)w * 3 * N + e] + gJ;
    if (A.grad) A.grad[(size_t)w * 3 * N + e] = gt;
    v[0] += gt * gt;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    cv += A.cval[(size_t)w * N + i];
    cl += A.clap[(size_t)w * N + i];
    ep += A.epot[(size_t)w * N + i];
  }
  v[2] = cv;
  v[3] = ep;
  double u[2] = {cl, 0.0};
  block_sum<4>(v, red);
  block_sum<2>(u, red);
  if (threadIdx.x == 0) {
    double lpJ = 0.0, sJ = 1.0, lapJ = 0.0;
    if (A.use_mlp) { lpJ += J0; lapJ += s_J[0][1]; }
    if (A.use_log) {
      lpJ += log(fabs(J1));
      sJ = J1 < 0 ? -1.0 : 1.0;
      lapJ += s_J[1][1] / J1 - v[1] / (J1 * J1);
    }
    if (A.cusp) { lpJ += v[2]; lapJ += u[0]; }
    const double logpsi = A.det_logpsi[w] + lpJ;
    if (A.logpsi) A.logpsi[w] = logpsi;
    if (A.sign) A.sign[w] = A.det_sign[w] * sJ;
    if (T) {
      const double lap = A.det_lap[w] + lapJ;
      if (A.lap) A.lap[w] = lap;
      const double ek = -0.5 * (lap + v[0]);
      const double epot = v[3] + A.e_nn;
      if (A.e_kin) A.e_kin[w] = ek;
      if (A.e_pot) A.e_pot[w] = epot;
      if (A.e_loc) A.e_loc[w] = ek + epot;
    }
  }
}

void launch_assemble(hipStream_t s, const AssembleArgs& a) { k_assemble<<<a.B, 256, 0, s>>>(a); }

}  // namespace fire
