// Pseudopotential (ECP) energy terms of a walker batch on the resident state.
//
// Local part: sum over electrons and ECP atoms of V_loc(|r_i - R_J|), each channel tabulated on
// a radial grid and linearly interpolated (make_local_pseudopotential,
// src/sparse_wf/pseudopotentials.py:204-218).
// Non-local part: every electron within the cutoff of its closest ECP atom is moved onto the
// randomly rotated spherical quadrature of that atom at its own distance; each point enters
// with sum_l (2l+1) V_l(|r' - R|) P_l(cos theta) w_q psi(r')/psi(r) (build_all_integration_points
// :269-336, make_nonlocal_pseudopotential :339-398).  k_ecp_points lays out the points and
// their coefficients; the psi ratios come from low-rank proposals of the resident state
// (fire_state_propose's path), one (electron, quadrature point) at a time for all walkers, and
// k_ecp_accum adds them up in a fixed order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fire_common.hpp"
#include "fire_kernels.hpp"

namespace fire {

namespace {

constexpr int kThreads = 128;

// jnp.interp(x, xp, fp): linear interpolation, constant beyond the ends
__device__ double interp(double x, const double* __restrict__ xp, const double* __restrict__ fp, int n) {
  int lo = 0, hi = n;  // first index with xp[idx] > x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] > x)
      hi = mid;
    else
      lo = mid + 1;
  }
  const int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
  const double dx = xp[i] - xp[i - 1];
  double f = dx == 0.0 ? fp[i] : fp[i - 1] + ((x - xp[i - 1]) / dx) * (fp[i] - fp[i - 1]);
  if (x < xp[0]) f = fp[0];
  if (x > xp[n - 1]) f = fp[n - 1];
  return f;
}

__device__ __forceinline__ double legendre(int l, double x) {
  switch (l) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * x * x - 1.0);
    default: return 0.5 * (5.0 * x * x * x - 3.0 * x);
  }
}

// Rotation (v1, v2, v3) of random_rot_mat (pseudopotentials.py:220-226): v1, v2 ~ N(0, I)
// from two Philox blocks under key (seed, walker), counter (electron, 0x45435000 + block);
// v1 normalised, v2 orthogonalised against v1 and normalised, v3 = v1 x v2.
__device__ void rotation(uint64_t seed, int w, int i, double v[3][3]) {
  const uint2 key = make_uint2((uint32_t)seed ^ (uint32_t)(seed >> 32), (uint32_t)w);
  double n[8];
  for (int b = 0; b < 2; ++b) {
    const uint4 a = philox4x32_10(make_uint4((uint32_t)i, 0x45435000u + (uint32_t)b, 0u, 0u), key);
    const double two_pi = 6.283185307179586;
    const double r1 = sqrt(-2.0 * log(u01(a.x))), r2 = sqrt(-2.0 * log(u01(a.z)));
    n[4 * b + 0] = r1 * cos(two_pi * u01(a.y));
    n[4 * b + 1] = r1 * sin(two_pi * u01(a.y));
    n[4 * b + 2] = r2 * cos(two_pi * u01(a.w));
    n[4 * b + 3] = r2 * sin(two_pi * u01(a.w));
  }
  double a1[3] = {n[0], n[1], n[2]}, a2[3] = {n[3], n[4], n[5]};
  const double l1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  for (int c = 0; c < 3; ++c) a1[c] /= l1;
  const double d = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
  for (int c = 0; c < 3; ++c) a2[c] -= a1[c] * d;
  const double l2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
  for (int c = 0; c < 3; ++c) a2[c] /= l2;
  for (int c = 0; c < 3; ++c) {
    v[0][c] = a1[c];
    v[1][c] = a2[c];
  }
  v[2][0] = a1[1] * a2[2] - a1[2] * a2[1];
  v[2][1] = a1[2] * a2[0] - a1[0] * a2[2];
  v[2][2] = a1[0] * a2[1] - a1[1] * a2[0];
}

// One workgroup per walker, threads over electrons.
__global__ __launch_bounds__(kThreads) void k_ecp_points(const float* __restrict__ r, const float* __restrict__ R, int B,
                                                         int N, EcpTables t, uint64_t seed, float* __restrict__ pts,
                                                         double* __restrict__ coef, int32_t* __restrict__ kq,
                                                         double* __restrict__ vloc, unsigned long long* __restrict__ n_pts) {
  __shared__ double red[kThreads];
  const int w = blockIdx.x;
  double vl = 0.0;
  unsigned np = 0;
  for (int i = threadIdx.x; i < N; i += kThreads) {
    const float* ri = r + ((size_t)w * N + i) * 3;
    const float x = ri[0], y = ri[1], z = ri[2];
    float best = 0.f;
    int jb = -1;
    for (int e = 0; e < t.n_ecp; ++e) {
      const float* RJ = R + 3 * t.atom[e];
      // float32 |r - R| rounded step by step (no fused multiply-add), as the oracle
      const float dx = __fsub_rn(x, RJ[0]), dy = __fsub_rn(y, RJ[1]), dz = __fsub_rn(z, RJ[2]);
      const float d = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)));
      vl += interp((double)d, t.grid, t.vloc + (size_t)e * t.n_grid, t.n_grid);
      if (jb < 0 || d < best) {
        best = d;
        jb = e;
      }
    }
    const float* RJ = R + 3 * t.atom[jb];
    const bool active = (double)best < t.cut[jb];
    const int n = t.nq[jb];
    double v[3][3];
    rotation(seed, w, i, v);
    const float dir[3] = {__fdiv_rn(__fsub_rn(x, RJ[0]), best), __fdiv_rn(__fsub_rn(y, RJ[1]), best),
                          __fdiv_rn(__fsub_rn(z, RJ[2]), best)};
    for (int q = 0; q < t.qmax; ++q) {
      const size_t slot = ((size_t)i * t.qmax + q) * B + w;
      float p[3] = {x, y, z};
      double c = 0.0;
      if (active && q < n) {
        const double* g = t.qpts + ((size_t)jb * t.qmax + q) * 3;
        double gr[3], cs = 0.0;
        for (int a = 0; a < 3; ++a) {
          gr[a] = g[0] * v[a][0] + g[1] * v[a][1] + g[2] * v[a][2];
          cs += gr[a] * (double)dir[a];
          p[a] = (float)((double)RJ[a] + (double)best * gr[a]);
        }
        const double ex = (double)p[0] - RJ[0], ey = (double)p[1] - RJ[1], ez = (double)p[2] - RJ[2];
        const double dae = sqrt(ex * ex + ey * ey + ez * ez);
        for (int l = 0; l < t.n_l; ++l)
          c += (2 * l + 1) * interp(dae, t.grid, t.vnl + ((size_t)jb * t.n_l + l) * t.n_grid, t.n_grid) * legendre(l, cs);
        c /= (double)n;
        ++np;
      }
      pts[slot * 3 + 0] = p[0];
      pts[slot * 3 + 1] = p[1];
      pts[slot * 3 + 2] = p[2];
      coef[slot] = c;
    }
    if (active) atomicMax(kq + i, n);
  }
  if (np) atomicAdd(n_pts, (unsigned long long)np);
  red[threadIdx.x] = vl;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) vloc[w] = red[0];
}

// v[w] += coef[w] * sign'(w) sign(w) exp(log|psi'|(w) - log|psi|(w))
__global__ void k_ecp_accum(int B, const double* __restrict__ coef, const double* __restrict__ lp_new,
                            const double* __restrict__ sg_new, const double* __restrict__ lp, const double* __restrict__ sg,
                            double* __restrict__ v) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  const double c = coef[w];
  if (c != 0.0) v[w] += c * sg_new[w] * sg[w] * exp(lp_new[w] - lp[w]);
}

}  // namespace

void launch_ecp_points(hipStream_t s, const float* r, const float* R, int B, int N, const EcpTables& t, uint64_t seed,
                       float* pts, double* coef, int32_t* kq, double* vloc, unsigned long long* n_pts) {
  hipLaunchKernelGGL(k_ecp_points, dim3(B), dim3(kThreads), 0, s, r, R, B, N, t, seed, pts, coef, kq, vloc, n_pts);
}

void launch_ecp_accum(hipStream_t s, int B, const double* coef, const double* lp_new, const double* sg_new,
                      const double* lp, const double* sg, double* v) {
  hipLaunchKernelGGL(k_ecp_accum, dim3((B + 255) / 256), dim3(256), 0, s, B, coef, lp_new, sg_new, lp, sg, v);
}

}  // namespace fire
