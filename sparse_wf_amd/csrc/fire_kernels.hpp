// Kernel launchers shared between the kernel translation units and the host API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fire_common.hpp"

namespace fire {

// Row layout of a family of tangent-row groups: group g starts at row a*g + b*ptr[g]
// (ptr = CSR pair offsets; b = 0 for uniform groups) and has
// row_start(g+1) - row_start(g) rows; row 0 = value, row 1 = Laplacian, rows >= 2 jacobian.
struct GroupDesc {
  int a, b;
  const int64_t* ptr;
  __host__ __device__ __forceinline__ int64_t row(int64_t g) const { return a * g + (b ? b * ptr[g] : 0); }
};

struct EnParams {  // e-n edges of ElecInit (new_sparse_model.py:139-158, 382-397)
  const float *K, *b, *sp, *hnuc;  // [M,4,w0], [M,w0], [M,E], [M,F]
  const float *Wd, *bd, *We;       // [w0,w1], [w1], [E,w1]
  const float *WG, *WF;            // [w1,F], [4,F]
  const float *ln_g, *ln_b;        // [F]
  int w0, w1, E;
};

struct EdgeParams {  // ElecElecEdges (new_sparse_model.py:161-187)
  FilterParams filt;
  const float *WG, *WF, *BF;  // [w1,F], [4,F], [F]
};

enum Act { ACT_SILU = 0, ACT_EXP_NEG_SCALED = 1 };

// graph
// list mode: n_list = capacity of the list; dn (optional, device) = entries actually in use;
// otherwise all B * N centres, indexed in group order or, with walker_major, as w * N + i
void launch_nbr_count(hipStream_t s, const float* r, int B, int N, int n_up, float cutoff, int32_t* cnt,
                      const int2* list = nullptr, int n_list = 0, const int32_t* dn = nullptr,
                      bool walker_major = false);
void launch_nbr_fill(hipStream_t s, const float* r, int B, int N, int n_up, float cutoff, const int64_t* ptr,
                     int32_t* nbr, const int2* list = nullptr, int n_list = 0, const int32_t* dn = nullptr,
                     bool walker_major = false);
void launch_scan(hipStream_t s, const int32_t* cnt, int G, int64_t* ptr, const int32_t* dn = nullptr);
void launch_rev(hipStream_t s, const int64_t* ptr, const int32_t* nbr, int B, int N, int n_up, int32_t* rev);

// embedding
// list (optional): (walker, electron) pairs to process, output row = list position
// with tangents: ef (G * M * pair_filter_floats(P.w1) floats) and eok (G * M bytes) are scratch
void launch_elec_init_en(hipStream_t s, const float* r, const float* R, int B, int N, int n_up, int M, float cutoff_1el,
                         int F, EnParams P, bool tangents, float* out, const int2* list = nullptr, int n_list = 0,
                         float* ef = nullptr, uint8_t* eok = nullptr);
void launch_glu(hipStream_t s, const float* Z, int G, int F, bool tangents, float* out);
// dG (optional, device): groups actually in use (G is the launch capacity)
// dcount (optional, device): only the first *dcount of the M rows
void launch_rowmax(hipStream_t s, const float* A, int64_t lda, int64_t M, int K, uint32_t* amax,
                   const int32_t* dcount = nullptr);
void launch_act(hipStream_t s, float* Y, int64_t ld, int ncols, int G, GroupDesc gd, int act, float act_scale,
                const int32_t* dG = nullptr);
// list mode: centres list[0..n_list) with CSR ptr/nbr over list positions; the message
// rows of electron `kmove` of each walker are read from M5k[walker] (value rows only)
void launch_messages(hipStream_t s, const float* r, int B, int N, int n_up, const int64_t* ptr, const int32_t* nbr,
                     const float* M5, EdgeParams same, EdgeParams diff, float cut, int F, bool tangents, float* out,
                     const int2* list = nullptr, int n_list = 0, int kmove = -1, const float* M5k = nullptr,
                     const int32_t* dn = nullptr, float* pf = nullptr, uint32_t* omax = nullptr);
// floats per pair of the pair-filter buffer pf that launch_messages needs with tangent rows
size_t pair_filter_floats(int w1);
void launch_scale(hipStream_t s, float* p, int64_t n, float a);
void launch_zero_f32(hipStream_t s, float* p, int64_t n);
void launch_rowflag(hipStream_t s, int G, GroupDesc gd, int64_t n_rows, uint8_t* flag);

// Optional device-side row range of a contraction (for launches sized by a capacity, e.g. under
// stream capture): rows [*offset, *offset + min(M, *count)) of A / C / rowflag; null = 0 / M.
struct DevRows {
  const int32_t* count = nullptr;
  const int32_t* offset = nullptr;
};

// Forward-Laplacian silu applied in the epilogue of a contraction over tangent rows (the
// rule of k_act, split in three):
//   before: z0 = the value row of every group (FP32 contraction of the gathered value rows),
//           f1[g][c] = silu'(z0);
//   epilogue (rowgrp[row] = g for jacobian rows, -1 value rows, -2 Laplacian rows):
//           value rows silu(z + bias), jacobian rows f1[g][c] z, Laplacian rows z unchanged,
//           and the sum of z^2 over the jacobian rows of group g that one wavefront holds for
//           one column, in part[g][slot][c] (slot = 2 (64-row block - first block of the
//           group's jacobian rows) + lane half; at most 8 slots);
//   after:  Laplacian rows f1 z + silu''(z0) * (sum of the group's slots), in k_act_lapfix.
// rowgrp == nullptr: plain store.
// FP16 planes (nsplit 4): amax[row] = float bits of max |A[row, :]|, winv[col] = inverse
// power-of-two scale of W's column.  omax (optional, fused epilogue only): row maxima of the
// output over its first omax_cols columns, accumulated with atomicMax into a zeroed array
// (Laplacian rows are left to k_act_lapfix), for the next contraction's FP16 planes.
struct ActEpi {
  const int32_t* rowgrp = nullptr;
  const float* f1 = nullptr;
  float* part = nullptr;
  GroupDesc gd{0, 0, nullptr};
  int ldf = 0;            // row stride of f1 and of each part slot
  int nfuse = 1 << 30;    // columns at or past nfuse are stored plain (+ bias on value rows)
  const uint32_t* amax = nullptr;
  const float* winv = nullptr;
  uint32_t* omax = nullptr;
  int omax_cols = 0;
};
constexpr int ACT_EPI_SLOTS = 8;

void launch_rowgrp(hipStream_t s, int G, GroupDesc gd, int32_t* rowgrp);
void launch_gather_rows(hipStream_t s, const float* A, int64_t lda, int K, int G, GroupDesc gd, float* out);
void launch_silu_grad(hipStream_t s, const float* z0, float* f1, int64_t n);
// omax (optional): sets the row maxima of the completed Laplacian rows (atomicMax)
void launch_act_lapfix(hipStream_t s, float* C, int64_t ldc, int N, int G, GroupDesc gd, const float* z0,
                       const float* part, int ldf, uint32_t* omax = nullptr);

// dense contractions: C[M,N] = A[M,K] W[K,N] (+ bias on rows with flag / row % period == 0)
// W is [K,N] (used for N <= 16), Wt its transpose [N,K] (used by the MFMA kernel)
void launch_gemm_f32(hipStream_t s, const float* A, int64_t lda, const float* W, const float* Wt, const float* bias,
                     float* C, int64_t ldc, int64_t M, int N, int K, const uint8_t* rowflag, int period,
                     DevRows dr = DevRows(), ActEpi ep = ActEpi());
void launch_gemm_f32_tail(hipStream_t s, const float* A, int64_t lda, const float* Wt, const float* bias, float* C,
                          int64_t ldc, int64_t M, int n0, int nr, int K, const uint8_t* rowflag, int period,
                          DevRows dr, int64_t sA, int64_t sW, int64_t sC, int nb);
void launch_gemm_f32_batched(hipStream_t s, const float* A, int64_t lda, int64_t sA, const float* Wt, int64_t sW,
                             float* C, int64_t ldc, int64_t sC, int64_t M, int N, int K, int nb);
// nsplit 1: plain BF16; 3: W = hi + lo planes, 3 products; 6: W = hi + mid + lo, 6 products;
// 4: two FP16 planes of the scaled operands (ep.amax, ep.winv), 3 products
// variant: tile shape index in [0, gemm_bf16_variants()) (0 = 128 x 128)
void launch_gemm_bf16(hipStream_t s, const float* A, int64_t lda, const uint16_t* Whi /*[N,K] bf16*/,
                      const uint16_t* Wlo, const uint16_t* Wlo2, const float* bias, float* C, int64_t ldc, int64_t M,
                      int N, int K, const uint8_t* rowflag, int period, int nsplit, int variant = 0,
                      DevRows dr = DevRows(), ActEpi ep = ActEpi());
int gemm_bf16_variants();
// variants that have a fused-silu (ActEpi) form; the others are never launched with one
bool gemm_bf16_fusable(int variant);
// variants that have an FP16-plane form (nsplit 4)
bool gemm_bf16_f16_ok(int variant);
bool gemm_bf16_spills(int variant, int nsplit);
bool gemm_bf16_persistent(int variant);
void gemm_bf16_tile(int variant, int* bm, int* bn);

// pseudopotential tables on the device (radial grids in FP64)
struct EcpTables {
  int n_ecp, n_l, n_grid, qmax;
  const int32_t* atom;  // [n_ecp] nucleus index
  const double* grid;   // [n_grid] radii
  const double* vloc;   // [n_ecp][n_grid] local channel
  const double* vnl;    // [n_ecp][n_l][n_grid] non-local channels l = 0 .. n_l - 1
  const double* cut;    // [n_ecp] non-local cutoff radius
  const int32_t* nq;    // [n_ecp] quadrature points
  const double* qpts;   // [n_ecp][qmax][3] unrotated quadrature directions
};
void launch_ecp_points(hipStream_t s, const float* r, const float* R, int B, int N, const EcpTables& t, uint64_t seed,
                       float* pts, double* coef, int32_t* kq, double* vloc, unsigned long long* n_pts);
void launch_ecp_accum(hipStream_t s, int B, const double* coef, const double* lp_new, const double* sg_new,
                      const double* lp, const double* sg, double* v);

// local-energy statistics and clipping of one batch (one workgroup)
void launch_energy_stats(hipStream_t s, const double* e, int B, double clip, int stat, double* clipped, double* diff,
                         double* stats);

}  // namespace fire
