// Launchers and argument blocks of the parameter-gradient kernels (k_grad.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fire_kernels.hpp"

namespace fire {

// Rows of walker b: [rng[4b], rng[4b+1]) and [rng[4b+2], rng[4b+3]) when rng is set, else
// [a0 + b as, + na) and [c0 + b cs, + nc).  avg_rows: the mean row count per walker when rng is
// set (sizes the split of each walker's rows over workgroups).
struct WRows {
  const int64_t* rng = nullptr;
  int64_t a0 = 0, as = 0;
  int na = 0;
  int64_t c0 = 0, cs = 0;
  int nc = 0;
  int64_t avg_rows = 0;
};

// per-edge forward record (e-e pairs and e-n edges)
constexpr int PR_FEAT = 0;   // [d, dx, dy, dz]
constexpr int PR_D = 4;      // d
constexpr int PR_PRE = 5;    // [w0] filter pre-activations
constexpr int PR_ENV = 37;   // [E] radial envelopes (e-e pairs, E = w0)
constexpr int PR_DIR = 69;   // [w1] directional part
constexpr int PR_ENVD = 85;  // [w1] radial part after Dense_1
constexpr int PR_FLOATS = 104;

struct GradHead {
  int B, N, nu, D, W, depth, H, Dk, HD, lwmax;
  int has_mlp, use_mlp, use_log, has_att, cusp;
  const double *ldet, *dsgn;
  const float *z3, *mlp_scale, *mlp_zb;
  const float *qv, *keys;
  const float* ro_W[8];
  const float* ro_b[8];
  const float *att_scale, *att_zb;
  const float* cusp_p;
  const float* r;
  double *wdet, *logpsi;
  float *dJ, *dval, *dnorm;
  float* gout;
  int64_t gstride;
  int64_t o_cusp, o_mlp_scale, o_mlp_zb, o_att_scale, o_att_zb;
  int64_t o_ro_W[8], o_ro_b[8];
};

struct GradRows {
  int B, N, nu, D, H, Dk, HD, has_mlp, has_att;
  const double *wdet, *Ainv;
  const float *orb, *env, *z3, *dJ, *mlp_scale, *qv, *keys, *dval, *dnorm;
  float *dorb, *denv, *dz3, *dqv, *dkey;
};

// part (optional, part_cap floats): scratch for per-slice partial products when a walker's
// rows are split over several workgroups; without it every walker's rows run in one.
void launch_wgrad(hipStream_t s, const float* X, int64_t ldx, const float* Y, int64_t ldy, int K, int J, int B,
                  const WRows& wr, float* out, int64_t so, bool accumulate = false, const uint8_t* rowmask = nullptr,
                  int want = 1, float* part = nullptr, int64_t part_cap = 0);
void launch_gr_en_cot(hipStream_t s, const float* beta, const float* lsf, const uint8_t* mask, const float* dhen,
                      const float* rec, float cutoff_1el, const EnParams& ep, int B, int N, int nu, int M, int F,
                      float* part, float* part2, float* gout, int64_t gstride, int64_t o_wg, int64_t o_wf,
                      int64_t o_hn, int64_t o_K, int64_t o_b, int64_t o_sp, int64_t o_wd, int64_t o_bd, int64_t o_we);
void launch_gr_pair_cot(hipStream_t s, const int64_t* ptr, const int32_t* nbr, const uint8_t* same, int want,
                        const float* beta, const float* lsf, const float* rec, const float* M5, const float* dh,
                        const EdgeParams& ep, float cut, int B, int N, int nu, int F, float* du, int64_t oK, int64_t ob,
                        int64_t osp, int64_t oWd, int64_t obd, int64_t oWe, int64_t oWG, int64_t oWF, int64_t oBF,
                        int64_t ocut, bool add_cut, float* part, float* gout, int64_t gstride);
void launch_silu_bwd(hipStream_t s, float* dz, int64_t ldo, const float* dy, int64_t ldy, const float* z, int64_t ldz,
                     int64_t rows, int cols);
void launch_silu_fwd(hipStream_t s, float* y, const float* z, int64_t n);
void launch_axpy(hipStream_t s, float* a, const float* b, int64_t n);
void launch_fill_f32(hipStream_t s, float* a, int64_t n, float v);
void launch_mul(hipStream_t s, float* o, const float* a, const float* b, int64_t n);
void launch_transpose2d(hipStream_t s, const float* A, int R, int C, float* AT);
void launch_gr_env(hipStream_t s, const float* r, const float* R, int B, int N, int nu, int M, int D, int E,
                   const float* sp_up, const float* pi_up, const float* sp_dn, const float* pi_dn, float* ex,
                   float* env, float* dist);
void launch_gr_phi(hipStream_t s, const float* orb, const float* env, int B, int N, int nu, int D, double* A);
size_t gr_head_lds(const GradHead& a);
void launch_gr_head(hipStream_t s, const GradHead& a);
void launch_gr_rows(hipStream_t s, const GradRows& a, int64_t G);
void launch_gr_sp(hipStream_t s, const float* ex, const float* tt, const float* dist, int B, int N, int nu, int M,
                  int D, int E, float* gout, int64_t gstride, int64_t o_up, int64_t o_dn);
void launch_gr_pair_fwd(hipStream_t s, const float* r, int B, int N, int nu, const int64_t* ptr, const int32_t* nbr,
                        const EdgeParams& ps, const EdgeParams& pd, float cut, int64_t P, float* rec, float* beta,
                        float* lsf, uint8_t* same, int32_t* ctr);
void launch_gr_msg_collect(hipStream_t s, const float* du, const int64_t* ptr, const int32_t* nbr, const int32_t* rev,
                           const uint8_t* same, int B, int N, int nu, int F, float* dM5);
void launch_gr_en_fwd(hipStream_t s, const float* r, const float* R, int B, int N, int nu, int M, float cutoff_1el,
                      const EnParams& ep, int64_t rows, float* rec, float* beta, float* lsf, uint8_t* mask);
void launch_gr_en_sum(hipStream_t s, const float* beta, const float* lsf, const EnParams& ep, const uint8_t* mask,
                      int64_t G, int M, int F, float* hen);
void launch_gr_ln_fwd(hipStream_t s, const float* h, int64_t G, int F, const float* gsc, const float* gb, float* y,
                      float* nrm, float* inv);
void launch_gr_ln_bwd(hipStream_t s, const float* dy, const float* nrm, const float* inv, const float* gsc, int64_t G,
                      int F, float* dh);
void launch_gr_glu_fwd(hipStream_t s, const float* Z2, int64_t G, int F, float* y);
void launch_gr_glu_bwd(hipStream_t s, const float* dy, const float* Z2, int64_t G, int F, float* dZ2);
}  // namespace fire
