// Launchers of the low-rank single-electron Metropolis kernels (k_mcmc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fire {

// step index of a Gaussian proposal = *dstep (if non-null) + step
void launch_mc_propose(hipStream_t s, const float* st_r, int B, int N, int k, float width, uint64_t seed, uint64_t step,
                       const int64_t* dstep, int mode, const float* rk_in, float* p_r, float* p_rk, double* p_u);
void launch_add_i64(hipStream_t s, int64_t* p, int64_t v);
void launch_set_i64(hipStream_t s, int64_t* p, int64_t v);
void launch_mc_affected(hipStream_t s, const float* st_r, const float* p_r, int B, int N, int n_up, int k, float cutoff,
                        int cap, int32_t* aff, int32_t* cnt_up, int32_t* cnt_dn, int32_t* overflow);
void launch_mc_scan(hipStream_t s, const int32_t* cu, const int32_t* cd, int B, int32_t* off_up, int32_t* off_dn,
                    int32_t* tail, int64_t* stats);
void launch_mc_rowlist(hipStream_t s, const int32_t* aff, const int32_t* cu, const int32_t* cd, const int32_t* off_up,
                       const int32_t* off_dn, const int32_t* tail, int B, int N, int n_up, int2* list, int32_t* rowof);
void launch_mc_klist(hipStream_t s, int B, int k, int2* list);

// V operand tiles per (walker, det), doubles: 4 row tiles x ceil(N/4) k steps x 64
__host__ __device__ inline size_t wb_vs_doubles(int N) { return (size_t)4 * 64 * ((N + 3) / 4); }
// scratch bytes per (walker, det): X operand tiles, then Q operand tiles (generic update)
__host__ __device__ inline size_t wb_update_scratch(int N) {
  return (size_t)2 * ((N + 15) / 16) * 16 * 64 * sizeof(double);
}
void launch_wb_prep(hipStream_t s, const float* p_orb, const float* p_envk, const float* st_env, const double* st_A,
                    const double* st_Ainv, const int32_t* aff, const int32_t* cu, const int32_t* cd,
                    const int32_t* rowof, int B, int N, int n_up, int D, int k, double* vs, double* sbuf, double* ldS,
                    double* sgS);
void launch_wb_update(hipStream_t s, const int32_t* accept, const int32_t* aff, const int32_t* cu, const int32_t* cd,
                      int B, int N, int D, const double* vs, const double* sbuf, double* scratch, double* st_A,
                      double* st_Ainv);

void launch_mc_jsum(hipStream_t s, const float* st_jr, const float* p_jr, const float* st_av, const float* p_av,
                    const int32_t* rowof, int B, int N, int n_up, int H, int Dk, int has_mlp, int has_att,
                    double* jsum, float* rin);
void launch_mc_accept(hipStream_t s, int B, int D, int use_mlp, int use_log, int use_att, int cusp, const double* st_ld,
                      const double* st_sg, const double* ldS, const double* sgS, const double* jsum,
                      const float* att_out, const float* att_scale, const float* att_zb, const double* ccusp,
                      const double* cdelta, const double* st_logpsi, const double* p_u, int mode, double* p_ld,
                      double* p_sg, double* p_logpsi, double* p_sign, int32_t* accept, int32_t* n_acc);
// change of the e-e cusp Jastrow when electron k moves (st_r -> p_r); per-walker cusp totals
void launch_mc_cusp_delta(hipStream_t s, const float* st_r, const float* p_r, int B, int N, int n_up, int k,
                          const float* cusp_params, double* cdelta);
void launch_mc_cusp_total(hipStream_t s, const double* cval, int B, int N, double* ccusp);
void launch_mc_commit(hipStream_t s, const int32_t* accept, const int2* list, int T, const int32_t* dT, int B, int N,
                      int n_up, int D,
                      int F, int ND, int nav, int k, const float* p_h, const float* p_orb, const float* p_jr,
                      const float* p_av, const float* p_rk, const float* p_M5k, const float* p_envk, const double* p_ld,
                      const double* p_sg, const double* p_logpsi, const double* p_sign, float* st_h, float* st_orb,
                      float* st_jr, float* st_av, float* st_r, float* st_M5, float* st_env, double* st_ld,
                      double* st_sg, double* st_logpsi, double* st_sign, float* p_r, const double* cdelta,
                      double* ccusp);
void launch_fill_i32(hipStream_t s, int32_t* p, int n, int v);
void launch_fill_f64(hipStream_t s, double* p, int n, double v);

}  // namespace fire
