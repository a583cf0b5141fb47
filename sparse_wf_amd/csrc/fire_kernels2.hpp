// Launchers of the orbital / determinant / Jastrow / energy kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fire_kernels.hpp"

namespace fire {

// orbitals and determinants
void launch_env_coef(hipStream_t s, const float* r, const float* R, int B, int N, int n_up, int M, int D, int E,
                     const float* sp_up, const float* sp_dn, bool tangents, float* coef, const int2* list = nullptr,
                     int n_list = 0);
void launch_orb_product(hipStream_t s, float* O, int ldo, int ND, int G, GroupDesc gd, const float* env, bool tangents);
void launch_gather_A(hipStream_t s, const float* O, int ldo, int ND, GroupDesc gd, int B, int N, int n_up, int D,
                     double* A);
void launch_lu_inv(hipStream_t s, double* A, int n_mat, int N, double* Ainv, float* AinvT32, double* logdet,
                   double* sign, bool want_inv, bool use_lds);
size_t lu_lds_bytes(int N);
hipError_t lu_set_smem_limit(size_t bytes);
void launch_det_trace(hipStream_t s, const float* O, int ldo, GroupDesc gd, int B, int N, int n_up, int D,
                      const double* Ainv, double* tr);
void launch_det_lap(hipStream_t s, const float* O, int ldo, int ND, GroupDesc gd, const float* env, const int64_t* ptr,
                    const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, int D, int smax,
                    const float* AinvT, double* trip, double* grad);
// FP64 form of launch_det_lap on the FP64 transposed inverse (det_lap64_supported: N even, ND
// and ldo multiples of 4 and 32)
void launch_det_lap64(hipStream_t s, const float* O, int ldo, int ND, GroupDesc gd, const float* env,
                      const int64_t* ptr, const int32_t* nbr, const int32_t* rev, int B, int N, int n_up, int D,
                      int smax, const double* AinvT, double* trip, double* grad);
bool det_lap64_supported(int N, int ND, int ldo);
void launch_transpose_f64(hipStream_t s, const double* X, int n_mat, int N, double* XT);
void launch_det_combine(hipStream_t s, int B, int N, int D, const double* logdet, const double* sgn, const double* tr,
                        const double* trip, const double* grad, double* out_logpsi, double* out_sign,
                        double* out_grad, double* out_lap, bool tangents);

// Jastrow
// Y: pre-activation rows of the last hidden layer (the silu is applied inside)
void launch_jmlp_last(hipStream_t s, const float* Y, int W, int G, GroupDesc gd, const float* W3, const float* b3,
                      const float* scale, const float* zb, float* JR, const int32_t* dG = nullptr);
void launch_node_sums(hipStream_t s, const float* X, int64_t ld, int ncols, GroupDesc gd, int B, int N, int n_up,
                      double* out /*[B][2 spin][2 (value, lap)][ncols]*/);
void launch_dep_gather(hipStream_t s, const float* X, int64_t ld, int ncols, GroupDesc gd, const int64_t* ptr,
                       const int32_t* nbr, const int32_t* rev, int B, int N, int n_up,
                       double* out /*[B][2 spin][N][3][ncols]*/);
void launch_att_group(hipStream_t s, const float* V /*[rows][H Dk]*/, int ldv, const float* QA /*[rows][H]*/, int ldq,
                      int H, int Dk, int G, GroupDesc gd, bool tangents, float* AV, const int32_t* dG = nullptr);
void launch_att_readin(hipStream_t s, const double* NS, const double* DG, int B, int N, int H, int Dk, bool tangents,
                       float* RIN);
void launch_cusp_coulomb(hipStream_t s, const float* r, const float* R, const float* Z, int B, int N, int n_up, int M,
                         int cusp, const float* cusp_params, double* cval, double* clap, double* cgrad, double* epot);

struct AssembleArgs {
  int B, N, use_mlp, use_log, use_att, cusp, tangents;
  const double *det_logpsi, *det_sign, *det_grad, *det_lap;
  const double *mlp_ns, *mlp_dg;  // node sums [B][2][2][2], dep gather [B][2][N][3][2]
  const float* att_out;           // [B][2+3N][2] (or [B][1][2] without tangents)
  const float *att_scale, *att_zb;
  const double *cval, *clap, *cgrad, *epot;
  double e_nn;
  double *e_loc, *e_kin, *e_pot, *logpsi, *sign, *grad, *lap;
};
void launch_assemble(hipStream_t s, const AssembleArgs& a);

}  // namespace fire
