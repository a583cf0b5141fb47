/*
 * fire.h -- C ABI of the MI355X-native FiRE hot path.
 *
 * One handle = one molecule + one FiRE model (``embedding: new_sparse``).  Every entry
 * point returns 0 on success and a non-zero status otherwise; the message of the last
 * failure on the calling thread is returned by fire_last_error().  No exception crosses
 * the boundary.  All array arguments named *_dev are device pointers on the handle's
 * device; calls are stream-ordered on the given hipStream_t (passed as void*).
 * Workspace is owned by the handle and reused between calls; a call grows it when a
 * batch needs more.  Calls on one handle must not overlap (they share the workspace).
 *
 * The reference binds this path through Python/JAX callables; each entry point below names
 * the reference function it replaces (paths relative to the sparse_wf repository).
 */
#ifndef FIRE_H
#define FIRE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fire_handle fire_handle;

/* Model hyper-parameters: model_args.embedding.new (src/sparse_wf/api.py:426-434),
 * model_args.jastrow (api.py:453-462), isotropic envelopes (api.py:475-478),
 * n_determinants (config/default.yaml:27). */
typedef struct fire_config {
  int32_t n_el, n_up, n_nuc;
  int32_t feature_dim;       /* F */
  int32_t filter_w0, filter_w1; /* pair_mlp_widths */
  int32_t n_env_en;          /* pair_n_envelopes */
  int32_t n_det;             /* n_determinants */
  int32_t n_env_iso;         /* envelopes.isotropic_args.n_envelopes */
  int32_t e_e_cusp;          /* 0 = none, 1 = psiformer */
  int32_t use_mlp_jastrow, use_log_jastrow, use_attention;
  int32_t mlp_width, mlp_depth, att_heads, att_dim;
  float cutoff, cutoff_1el;
} fire_config;

enum fire_precision {
  FIRE_FP32 = 0,  /* every embedding / orbital / Jastrow contraction in FP32 (FP32 MFMA) */
  FIRE_MIXED = 1, /* dense contractions as split BF16 (hi*hi + hi*lo + lo*hi) on BF16 MFMA,
                     FP32 accumulation; element-wise rules FP32; determinants FP64 */
  FIRE_BF16 = 2,  /* dense contractions in single BF16 on MFMA (fastest, least accurate) */
};

/* Creates a handle for molecule (R [n_nuc*3] bohr, Z [n_nuc]) on HIP device `device`.
 * Replaces MoonLikeWaveFunction.create (src/sparse_wf/model/wave_function.py:102-150). */
int fire_create(const fire_config* cfg, const float* R, const float* Z, int device, fire_handle** out);
int fire_destroy(fire_handle* h);

/* Uploads parameters from host memory: `n_tensors` float32 tensors back to back in the
 * canonical order listed in sparse_wf_amd/params.py (pack_params), `sizes[k]` = element
 * count of tensor k; sizes are validated against the config.  Replaces the params pytree
 * MoonLikeParams (wave_function.py:54-57) handed to every apply call. */
int fire_upload_params(fire_handle* h, const float* packed, const int64_t* sizes, int32_t n_tensors);

/* Neighbour lists of B walkers r_dev [B, n_el, 3]: counts_dev [B*n_el] and, if
 * nbr_dev != NULL, the neighbour indices in CSR order (walker, centre ascending,
 * neighbour ascending) with capacity `cap`; *n_pairs receives the total.  Strict
 * d < cutoff on the float32 distance sqrt((dx*dx+dy*dy)+dz*dz).
 * Replaces get_pair_indices / is_neighbour (src/sparse_wf/model/new_sparse_model.py:238-263,
 * src/sparse_wf/model/graph_utils.py:179-184). */
int fire_neighbours(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t* counts_dev,
                    int32_t* nbr_dev, int64_t cap, int64_t* n_pairs);

/* Local energy of B walkers: E_loc = E_kin + E_pot with the sparse forward-Laplacian
 * kinetic energy.  Outputs (device, double, [B]); any may be NULL except e_loc_dev.
 * grad_dev (optional, double [B, n_el, 3]) and lap_dev (optional, double [B]) receive the
 * gradient and Laplacian of log|psi|.
 * Replaces make_local_energy(...)(key, params, r, static) (src/sparse_wf/hamiltonian.py:49-81)
 * with MoonLikeWaveFunction.kinetic_energy (wave_function.py:215-221) and potential_energy
 * (hamiltonian.py:15-26). */
int fire_local_energy(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision,
                      double* e_loc_dev, double* e_kin_dev, double* e_pot_dev, double* logpsi_dev,
                      double* sign_dev, double* grad_dev, double* lap_dev);

/* log|psi| and sign of B walkers (value path).
 * Replaces MoonLikeWaveFunction.signed (wave_function.py:183-190). */
int fire_signed(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision,
                double* logpsi_dev, double* sign_dev);

/* Embeddings h [B, n_el, F] (float32) after message passing (`h_contr_dev`, optional)
 * and after the update MLP and scaling (`h_out_dev`, optional).
 * Replaces NewSparseEmbedding.apply (new_sparse_model.py:483-551); h_contr is the input of
 * the first normalize() used by init (:535, :378). */
int fire_embedding(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision,
                   float* h_contr_dev, float* h_out_dev);

/* Instrumentation (no counterpart in the reference): enables / disables HIP-event timing of
 * every dense-contraction launch of this handle.  Each call first reports what was recorded
 * since the previous call -- total kernel milliseconds, algorithmic FLOPs (2*M*N*K per
 * launch) and the launch count -- then resets and applies `enable`.  Synchronises. */
int fire_profile_gemm(fire_handle* h, int32_t enable, double* total_ms, double* total_flops, int64_t* n_launches);

/* Message of the last failure on this thread ("" if none). */
const char* fire_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FIRE_H */
