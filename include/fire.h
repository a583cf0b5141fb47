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
  FIRE_MIXED = 1, /* dense contractions on the 16-bit matrix cores with FP32 accumulation
                     (see fire_precision_plan): on the tangent rows of the local energy the
                     ElecInit MLP, update-MLP and orbital contractions carry each operand as two
                     FP16 planes of a power-of-two-scaled value (per row of A, per column of W;
                     three products, about 2^-22 relative each); the value-path contractions
                     carry three BF16 planes (six products, FP32-level), the Jastrow and
                     attention contractions a BF16 hi + lo pair (three products);
                     element-wise forward-Laplacian rules, envelopes and Jastrow sums FP32;
                     Slater determinants, their inverses, updates and energy sums FP64 */
  FIRE_BF16 = 2,  /* dense contractions in single BF16 on MFMA (fastest, least accurate) */
  FIRE_MIXED_FAST = 3, /* as FIRE_MIXED with two BF16 planes (hi + lo) and three products
                          (hi*hi + hi*lo + lo*hi): ~2^-17 relative per product */
};

/* Creates a handle for molecule (R [n_nuc*3] bohr, Z [n_nuc]) on HIP device `device`.
 * Replaces MoonLikeWaveFunction.create (src/sparse_wf/model/wave_function.py:102-150). */
int fire_create(const fire_config* cfg, const float* R, const float* Z, int device, fire_handle** out);
int fire_destroy(fire_handle* h);

/* Gives the handle a new molecule and model (as fire_create) while keeping its device
 * workspace, which only grows: a serving loop reuses a pool of handles across requests
 * without device allocations for buffers that are already large enough.  Waits for the
 * device; the parameters and the walker state belong to the old model and are dropped
 * (fire_upload_params is needed again).  No counterpart in the reference (a serving-side
 * extension of MoonLikeWaveFunction.create, wave_function.py:102-150). */
int fire_reconfigure(fire_handle* h, const fire_config* cfg, const float* R, const float* Z);

/* Device bytes held by the handle: workspace, walker state and parameters (host output).
 * No counterpart in the reference (admission by HBM footprint in the serving loop). */
int fire_workspace_bytes(fire_handle* h, int64_t* bytes);

/* Uploads parameters from host memory: `n_tensors` float32 tensors back to back in the
 * canonical order listed in sparse_wf_amd/params.py (pack_params), `sizes[k]` = element
 * count of tensor k; sizes are validated against the config.  The resident walker state
 * (fire_state_init) belongs to the previous parameters and is dropped: state calls fail until
 * the next fire_state_init.  Replaces the params pytree MoonLikeParams (wave_function.py:54-57)
 * handed to every apply call. */
int fire_upload_params(fire_handle* h, const float* packed, const int64_t* sizes, int32_t n_tensors);

/* Neighbour lists of B walkers r_dev [B, n_el, 3]: counts_dev [B*n_el] and, if
 * nbr_dev != NULL, the neighbour indices in CSR order (walker, centre ascending,
 * neighbour ascending) with capacity `cap`; *n_pairs receives the total.  Strict
 * d < cutoff on the float32 distance sqrt((dx*dx+dy*dy)+dz*dz).
 * Replaces get_pair_indices / is_neighbour (src/sparse_wf/model/new_sparse_model.py:238-263,
 * src/sparse_wf/model/graph_utils.py:179-184). */
int fire_neighbours(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t* counts_dev,
                    int32_t* nbr_dev, int64_t cap, int64_t* n_pairs);

/* As fire_neighbours, without any device allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: tight serving loops and harnesses that forbid
 * hipMalloc).  The CSR row pointer lives in a static device array of 16 MiB per GPU (held
 * until the process ends): B * n_el must be at most 2097150, larger batches are rejected.
 * Every call on one GPU shares that array, so calls that overlap on one GPU (any handles,
 * any streams) race; the caller orders them.  Synchronises the stream once (to read the
 * pair count). */
int fire_neighbours_noalloc(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t* counts_dev,
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

/* ---- resident walker state and single-electron low-rank updates -------------------------
 * The handle keeps one walker state: positions, the h0/message rows, embeddings, raw
 * orbitals and envelopes, FP64 Slater matrices with their inverses and log-determinants,
 * and the per-electron Jastrow rows of B walkers.  All entry points below are
 * stream-ordered on `stream`; calls that overlap on one handle race. */

/* Evaluates B walkers r_dev [B, n_el, 3] (value path) and keeps their state; logpsi_dev /
 * sign_dev (optional, double [B]) receive log|psi| and sign.  `precision` applies to every
 * later update of this state.
 * Replaces MoonLikeWaveFunction.log_psi_with_state (src/sparse_wf/model/wave_function.py:226-249). */
int fire_state_init(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision,
                    double* logpsi_dev, double* sign_dev);

/* Proposal: electron k of every walker moved to rk_dev [B, 3].  Embeddings of the electrons
 * within the cutoff of k before or after the move, the orbital rows of those electrons, the
 * envelope row of k, the Jastrow and the determinants (Woodbury, FP64) are updated; the
 * state is unchanged until fire_state_commit.  logpsi_dev / sign_dev (optional, [B])
 * receive log|psi| and sign of the proposals.  Synchronises the stream to check the static
 * capacity: a move that changes more than min(n_el, 64) electrons of a walker fails and
 * drops the state.  Chains of FP64 Woodbury updates accumulate rounding in the inverses; on
 * ill-conditioned walkers log|psi| of the state drifts from a fresh evaluation (C16H34, 50
 * sweeps without a rebuild: median 1e-9, maximum 0.1), so long chains call fire_state_init
 * on the current positions every few sweeps (the reference rebuilds at every
 * mcmc_steps_low_rank call).
 * Replaces MoonLikeWaveFunction.log_psi_low_rank_update (wave_function.py:251-287) with
 * NewSparseEmbedding.low_rank_update (new_sparse_model.py:553-629), Orbitals.low_rank_update
 * (orbitals.py:121-146), woodburry_update (model/utils.py:216-244) and
 * Jastrow.low_rank_update (jastrow.py:309-361). */
int fire_state_propose(fire_handle* h, void* stream, int32_t k, const float* rk_dev, double* logpsi_dev,
                       double* sign_dev);

/* Applies the pending proposal to the walkers with accept_dev[w] != 0 (int32 [B]).
 * Replaces the jnp.where(accept, new, old) selection of the state in mcmc_steps_low_rank
 * (src/sparse_wf/mcmc.py:240-244). */
int fire_state_commit(fire_handle* h, void* stream, const int32_t* accept_dev);

/* n_steps Metropolis steps on the state: step s moves electron s mod n_el of every walker by
 * width * N(0, 1) per coordinate and accepts when 2 (log|psi'| - log|psi|) > log u,
 * u ~ U(0, 1).  Random numbers are Philox4x32-10 keyed by (seed, walker) with counter s, so
 * a run is reproducible and independent of how steps are split over calls (pass the running
 * step index as step0).  n_accept_dev (optional, int32 [B]) receives the accepted moves of
 * each walker in this call.  Every proposal runs at static capacities (at most min(n_el, 64)
 * changed electrons per move) with row counts kept on the device, so the call is
 * asynchronous (it synchronises only when it captures a new sweep graph).  A move that
 * changes more electrons than the capacity invalidates the state; fire_state_get reports it
 * as an error.
 * Replaces mcmc_steps_low_rank with proposal_single_electron (src/sparse_wf/mcmc.py:100-111,
 * 196-275). */
int fire_mcmc_steps(fire_handle* h, void* stream, int32_t n_steps, float width, uint64_t seed, int64_t step0,
                    int32_t* n_accept_dev);

/* Options of fire_mcmc_steps (no counterpart in the reference): use_graphs = 1 (default)
 * replays whole sweeps (n_el steps starting at a multiple of n_el) from one hipGraph captured
 * on an internal stream, 0 issues every step's kernels one by one.  Results are identical. */
int fire_mcmc_options(fire_handle* h, int32_t use_graphs);

/* Copies the state's positions (float [B, n_el, 3]), log|psi| and sign (double [B]) out; any
 * may be NULL.  A pending proposal is dropped.  (No counterpart: the reference returns the
 * electrons from mcmc_steps_low_rank, mcmc.py:260-275.) */
int fire_state_get(fire_handle* h, void* stream, float* r_dev, double* logpsi_dev, double* sign_dev);

/* Counters of the state's proposals since fire_state_init: walkers, proposals (walker-moves),
 * sum and maximum of the number of changed electrons per move (the reference's static
 * statistics of low-rank updates, mcmc.py:225-231).  Host outputs; any may be NULL. */
int fire_state_stats(fire_handle* h, int32_t* n_walkers, int64_t* n_proposals, int64_t* sum_changed,
                     int64_t* max_changed);

/* Precision plan of FIRE_MIXED (no counterpart in the reference): operand planes for each
 * contraction class, in the order ElecInit MLP, update MLP, orbitals, MLP Jastrow, attention
 * Q/V, attention readout.  3: three BF16 planes, six products accumulated down to 2^-16
 * relative (FP32-level); 2: a BF16 hi + lo pair, three products; 1: plain BF16; 4: two FP16
 * planes of operands scaled by powers of two (each row of A by its maximum, which the kernel
 * producing A leaves or a row-maximum pass takes, each column of W at upload), three products,
 * about 2^-22 relative per product; the value path and the sampler's proposals take the row
 * maxima from the rows themselves, so a moved electron's rows match a fresh evaluation.  Default
 * {4,4,4,2,2,2}.  Sets the first n classes from planes (n may be 0); current (optional,
 * int32 [6]) receives the plan. */
int fire_precision_plan(fire_handle* h, const int32_t* planes, int32_t n, int32_t* current);

/* Precision of the determinant Laplacian's triplet products B_x = J_x A^-1[:, S] (no counterpart
 * in the reference, which runs the whole model in one precision): fp64_laplacian = 1 multiplies
 * the FP32 orbital jacobian rows, widened, with an FP64 transposed copy of the FP64 inverse on the
 * FP64 matrix cores; 0 uses an FP32 copy of the inverse on the FP32 matrix cores.  The inverse
 * itself, the trace with the Laplacian rows and every sum are FP64 either way.  The FP64 form
 * needs n_el even and n_el * n_det a multiple of 4 (otherwise the FP32 form runs).  Applies to
 * every precision mode. */
int fire_det_options(fire_handle* h, int32_t fp64_laplacian);

/* Tile-shape autotuner of the dense contractions (no counterpart in the reference; the
 * reference leaves kernel selection to XLA).  On by default: the first BF16 contraction of
 * each (split, N, K, floor(log2 M)) shape times every tile shape on its real operands
 * (synchronising) and keeps the fastest for the life of the process, shared by every handle
 * on the same device; a shape with a tuned neighbour of the same split, K, floor(log2 M) and
 * number of 128-column tiles takes the nearest neighbour's choice without timing.  enable: 1 on,
 * 0 off (tile 128 x 128), 2 on and forget earlier choices (this device's), 3 on and also reuse
 * a neighbour up to two row buckets away (for ragged and serving batches, whose geometries each
 * bring new shapes), 100 + v every contraction of this handle on tile form v (a diagnostic: where
 * form v does not apply to a contraction, the base tile; modes 0-3 end it), -1 unchanged.  table (optional, int32 [cap][7]) receives this handle's
 * choices as (split, N, K, log2 M, variant, BM, BN); log2 M is raised by 32 for row counts held
 * on the device and by 64 for contractions with the fused activation epilogue;
 * *n_entries the number of choices. */
int fire_autotune(fire_handle* h, int32_t enable, int32_t* table, int32_t cap, int32_t* n_entries);

/* Number of tile forms of the split contractions: fire_autotune's mode 100 + v takes
 * v = 0 .. fire_tile_forms() - 1 (no counterpart in the reference). */
int fire_tile_forms(void);

/* Instrumentation (no counterpart in the reference): enables / disables HIP-event timing of
 * every dense-contraction launch of this handle.  Each call first reports what was recorded
 * since the previous call -- total kernel milliseconds, algorithmic FLOPs (2*M*N*K per
 * launch), algorithmic bytes (4*M*(K + N) for A read and C written once, plus the weight
 * planes once) and the launch count -- then resets and applies `enable`.  Synchronises. */
int fire_profile_gemm(fire_handle* h, int32_t enable, double* total_ms, double* total_flops, double* total_bytes,
                      int64_t* n_launches);

/* ---- pseudopotentials ----------------------------------------------------------------------
 * Effective-core potentials of n_ecp nuclei (atom[e] indexes the molecule's nuclei), tabulated
 * on one ascending radial grid (double [n_grid]): the local channel v_loc [n_ecp][n_grid] and the
 * non-local channels l = 0 .. n_l - 1 (n_l <= 4) v_nl [n_ecp][n_l][n_grid], both already divided by
 * r^2 as eval_ecp_on_grid does; cutoff [n_ecp] is the radius beyond which the non-local channels
 * are neglected, n_quad [n_ecp] the points of the atom's spherical quadrature and quad_dirs
 * [n_ecp][q_max][3] its unit directions (weights 1 / n_quad).  z_eff (optional, [n_nuc]) replaces
 * the nuclear charges of the Coulomb terms of fire_local_energy (Z - core electrons).  n_ecp = 0
 * removes the pseudopotentials.  Synchronises the device.
 * Replaces make_pseudopotential / eval_ecp_on_grid (src/sparse_wf/pseudopotentials.py:136-202,
 * 411-454) given the ECP coefficients (the reference loads them with pyscf). */
int fire_ecp_upload(fire_handle* h, int32_t n_ecp, const int32_t* atom, int32_t n_l, int32_t n_grid,
                    const double* grid, const double* v_loc, const double* v_nl, const double* cutoff,
                    const int32_t* n_quad, int32_t q_max, const double* quad_dirs, const float* z_eff);

/* Pseudopotential energy of the walkers of the resident state (fire_state_init): v_loc_dev
 * (double [B]) the local part, sum over electrons and ECP nuclei of the interpolated local
 * channel; v_nl_dev (double [B]) the non-local part: each electron within the cutoff of its
 * closest ECP nucleus is placed on that nucleus's quadrature, rotated at random (Philox4x32-10
 * normals under (seed, walker) and counter (electron, 0x45435000 + block)), at its own distance;
 * each point adds sum_l (2l+1) V_l(r') P_l(cos theta) / n_quad times psi(r')/psi(r), the ratio
 * from a low-rank proposal of the state (one per electron and quadrature point for all walkers,
 * each a fire_state_propose).  *n_points (optional, host) receives the number of active
 * integration points.  The state is unchanged; proposals are counted in fire_state_stats.
 * Synchronises the stream (the move's capacity check, the active points).
 * Replaces the pp_local / pp_nonloc calls of make_local_energy (src/sparse_wf/hamiltonian.py:
 * 69-81) with make_local_pseudopotential, build_all_integration_points and
 * make_nonlocal_pseudopotential (pseudopotentials.py:204-218, 269-398). */
int fire_ecp_energy(fire_handle* h, void* stream, uint64_t seed, double* v_loc_dev, double* v_nl_dev,
                    int64_t* n_points);

/* Statistics and clipping of a batch of local energies e_loc_dev (double [B], B > 0) on the
 * handle's device (h may be NULL: the calling thread's current device), one workgroup:
 * stats_dev (double [5]) = E_mean, E_std (population), clip centre, mean absolute deviation
 * from it, mean of the clipped energies.  With clip_local_energy > 0 the energies are clipped
 * to centre +- clip_local_energy * deviation, the centre being the median (stat = 0, the
 * midpoint of the elements of rank (B-1)/2 and B/2, selected exactly) or the mean (stat = 1);
 * with clip_local_energy <= 0 nothing is clipped and centre and deviation are NaN.
 * e_clipped_dev (double [B], optional) receives the clipped energies, e_diff_dev (optional) the
 * clipped energies minus their mean (dloss/dlogpsi of the energy loss).  A NaN energy makes
 * the mean, median and deviation NaN and stays NaN when clipped.  Sums are deterministic.
 * Replaces clip_local_energies / local_energy_diff (src/sparse_wf/update.py:45-69) and the
 * energy mean / std of Trainer.step (update.py:166-168), on one GPU's batch. */
int fire_energy_stats(fire_handle* h, void* stream, const double* e_loc_dev, int32_t B, double clip_local_energy,
                      int32_t stat, double* e_clipped_dev, double* e_diff_dev, double* stats_dev);

/* Per-walker gradient of log|psi| with respect to the uploaded parameters: grad_dev (float
 * [B][n], n from fire_param_count) holds, for each walker of r_dev (float [B][N][3]),
 * d log|psi| / d theta for every float of the packed parameter tensors, in upload order
 * (fire_upload_params), each tensor in its packed layout (folded scales, softplus'd envelope
 * widths, transposed pi); logpsi_dev (double [B], optional) receives log|psi|.  The forward
 * pass runs in FP32 with FP64 determinants (Gauss-Jordan with partial pivoting); the backward
 * pass is the reverse mode of that value path, determinants through their FP64 inverses.
 * Walkers are processed in chunks; synchronises the stream; B = 0 does nothing.  Replaces the
 * per-walker jax.grad(log_p) of the natural-gradient preconditioners
 * (src/sparse_wf/preconditioner.py:191-194, 290-294, 382-383); the adjoint of the packing to
 * the reference's parameter tree is host work. */
int fire_grad_logpsi(fire_handle* h, void* stream, const float* r_dev, int32_t B, float* grad_dev,
                     double* logpsi_dev);

/* *n_floats = total number of floats of the packed parameter tensors (the row length of
 * fire_grad_logpsi's output). */
int fire_param_count(fire_handle* h, int64_t* n_floats);

/* Message of the last failure on this thread ("" if none). */
const char* fire_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FIRE_H */
