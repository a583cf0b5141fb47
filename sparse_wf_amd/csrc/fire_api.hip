// Host side of the C ABI (include/fire.h): handle, parameters, workspace and the
// stream-ordered launch sequence of the FiRE hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fire.h"
#include "fire_kernels.hpp"
#include "fire_kernels2.hpp"

namespace {

thread_local std::string g_last_error;

struct FireError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
      throw FireError(std::string(#x) + " failed: " + hipGetErrorString(e_) + " at " __FILE__ ":" + \
                      std::to_string(__LINE__));                                                   \
  } while (0)

#define REQUIRE(c, msg) \
  do {                  \
    if (!(c)) throw FireError(msg); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t n) {
    if (n <= bytes) return;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    bytes = 0;
    size_t want = std::max<size_t>(n + n / 8, 256);
    HIPCHK(hipMalloc(&p, want));
    bytes = want;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

uint16_t host_f2bf(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

}  // namespace

struct fire_handle {
  fire_config cfg;
  int device = 0;
  int N, n_up, n_dn, M, F, w0, w1, Een, D, Eiso, ND, W, H, Dk, HD, depth;
  bool has_mlp, has_att;
  int cusp;
  double e_nn = 0.0;
  float* R = nullptr;
  float* Z = nullptr;
  // parameters
  std::vector<int64_t> expect;  // element counts in canonical order
  std::vector<float*> P;
  std::vector<uint16_t*> Pbf;   // BF16 transposed copies of GEMM weights (null if none): hi part
  std::vector<uint16_t*> Pbf_lo;  // ... and the BF16 remainder W - hi
  std::vector<float*> Pt32;       // FP32 transposed copies [N,K] of GEMM weights
  std::vector<std::pair<int, int>> shape2;  // (K, N) of GEMM weights
  float cut_learned = 0.f;
  float s1 = 1.f;
  bool params_ok = false;
  // tensor indices
  int iEN, iEEs, iEEd, iMLP, iORB, iJM, iATT, iCUSP, iSCAL;
  // workspace
  DevBuf cnt, ptr, nbr, rev, rowflag;
  DevBuf T5, Z2, T5b, M5, XA, XB, XC, O, ENV, ENVC;
  DevBuf A64, Ainv, AinvT, ldet, dsgn, tr, trip, dgrad;
  DevBuf det_lp, det_sg, det_gr, det_lap;
  DevBuf JR, mlp_ns, mlp_dg, QV, AV, att_ns, att_dg, RIN, RY1, RY2, att_out;
  DevBuf cval, clap, cgrad, epot;
  int chunk = 256;
  size_t smem_set = 0;
  bool lu_smem_set = false;
  // optional HIP-event timing of the dense contractions (fire_profile_gemm)
  bool gemm_timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> gemm_events;
  double gemm_flops = 0.0;
};

namespace {

using namespace fire;

void build_expect(fire_handle* h) {
  auto& e = h->expect;
  e.clear();
  const int64_t M = h->M, F = h->F, w0 = h->w0, w1 = h->w1, E = h->Een, D = h->D, Ei = h->Eiso, N = h->N,
                ND = h->ND, W = h->W, HD = h->HD;
  h->iEN = (int)e.size();
  for (int64_t v : {M * 4 * w0, M * w0, M * E, M * F, w0 * w1, w1, E * w1, w1 * F, 4 * F, F, F, F * 2 * F, F * F, F,
                    F * 5 * F, 5 * F})
    e.push_back(v);
  h->iEEs = (int)e.size();
  for (int k = 0; k < 2; ++k) {
    if (k == 1) h->iEEd = (int)e.size();
    for (int64_t v : {4 * w0, w0, w0, w0 * w1, w1, w0 * w1, w1 * F, 4 * F, F}) e.push_back(v);
  }
  h->iMLP = (int)e.size();
  for (int64_t v : {F * F, F, F * F, F}) e.push_back(v);
  h->iORB = (int)e.size();
  for (int64_t v : {F * ND, ND, F * ND, ND, M * D * Ei, M * D * Ei * N, M * D * Ei, M * D * Ei * N}) e.push_back(v);
  h->iJM = (int)e.size();
  if (h->has_mlp) {
    int64_t fan = F;
    for (int l = 0; l < h->depth; ++l) {
      e.push_back(fan * W);
      e.push_back(W);
      fan = W;
    }
    e.push_back(W * 2);
    e.push_back(2);
    e.push_back(2);  // mlp_scale
    e.push_back(2);  // [0, log_bias]
  }
  h->iATT = (int)e.size();
  if (h->has_att) {
    e.push_back(F * 2 * HD);
    e.push_back(2 * HD);
    e.push_back(HD);  // register keys
    int64_t fan = 2 * HD;
    for (int l = 0; l < h->depth; ++l) {
      e.push_back(fan * W);
      e.push_back(W);
      fan = W;
    }
    e.push_back(W * 2);
    e.push_back(2);
    e.push_back(2);  // scale
    e.push_back(2);  // [0, bias]
  }
  h->iCUSP = (int)e.size();
  if (h->cusp) e.push_back(4);
  h->iSCAL = (int)e.size();
  e.push_back(2);  // learned cutoff, output scale s1 of the embedding
}

// GEMM weights get a BF16 transposed copy for the mixed-precision path.
void gemm_weights(fire_handle* h, std::vector<std::pair<int, std::pair<int, int>>>& out) {
  const int F = h->F, W = h->W, HD = h->HD, ND = h->ND;
  out.push_back({h->iEN + 11, {F, 2 * F}});
  out.push_back({h->iEN + 12, {F, F}});
  out.push_back({h->iEN + 14, {F, 5 * F}});
  out.push_back({h->iMLP + 0, {F, F}});
  out.push_back({h->iMLP + 2, {F, F}});
  out.push_back({h->iORB + 0, {F, ND}});
  out.push_back({h->iORB + 2, {F, ND}});
  if (h->has_mlp) {
    int fan = F;
    for (int l = 0; l < h->depth; ++l) {
      out.push_back({h->iJM + 2 * l, {fan, W}});
      fan = W;
    }
  }
  if (h->has_att) {
    out.push_back({h->iATT + 0, {F, 2 * HD}});
    int fan = 2 * HD;
    for (int l = 0; l < h->depth; ++l) {
      out.push_back({h->iATT + 3 + 2 * l, {fan, W}});
      fan = W;
    }
  }
}

struct Ctx {
  fire_handle* h;
  hipStream_t s;
  int prec;
};

// C = A @ W(tensor wi) + bias(tensor bi, optional) on value rows
void gemm(Ctx& c, const float* A, int64_t lda, int wi, int bi, float* C, int64_t ldc, int64_t M, int N, int K,
          const uint8_t* rowflag, int period) {
  fire_handle* h = c.h;
  const float* bias = bi >= 0 ? h->P[bi] : nullptr;
  const bool bf_ok = h->Pbf[wi] != nullptr && (K % 32) == 0 && (lda % 4) == 0 && N > 16;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->gemm_timing) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c.s));
  }
  if ((c.prec == FIRE_MIXED || c.prec == FIRE_BF16) && bf_ok)
    launch_gemm_bf16(c.s, A, lda, h->Pbf[wi], h->Pbf_lo[wi], bias, C, ldc, M, N, K, rowflag, period,
                     c.prec == FIRE_MIXED ? 3 : 1);
  else
    launch_gemm_f32(c.s, A, lda, h->P[wi], h->Pt32[wi], bias, C, ldc, M, N, K, rowflag, period);
  if (h->gemm_timing) {
    HIPCHK(hipEventRecord(e1, c.s));
    h->gemm_events.push_back({e0, e1});
    h->gemm_flops += 2.0 * (double)M * N * K;
  }
}

EnParams en_params(fire_handle* h) {
  const int b = h->iEN;
  EnParams P;
  P.K = h->P[b + 0];
  P.b = h->P[b + 1];
  P.sp = h->P[b + 2];
  P.hnuc = h->P[b + 3];
  P.Wd = h->P[b + 4];
  P.bd = h->P[b + 5];
  P.We = h->P[b + 6];
  P.WG = h->P[b + 7];
  P.WF = h->P[b + 8];
  P.ln_g = h->P[b + 9];
  P.ln_b = h->P[b + 10];
  P.w0 = h->w0;
  P.w1 = h->w1;
  P.E = h->Een;
  return P;
}

EdgeParams edge_params(fire_handle* h, int b) {
  EdgeParams P;
  P.filt.K = h->P[b + 0];
  P.filt.b = h->P[b + 1];
  P.filt.sp = h->P[b + 2];
  P.filt.Wd = h->P[b + 3];
  P.filt.bd = h->P[b + 4];
  P.filt.We = h->P[b + 5];
  P.filt.w0 = h->w0;
  P.filt.w1 = h->w1;
  P.filt.E = h->w0;
  P.WG = h->P[b + 6];
  P.WF = h->P[b + 7];
  P.BF = h->P[b + 8];
  return P;
}

struct Graph {
  int64_t n_pairs = 0;
  int max_nb = 0;
};

Graph build_graph(Ctx& c, const float* r, int B) {
  fire_handle* h = c.h;
  const int N = h->N, G = B * N;
  h->cnt.ensure(sizeof(int32_t) * G);
  h->ptr.ensure(sizeof(int64_t) * (G + 2));
  launch_nbr_count(c.s, r, B, N, h->n_up, h->cfg.cutoff, h->cnt.as<int32_t>());
  launch_scan(c.s, h->cnt.as<int32_t>(), G, h->ptr.as<int64_t>());
  int64_t tail[2];
  HIPCHK(hipMemcpyAsync(tail, h->ptr.as<int64_t>() + G, sizeof(tail), hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  Graph g;
  g.n_pairs = tail[0];
  g.max_nb = (int)tail[1];
  h->nbr.ensure(sizeof(int32_t) * std::max<int64_t>(g.n_pairs, 1));
  h->rev.ensure(sizeof(int32_t) * std::max<int64_t>(g.n_pairs, 1));
  launch_nbr_fill(c.s, r, B, N, h->n_up, h->cfg.cutoff, h->ptr.as<int64_t>(), h->nbr.as<int32_t>());
  launch_rev(c.s, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), B, N, h->n_up, h->rev.as<int32_t>());
  return g;
}

// Embedding (value + tangent rows when tan): result in h->XA with group layout gd.
void run_embedding(Ctx& c, const float* r, int B, const Graph& gr, bool tan, GroupDesc& gd, int64_t& R) {
  fire_handle* h = c.h;
  const int N = h->N, F = h->F, G = B * N;
  const int NR = tan ? 5 : 1;
  h->T5.ensure(sizeof(float) * (size_t)G * NR * F);
  h->T5b.ensure(sizeof(float) * (size_t)G * NR * F);
  h->Z2.ensure(sizeof(float) * (size_t)G * NR * 2 * F);
  h->M5.ensure(sizeof(float) * (size_t)G * NR * 5 * F);
  const int b = h->iEN;
  launch_elec_init_en(c.s, r, h->R, B, N, h->n_up, h->M, h->cfg.cutoff_1el, F, en_params(h), tan, h->T5.as<float>());
  const int64_t rows5 = (int64_t)G * NR;
  gemm(c, h->T5.as<float>(), F, b + 11, -1, h->Z2.as<float>(), 2 * F, rows5, 2 * F, F, nullptr, 0);
  launch_glu(c.s, h->Z2.as<float>(), G, F, tan, h->T5b.as<float>());
  gemm(c, h->T5b.as<float>(), F, b + 12, b + 13, h->T5.as<float>(), F, rows5, F, F, nullptr, NR);
  GroupDesc g5{NR, 0, nullptr};
  launch_act(c.s, h->T5.as<float>(), F, F, G, g5, ACT_SILU, 0.f);
  gemm(c, h->T5.as<float>(), F, b + 14, b + 15, h->M5.as<float>(), 5 * F, rows5, 5 * F, F, nullptr, NR);

  if (tan) {
    gd = GroupDesc{5, 3, h->ptr.as<int64_t>()};
    R = 5 * (int64_t)G + 3 * gr.n_pairs;
  } else {
    gd = GroupDesc{1, 0, nullptr};
    R = G;
  }
  h->XA.ensure(sizeof(float) * (size_t)R * F);
  h->XB.ensure(sizeof(float) * (size_t)R * F);
  launch_messages(c.s, r, B, N, h->n_up, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), h->M5.as<float>(),
                  edge_params(h, h->iEEs), edge_params(h, h->iEEd), h->cut_learned, F, tan, h->XA.as<float>());
  const uint8_t* flag = nullptr;
  int period = 1;
  if (tan) {
    h->rowflag.ensure((size_t)R);
    launch_rowflag(c.s, G, gd, R, h->rowflag.as<uint8_t>());
    flag = h->rowflag.as<uint8_t>();
    period = 0;
  }
  const int m = h->iMLP;
  gemm(c, h->XA.as<float>(), F, m + 0, m + 1, h->XB.as<float>(), F, R, F, F, flag, period);
  launch_act(c.s, h->XB.as<float>(), F, F, G, gd, ACT_SILU, 0.f);
  gemm(c, h->XB.as<float>(), F, m + 2, m + 3, h->XA.as<float>(), F, R, F, F, flag, period);
  launch_act(c.s, h->XA.as<float>(), F, F, G, gd, ACT_SILU, 0.f);
}

struct DetOut {
  double *logpsi, *sign, *grad, *lap;
};

void run_orbitals_det(Ctx& c, const float* r, int B, const Graph& gr, bool tan, const GroupDesc& gd, int64_t R,
                      DetOut& out) {
  fire_handle* h = c.h;
  const int N = h->N, F = h->F, G = B * N, ND = h->ND, D = h->D;
  const int NR = tan ? 5 : 1;
  const size_t ocols = std::max(ND, h->has_att ? 2 * h->HD : 0);
  h->O.ensure(sizeof(float) * (size_t)R * ocols);
  h->ENV.ensure(sizeof(float) * (size_t)G * NR * ND);
  const int o = h->iORB;
  const int64_t R_up = gd.a * (int64_t)B * h->n_up;  // rows of up-spin groups (value part)
  int64_t R_up_rows = R_up;
  if (gd.b) {
    int64_t p_up;
    HIPCHK(hipMemcpyAsync(&p_up, h->ptr.as<int64_t>() + (size_t)B * h->n_up, sizeof(int64_t), hipMemcpyDeviceToHost,
                          c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    R_up_rows = R_up + 3 * p_up;
  }
  const uint8_t* flag = tan ? h->rowflag.as<uint8_t>() : nullptr;
  const int period = tan ? 0 : 1;
  float* Ob = h->O.as<float>();
  gemm(c, h->XA.as<float>(), F, o + 0, o + 1, Ob, ND, R_up_rows, ND, F, flag, period);
  if (R > R_up_rows)
    gemm(c, h->XA.as<float>() + (size_t)R_up_rows * F, F, o + 2, o + 3, Ob + (size_t)R_up_rows * ND, ND,
         R - R_up_rows, ND, F, flag ? flag + R_up_rows : nullptr, period);
  // envelopes: coefficients per (electron, row, nucleus, envelope), then one FP32 GEMM per
  // (spin block, determinant) against pi stored as [det][nucleus * envelope][orbital]
  {
    const int ME = h->M * h->Eiso;
    const int64_t rows_all = (int64_t)G * NR, rows_up = (int64_t)B * h->n_up * NR;
    h->ENVC.ensure(sizeof(float) * (size_t)D * rows_all * ME);
    launch_env_coef(c.s, r, h->R, B, N, h->n_up, h->M, D, h->Eiso, h->P[o + 4], h->P[o + 6], tan,
                    h->ENVC.as<float>());
    for (int sp = 0; sp < 2; ++sp) {
      const int64_t r0 = sp == 0 ? 0 : rows_up, nr = sp == 0 ? rows_up : rows_all - rows_up;
      if (nr == 0) continue;
      const float* pi = h->P[o + (sp == 0 ? 5 : 7)];
      const float* piT = h->Pt32[o + (sp == 0 ? 5 : 7)];
      for (int d = 0; d < D; ++d)
        launch_gemm_f32(c.s, h->ENVC.as<float>() + ((size_t)d * rows_all + r0) * ME, ME, pi + (size_t)d * ME * N,
                        piT + (size_t)d * ME * N, nullptr, h->ENV.as<float>() + (size_t)r0 * ND + (size_t)d * N, ND,
                        nr, N, ME, nullptr, 0);
    }
  }
  launch_orb_product(c.s, Ob, ND, G, gd, h->ENV.as<float>(), tan);
  const size_t nmat = (size_t)B * D;
  h->A64.ensure(sizeof(double) * nmat * N * N);
  h->ldet.ensure(sizeof(double) * nmat);
  h->dsgn.ensure(sizeof(double) * nmat);
  launch_gather_A(c.s, Ob, ND, gd, B, N, h->n_up, D, h->A64.as<double>());
  h->det_lp.ensure(sizeof(double) * B);
  h->det_sg.ensure(sizeof(double) * B);
  out.logpsi = h->det_lp.as<double>();
  out.sign = h->det_sg.as<double>();
  const bool lu_lds = lu_lds_bytes(N) <= 160 * 1024;
  if (lu_lds && lu_lds_bytes(N) > 64 * 1024 && !h->lu_smem_set) {
    HIPCHK(lu_set_smem_limit(lu_lds_bytes(N)));
    h->lu_smem_set = true;
  }
  if (tan) {
    h->Ainv.ensure(sizeof(double) * nmat * N * N);
    h->tr.ensure(sizeof(double) * nmat);
    h->trip.ensure(sizeof(double) * nmat * N);
    h->dgrad.ensure(sizeof(double) * nmat * N * 3);
    h->det_gr.ensure(sizeof(double) * (size_t)B * 3 * N);
    h->det_lap.ensure(sizeof(double) * B);
    h->AinvT.ensure(sizeof(float) * nmat * N * N);
    launch_lu_inv(c.s, h->A64.as<double>(), (int)nmat, N, h->Ainv.as<double>(), h->AinvT.as<float>(),
                  h->ldet.as<double>(), h->dsgn.as<double>(), true, lu_lds);
    launch_det_trace(c.s, Ob, ND, gd, B, N, h->n_up, D, h->Ainv.as<double>(), h->tr.as<double>());
    const int smax = gr.max_nb + 1;
    REQUIRE(smax <= FIRE_SMAX, "an electron has more than 63 neighbours within the cutoff (determinant Laplacian limit)");
    const size_t sm = det_lap_smem(smax);
    if (sm > 64 * 1024 && sm > h->smem_set) {
      HIPCHK(det_lap_set_smem_limit(sm));
      h->smem_set = sm;
    }
    launch_det_lap(c.s, Ob, ND, gd, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), h->rev.as<int32_t>(), B, N, h->n_up, D,
                   smax, h->AinvT.as<float>(), h->trip.as<double>(), h->dgrad.as<double>());
    out.grad = h->det_gr.as<double>();
    out.lap = h->det_lap.as<double>();
  } else {
    launch_lu_inv(c.s, h->A64.as<double>(), (int)nmat, N, nullptr, nullptr, h->ldet.as<double>(),
                  h->dsgn.as<double>(), false, lu_lds);
    out.grad = nullptr;
    out.lap = nullptr;
  }
  launch_det_combine(c.s, B, N, D, h->ldet.as<double>(), h->dsgn.as<double>(), tan ? h->tr.as<double>() : nullptr,
                     tan ? h->trip.as<double>() : nullptr, tan ? h->dgrad.as<double>() : nullptr, out.logpsi, out.sign,
                     out.grad, out.lap, tan);
}

void run_jastrow_and_assemble(Ctx& c, const float* r, int B, bool tan, const GroupDesc& gd, int64_t R,
                              const DetOut& det, double* e_loc, double* e_kin, double* e_pot, double* logpsi,
                              double* sign, double* grad, double* lap) {
  fire_handle* h = c.h;
  const int N = h->N, F = h->F, G = B * N, W = h->W;
  const uint8_t* flag = tan ? h->rowflag.as<uint8_t>() : nullptr;
  const int period = tan ? 0 : 1;
  AssembleArgs A{};
  A.B = B;
  A.N = N;
  A.use_mlp = h->cfg.use_mlp_jastrow;
  A.use_log = h->cfg.use_log_jastrow;
  A.use_att = h->has_att;
  A.cusp = h->cusp;
  A.tangents = tan;
  A.det_logpsi = det.logpsi;
  A.det_sign = det.sign;
  A.det_grad = det.grad;
  A.det_lap = det.lap;
  if (h->has_mlp) {
    h->XC.ensure(sizeof(float) * (size_t)R * W);
    h->XB.ensure(sizeof(float) * (size_t)R * std::max(W, F));
    h->JR.ensure(sizeof(float) * (size_t)R * 2);
    h->mlp_ns.ensure(sizeof(double) * (size_t)B * 8);
    const int j = h->iJM;
    float* cur = h->XB.as<float>();
    float* nxt = h->XC.as<float>();
    gemm(c, h->XA.as<float>(), F, j + 0, j + 1, cur, W, R, W, F, flag, period);
    launch_act(c.s, cur, W, W, G, gd, ACT_SILU, 0.f);
    for (int l = 1; l < h->depth; ++l) {
      gemm(c, cur, W, j + 2 * l, j + 2 * l + 1, nxt, W, R, W, W, flag, period);
      launch_act(c.s, nxt, W, W, G, gd, ACT_SILU, 0.f);
      std::swap(cur, nxt);
    }
    const int jl = j + 2 * h->depth;
    launch_jmlp_last(c.s, cur, W, G, gd, h->P[jl], h->P[jl + 1], h->P[jl + 2], h->P[jl + 3], h->JR.as<float>());
    launch_node_sums(c.s, h->JR.as<float>(), 2, 2, gd, B, N, h->n_up, h->mlp_ns.as<double>());
    A.mlp_ns = h->mlp_ns.as<double>();
    if (tan) {
      h->mlp_dg.ensure(sizeof(double) * (size_t)B * 2 * N * 3 * 2);
      launch_dep_gather(c.s, h->JR.as<float>(), 2, 2, gd, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(),
                        h->rev.as<int32_t>(), B, N, h->n_up, h->mlp_dg.as<double>());
      A.mlp_dg = h->mlp_dg.as<double>();
    }
  }
  if (h->has_att) {
    const int HD = h->HD, H = h->H, ac = h->iATT;
    const int nc = H + HD;
    float* QV = h->O.as<float>();  // orbital rows are no longer needed
    h->AV.ensure(sizeof(float) * (size_t)R * nc);
    h->att_ns.ensure(sizeof(double) * (size_t)B * 2 * 2 * nc);
    gemm(c, h->XA.as<float>(), F, ac + 0, ac + 1, QV, 2 * HD, R, 2 * HD, F, flag, period);
    launch_att_group(c.s, QV, H, h->Dk, G, gd, h->P[ac + 2], tan, h->AV.as<float>());
    launch_node_sums(c.s, h->AV.as<float>(), nc, nc, gd, B, N, h->n_up, h->att_ns.as<double>());
    const int nrow = tan ? 2 + 3 * N : 1;
    if (tan) {
      h->att_dg.ensure(sizeof(double) * (size_t)B * 2 * N * 3 * nc);
      launch_dep_gather(c.s, h->AV.as<float>(), nc, nc, gd, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(),
                        h->rev.as<int32_t>(), B, N, h->n_up, h->att_dg.as<double>());
    }
    h->RIN.ensure(sizeof(float) * (size_t)B * nrow * 2 * HD);
    h->RY1.ensure(sizeof(float) * (size_t)B * nrow * W);
    h->RY2.ensure(sizeof(float) * (size_t)B * nrow * W);
    h->att_out.ensure(sizeof(float) * (size_t)B * nrow * 2);
    launch_att_readin(c.s, h->att_ns.as<double>(), tan ? h->att_dg.as<double>() : nullptr, B, N, H, h->Dk, tan,
                      h->RIN.as<float>());
    GroupDesc gw{nrow, 0, nullptr};
    const int64_t RR = (int64_t)B * nrow;
    float* cur = h->RY1.as<float>();
    float* nxt = h->RY2.as<float>();
    gemm(c, h->RIN.as<float>(), 2 * HD, ac + 3, ac + 4, cur, W, RR, W, 2 * HD, nullptr, nrow);
    launch_act(c.s, cur, W, W, B, gw, ACT_SILU, 0.f);
    for (int l = 1; l < h->depth; ++l) {
      gemm(c, cur, W, ac + 3 + 2 * l, ac + 4 + 2 * l, nxt, W, RR, W, W, nullptr, nrow);
      launch_act(c.s, nxt, W, W, B, gw, ACT_SILU, 0.f);
      std::swap(cur, nxt);
    }
    const int al = ac + 3 + 2 * h->depth;
    launch_gemm_f32(c.s, cur, W, h->P[al], nullptr, h->P[al + 1], h->att_out.as<float>(), 2, RR, 2, W, nullptr,
                    nrow);
    A.att_out = h->att_out.as<float>();
    A.att_scale = h->P[al + 2];
    A.att_zb = h->P[al + 3];
  }
  h->cval.ensure(sizeof(double) * (size_t)G);
  h->clap.ensure(sizeof(double) * (size_t)G);
  h->cgrad.ensure(sizeof(double) * (size_t)G * 3);
  h->epot.ensure(sizeof(double) * (size_t)G);
  launch_cusp_coulomb(c.s, r, h->R, h->Z, B, N, h->n_up, h->M, h->cusp, h->cusp ? h->P[h->iCUSP] : nullptr,
                      h->cval.as<double>(), h->clap.as<double>(), h->cgrad.as<double>(), h->epot.as<double>());
  A.cval = h->cval.as<double>();
  A.clap = h->clap.as<double>();
  A.cgrad = h->cgrad.as<double>();
  A.epot = h->epot.as<double>();
  A.e_nn = h->e_nn;
  A.e_loc = e_loc;
  A.e_kin = e_kin;
  A.e_pot = e_pot;
  A.logpsi = logpsi;
  A.sign = sign;
  A.grad = grad;
  A.lap = lap;
  launch_assemble(c.s, A);
}

int fail(const std::exception& e) {
  g_last_error = e.what();
  return 1;
}

void check_ready(fire_handle* h) {
  REQUIRE(h != nullptr, "null handle");
  REQUIRE(h->params_ok, "parameters not uploaded (call fire_upload_params first)");
}

}  // namespace

extern "C" {

const char* fire_last_error(void) { return g_last_error.c_str(); }

int fire_create(const fire_config* cfg, const float* R, const float* Z, int device, fire_handle** out) {
  try {
    REQUIRE(cfg && R && Z && out, "null argument");
    const fire_config& c = *cfg;
    REQUIRE(c.n_el > 0 && c.n_up >= 0 && c.n_up <= c.n_el && c.n_nuc > 0, "invalid electron / nucleus counts");
    REQUIRE(c.feature_dim % 64 == 0 && c.feature_dim <= 256,
            "feature_dim must be a multiple of 64, at most 256 (one feature column per thread of a 256-thread "
            "workgroup in the embedding kernels)");
    REQUIRE(!c.use_attention || (c.att_dim <= 64 && (c.att_dim & (c.att_dim - 1)) == 0 && c.att_heads * c.att_dim <= 256),
            "attention_dim must be a power of two <= 64 and attention_heads * attention_dim <= 256");
    REQUIRE(c.filter_w0 > 0 && c.filter_w0 <= FIRE_W0MAX &&
                (c.filter_w1 == 4 || c.filter_w1 == 8 || c.filter_w1 == 16),
            "pair_mlp_widths out of range (w0 <= 32; w1 zero-padded to 4, 8 or 16)");
    REQUIRE(c.n_env_en > 0 && c.n_env_en <= FIRE_EEMAX, "pair_n_envelopes out of range (<= 64)");
    REQUIRE(c.n_det > 0 && c.n_det <= 16 && c.n_el * c.n_det <= 2048,
            "n_determinants must be <= 16 and n_el * n_determinants <= 2048");
    REQUIRE(c.n_det * c.n_env_iso <= 1600, "n_determinants * n_envelopes must be <= 1600");
    REQUIRE(c.e_e_cusp == 0 || c.e_e_cusp == 1, "e_e_cusp must be 0 (none) or 1 (psiformer)");

    REQUIRE(c.mlp_depth >= 1 && c.mlp_width > 0, "mlp_depth >= 1 and mlp_width > 0 required");
    HIPCHK(hipSetDevice(device));
    fire_handle* h = new fire_handle();
    h->cfg = c;
    h->device = device;
    h->N = c.n_el;
    h->n_up = c.n_up;
    h->n_dn = c.n_el - c.n_up;
    h->M = c.n_nuc;
    h->F = c.feature_dim;
    h->w0 = c.filter_w0;
    h->w1 = c.filter_w1;
    h->Een = c.n_env_en;
    h->D = c.n_det;
    h->Eiso = c.n_env_iso;
    h->ND = c.n_el * c.n_det;
    h->W = c.mlp_width;
    h->H = c.att_heads;
    h->Dk = c.att_dim;
    h->HD = c.att_heads * c.att_dim;
    h->depth = c.mlp_depth;
    h->has_mlp = c.use_mlp_jastrow || c.use_log_jastrow;
    h->has_att = c.use_attention != 0;
    h->cusp = c.e_e_cusp;
    double enn = 0.0;
    for (int I = 0; I < c.n_nuc; ++I)
      for (int J = I + 1; J < c.n_nuc; ++J) {
        const double dx = R[3 * I] - R[3 * J], dy = R[3 * I + 1] - R[3 * J + 1], dz = R[3 * I + 2] - R[3 * J + 2];
        enn += (double)Z[I] * Z[J] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    h->e_nn = enn;
    HIPCHK(hipMalloc(&h->R, sizeof(float) * 3 * c.n_nuc));
    HIPCHK(hipMalloc(&h->Z, sizeof(float) * c.n_nuc));
    HIPCHK(hipMemcpy(h->R, R, sizeof(float) * 3 * c.n_nuc, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->Z, Z, sizeof(float) * c.n_nuc, hipMemcpyHostToDevice));
    build_expect(h);
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_destroy(fire_handle* h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  for (float* p : h->P)
    if (p) (void)hipFree(p);
  for (uint16_t* p : h->Pbf)
    if (p) (void)hipFree(p);
  for (uint16_t* p : h->Pbf_lo)
    if (p) (void)hipFree(p);
  for (float* p : h->Pt32)
    if (p) (void)hipFree(p);
  if (h->R) (void)hipFree(h->R);
  if (h->Z) (void)hipFree(h->Z);
  DevBuf* bufs[] = {&h->cnt, &h->ptr, &h->nbr, &h->rev, &h->rowflag, &h->T5, &h->Z2, &h->T5b, &h->M5,
                    &h->XA, &h->XB, &h->XC, &h->O, &h->ENV, &h->ENVC, &h->A64, &h->Ainv, &h->AinvT, &h->ldet, &h->dsgn,
                    &h->tr, &h->trip, &h->dgrad, &h->det_lp, &h->det_sg, &h->det_gr, &h->det_lap, &h->JR,
                    &h->mlp_ns, &h->mlp_dg, &h->QV, &h->AV, &h->att_ns, &h->att_dg, &h->RIN, &h->RY1,
                    &h->RY2, &h->att_out, &h->cval, &h->clap, &h->cgrad, &h->epot};
  for (DevBuf* b : bufs) b->release();
  delete h;
  return 0;
}

int fire_upload_params(fire_handle* h, const float* packed, const int64_t* sizes, int32_t n_tensors) {
  try {
    REQUIRE(h && packed && sizes, "null argument");
    HIPCHK(hipSetDevice(h->device));
    REQUIRE(n_tensors == (int32_t)h->expect.size(),
            "parameter tensor count " + std::to_string(n_tensors) + " != expected " + std::to_string(h->expect.size()));
    for (int k = 0; k < n_tensors; ++k)
      REQUIRE(sizes[k] == h->expect[k], "parameter tensor " + std::to_string(k) + " has " + std::to_string(sizes[k]) +
                                            " elements, expected " + std::to_string(h->expect[k]));
    for (float* p : h->P)
      if (p) HIPCHK(hipFree(p));
    for (uint16_t* p : h->Pbf)
      if (p) HIPCHK(hipFree(p));
    for (uint16_t* p : h->Pbf_lo)
      if (p) HIPCHK(hipFree(p));
    for (float* p : h->Pt32)
      if (p) HIPCHK(hipFree(p));
    h->P.assign(n_tensors, nullptr);
    h->Pbf.assign(n_tensors, nullptr);
    h->Pbf_lo.assign(n_tensors, nullptr);
    h->Pt32.assign(n_tensors, nullptr);
    size_t off = 0;
    for (int k = 0; k < n_tensors; ++k) {
      HIPCHK(hipMalloc(&h->P[k], sizeof(float) * std::max<int64_t>(sizes[k], 1)));
      HIPCHK(hipMemcpy(h->P[k], packed + off, sizeof(float) * sizes[k], hipMemcpyHostToDevice));
      off += sizes[k];
    }
    std::vector<std::pair<int, std::pair<int, int>>> gw;
    gemm_weights(h, gw);
    size_t base = 0;
    std::vector<size_t> offs(n_tensors);
    for (int k = 0; k < n_tensors; ++k) {
      offs[k] = base;
      base += sizes[k];
    }
    for (auto& t : gw) {
      const int k = t.first, K = t.second.first, Nn = t.second.second;
      std::vector<uint16_t> tb((size_t)K * Nn), tl((size_t)K * Nn);
      const float* src = packed + offs[k];
      for (int kk = 0; kk < K; ++kk)
        for (int n = 0; n < Nn; ++n) {
          const float x = src[(size_t)kk * Nn + n];
          const uint16_t hi = host_f2bf(x);
          uint32_t u = (uint32_t)hi << 16;
          float xh;
          std::memcpy(&xh, &u, 4);
          tb[(size_t)n * K + kk] = hi;
          tl[(size_t)n * K + kk] = host_f2bf(x - xh);
        }
      HIPCHK(hipMalloc(&h->Pbf[k], sizeof(uint16_t) * tb.size()));
      HIPCHK(hipMemcpy(h->Pbf[k], tb.data(), sizeof(uint16_t) * tb.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&h->Pbf_lo[k], sizeof(uint16_t) * tl.size()));
      HIPCHK(hipMemcpy(h->Pbf_lo[k], tl.data(), sizeof(uint16_t) * tl.size(), hipMemcpyHostToDevice));
    }
    // FP32 transposed copies [N,K] for the FP32 MFMA kernel: every GEMM weight, and the
    // envelope weights pi per determinant ([det][ME][N] -> [det][N][ME])
    auto upload_t = [&](int k, int nb, int K, int Nn) {
      std::vector<float> tt((size_t)nb * K * Nn);
      const float* src = packed + offs[k];
      for (int b = 0; b < nb; ++b)
        for (int kk = 0; kk < K; ++kk)
          for (int n = 0; n < Nn; ++n)
            tt[(size_t)b * K * Nn + (size_t)n * K + kk] = src[(size_t)b * K * Nn + (size_t)kk * Nn + n];
      HIPCHK(hipMalloc(&h->Pt32[k], sizeof(float) * tt.size()));
      HIPCHK(hipMemcpy(h->Pt32[k], tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice));
    };
    for (auto& t : gw) upload_t(t.first, 1, t.second.first, t.second.second);
    for (int k : {h->iORB + 5, h->iORB + 7}) upload_t(k, h->D, h->M * h->Eiso, h->N);
    h->cut_learned = packed[offs[h->iSCAL]];
    h->s1 = packed[offs[h->iSCAL] + 1];
    h->params_ok = true;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_profile_gemm(fire_handle* h, int32_t enable, double* total_ms, double* total_flops, int64_t* n_launches) {
  try {
    REQUIRE(h, "null handle");
    HIPCHK(hipSetDevice(h->device));
    double ms = 0.0;
    for (auto& e : h->gemm_events) {
      HIPCHK(hipEventSynchronize(e.second));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, e.first, e.second));
      ms += t;
      HIPCHK(hipEventDestroy(e.first));
      HIPCHK(hipEventDestroy(e.second));
    }
    if (total_ms) *total_ms = ms;
    if (total_flops) *total_flops = h->gemm_flops;
    if (n_launches) *n_launches = (int64_t)h->gemm_events.size();
    h->gemm_events.clear();
    h->gemm_flops = 0.0;
    h->gemm_timing = enable != 0;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_neighbours(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t* counts_dev, int32_t* nbr_dev,
                    int64_t cap, int64_t* n_pairs) {
  try {
    REQUIRE(h && r_dev && counts_dev && n_pairs, "null argument");
    REQUIRE(B >= 0, "negative batch");
    HIPCHK(hipSetDevice(h->device));
    if (B == 0) {
      *n_pairs = 0;
      return 0;
    }
    Ctx c{h, (hipStream_t)stream, FIRE_FP32};
    Graph g = build_graph(c, r_dev, B);
    *n_pairs = g.n_pairs;
    // counts in walker-major order (the group order is up-spin first)
    std::vector<int32_t> cnt_g((size_t)B * h->N);
    HIPCHK(hipMemcpyAsync(cnt_g.data(), h->cnt.p, sizeof(int32_t) * cnt_g.size(), hipMemcpyDeviceToHost, c.s));
    std::vector<int64_t> ptr_g((size_t)B * h->N + 1);
    HIPCHK(hipMemcpyAsync(ptr_g.data(), h->ptr.p, sizeof(int64_t) * ptr_g.size(), hipMemcpyDeviceToHost, c.s));
    std::vector<int32_t> nb_g((size_t)std::max<int64_t>(g.n_pairs, 1));
    if (g.n_pairs) HIPCHK(hipMemcpyAsync(nb_g.data(), h->nbr.p, sizeof(int32_t) * g.n_pairs, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    std::vector<int32_t> cnt_w(cnt_g.size());
    std::vector<int32_t> nb_w;
    nb_w.reserve(g.n_pairs);
    const int N = h->N, nu = h->n_up, nd = N - nu;
    for (int w = 0; w < B; ++w)
      for (int i = 0; i < N; ++i) {
        const int gi = i < nu ? w * nu + i : B * nu + w * nd + (i - nu);
        cnt_w[(size_t)w * N + i] = cnt_g[gi];
        for (int64_t p = ptr_g[gi]; p < ptr_g[gi + 1]; ++p) nb_w.push_back(nb_g[p]);
      }
    HIPCHK(hipMemcpyAsync(counts_dev, cnt_w.data(), sizeof(int32_t) * cnt_w.size(), hipMemcpyHostToDevice, c.s));
    if (nbr_dev) {
      REQUIRE(cap >= g.n_pairs, "neighbour capacity too small: need " + std::to_string(g.n_pairs));
      if (g.n_pairs)
        HIPCHK(hipMemcpyAsync(nbr_dev, nb_w.data(), sizeof(int32_t) * g.n_pairs, hipMemcpyHostToDevice, c.s));
    }
    HIPCHK(hipStreamSynchronize(c.s));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_local_energy(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision, double* e_loc_dev,
                      double* e_kin_dev, double* e_pot_dev, double* logpsi_dev, double* sign_dev, double* grad_dev,
                      double* lap_dev) {
  try {
    check_ready(h);
    REQUIRE(r_dev && e_loc_dev, "null argument");
    REQUIRE(precision == FIRE_FP32 || precision == FIRE_MIXED || precision == FIRE_BF16, "unknown precision mode");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, precision};
    const int N = h->N;
    for (int w0 = 0; w0 < B; w0 += h->chunk) {
      const int Bc = std::min(h->chunk, B - w0);
      const float* r = r_dev + (size_t)w0 * N * 3;
      Graph g = build_graph(c, r, Bc);
      GroupDesc gd;
      int64_t R;
      run_embedding(c, r, Bc, g, true, gd, R);
      DetOut det;
      run_orbitals_det(c, r, Bc, g, true, gd, R, det);
      auto off = [&](double* p, int per) { return p ? p + (size_t)w0 * per : nullptr; };
      run_jastrow_and_assemble(c, r, Bc, true, gd, R, det, off(e_loc_dev, 1), off(e_kin_dev, 1), off(e_pot_dev, 1),
                               off(logpsi_dev, 1), off(sign_dev, 1), off(grad_dev, 3 * N), off(lap_dev, 1));
    }
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_signed(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision, double* logpsi_dev,
                double* sign_dev) {
  try {
    check_ready(h);
    REQUIRE(r_dev && logpsi_dev, "null argument");
    REQUIRE(precision == FIRE_FP32 || precision == FIRE_MIXED || precision == FIRE_BF16, "unknown precision mode");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, precision};
    const int N = h->N;
    for (int w0 = 0; w0 < B; w0 += h->chunk) {
      const int Bc = std::min(h->chunk, B - w0);
      const float* r = r_dev + (size_t)w0 * N * 3;
      Graph g = build_graph(c, r, Bc);
      GroupDesc gd;
      int64_t R;
      run_embedding(c, r, Bc, g, false, gd, R);
      DetOut det;
      run_orbitals_det(c, r, Bc, g, false, gd, R, det);
      run_jastrow_and_assemble(c, r, Bc, false, gd, R, det, nullptr, nullptr, nullptr, logpsi_dev + w0,
                               sign_dev ? sign_dev + w0 : nullptr, nullptr, nullptr);
    }
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_embedding(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision, float* h_contr_dev,
                   float* h_out_dev) {
  try {
    check_ready(h);
    REQUIRE(r_dev, "null argument");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, precision};
    const int N = h->N, F = h->F;
    for (int w0 = 0; w0 < B; w0 += h->chunk) {
      const int Bc = std::min(h->chunk, B - w0);
      const float* r = r_dev + (size_t)w0 * N * 3;
      Graph g = build_graph(c, r, Bc);
      GroupDesc gd;
      int64_t R;
      run_embedding(c, r, Bc, g, false, gd, R);
      // value rows are in group order (up-spin first); scatter back to walker-major order
      auto to_walker_major = [&](const float* src, float* dst) {
        const int nu = h->n_up, nd = N - nu;
        for (int w = 0; w < Bc; ++w) {
          HIPCHK(hipMemcpyAsync(dst + ((size_t)(w0 + w) * N) * F, src + (size_t)w * nu * F, sizeof(float) * nu * F,
                                hipMemcpyDeviceToDevice, c.s));
          if (nd)
            HIPCHK(hipMemcpyAsync(dst + ((size_t)(w0 + w) * N + nu) * F, src + ((size_t)Bc * nu + (size_t)w * nd) * F,
                                  sizeof(float) * nd * F, hipMemcpyDeviceToDevice, c.s));
        }
      };
      if (h_out_dev) {
        // the consumers of h hold the output scale s1 in their weights; apply it here
        launch_scale(c.s, h->XA.as<float>(), (int64_t)Bc * N * F, h->s1);
        to_walker_major(h->XA.as<float>(), h_out_dev);
      }
      if (h_contr_dev) {
        // message-passing output before the update MLP: recompute it into XB
        launch_messages(c.s, r, Bc, N, h->n_up, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), h->M5.as<float>(),
                        edge_params(h, h->iEEs), edge_params(h, h->iEEd), h->cut_learned, F, false, h->XB.as<float>());
        to_walker_major(h->XB.as<float>(), h_contr_dev);
      }
    }
    HIPCHK(hipStreamSynchronize(c.s));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

}  // extern "C"
