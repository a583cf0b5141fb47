// Host side of the C ABI (include/fire.h): handle, parameters, workspace and the
// stream-ordered launch sequence of the FiRE hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/fire.h"
#include "fire_kernels.hpp"
#include "fire_kernels2.hpp"
#include "fire_mcmc.hpp"
#include "fire_grad.hpp"

static constexpr int64_t kScratchInt64 = (int64_t)1 << 21;  // 16 MiB
static __device__ __align__(256) int64_t scratch[kScratchInt64];

namespace {

thread_local std::string g_last_error;

// Tile shapes chosen by the autotuner, kept for the process per device: every handle (one per
// geometry in ragged batches and in the serving loop) reuses a shape another one measured.
struct TileCache {
  std::mutex mu;
  std::map<std::tuple<int, int, int, int, int>, int> best;  // (device, split, N, K, log2 M) -> variant
};
TileCache& tile_cache() {
  static TileCache c;
  return c;
}

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

// Row-pointer scratch of fire_neighbours_noalloc: a fixed device array (one copy per GPU,
// held until the process ends), so that entry point never allocates.  Host code must use the
// address hipGetSymbolAddress returns for the current device, never the array's name.
struct ScratchAddrs {
  std::mutex mu;
  std::map<int, int64_t*> ptr;  // device -> address of `scratch` on that device
};
ScratchAddrs& scratch_addrs() {
  static ScratchAddrs a;
  return a;
}

// Address of `scratch` on `device` (which must be current), looked up once per device.
int64_t* scratch_address(int device) {
  auto& a = scratch_addrs();
  std::lock_guard<std::mutex> lk(a.mu);
  auto it = a.ptr.find(device);
  if (it != a.ptr.end()) return it->second;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  a.ptr[device] = static_cast<int64_t*>(p);
  return static_cast<int64_t*>(p);
}

uint16_t host_f2bf(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// IEEE binary16 of x, rounded to nearest even (normal and subnormal results; |x| below the
// overflow threshold, which the power-of-two column scales guarantee)
uint16_t host_f2h(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
  u &= 0x7fffffffu;
  if (u >= 0x47800000u) return sign | 0x7c00u;  // |x| >= 2^16: infinity
  if (u < 0x38800000u) {                         // |x| < 2^-14: a subnormal half, units of 2^-24
    float ax;
    std::memcpy(&ax, &u, 4);
    const float q = ax * 16777216.0f;  // exact: scaling by a power of two
    uint32_t m = (uint32_t)q;
    const float rem = q - (float)m;
    if (rem > 0.5f || (rem == 0.5f && (m & 1u))) ++m;
    return sign | (uint16_t)m;  // m = 0x400 is the smallest normal, correctly encoded
  }
  uint32_t hb = (u >> 13) - ((127u - 15u) << 10);  // rebias the exponent, keep 10 mantissa bits
  const uint32_t rest = u & 0x1fffu;
  if (rest > 0x1000u || (rest == 0x1000u && (hb & 1u))) ++hb;  // a carry bumps the exponent
  return sign | (uint16_t)hb;
}

float host_h2f(uint16_t hv) {
  const uint32_t sign = (uint32_t)(hv & 0x8000u) << 16;
  const uint32_t e = (hv >> 10) & 0x1fu, m = hv & 0x3ffu;
  float f;
  if (e == 0) {
    f = (float)m * (1.0f / 16777216.0f);  // subnormal: m 2^-24
  } else {
    const uint32_t u = ((e + 112u) << 23) | (m << 13);
    std::memcpy(&f, &u, 4);
  }
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u |= sign;
  std::memcpy(&f, &u, 4);
  return f;
}

// The power-of-two scale 2^(141 - e) of a column whose largest |entry| has biased exponent e
// (clamped to [15, 254]): the maximum lands in [2^14, 2^15).  Same rule as the device's rows.
float host_f16_scale(float colmax) {
  uint32_t u;
  std::memcpy(&u, &colmax, 4);
  int e = (int)((u >> 23) & 0xffu);
  e = e < 15 ? 15 : (e > 254 ? 254 : e);
  const uint32_t sb = (uint32_t)(268 - e) << 23;
  float sc;
  std::memcpy(&sc, &sb, 4);
  return sc;
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
  // 16-bit planes of the GEMM weights (null if none) in the K-blocked layout the split
  // contractions read: element (k, n) at ((k / 16) N + n) 16 + k % 16, so the 16-wide K chunk
  // of every column is one contiguous run
  std::vector<uint16_t*> Pbf;   // BF16 planes: hi part
  std::vector<uint16_t*> Pbf_lo;  // ... the BF16 remainder W - hi (mid plane)
  std::vector<uint16_t*> Pbf_lo2;  // ... and the BF16 remainder W - hi - mid (lo plane)
  std::vector<float*> Pt32;       // FP32 transposed copies [N,K] of GEMM weights
  // FP16 planes (same layout) of every GEMM weight with each column scaled by a power of two
  // (its maximum into [2^14, 2^15)), and the inverse scales [N]
  std::vector<uint16_t*> Pf16, Pf16_lo;
  std::vector<float*> Pwinv;
  std::vector<int> Pwn;  // columns N of each GEMM weight with planes (the planes' row length)
  std::vector<std::pair<int, int>> shape2;  // (K, N) of GEMM weights
  float cut_learned = 0.f;
  float s1 = 1.f;
  bool params_ok = false;
  // tensor indices
  int iEN, iEEs, iEEd, iMLP, iORB, iJM, iATT, iCUSP, iSCAL;
  // tensors derived at upload (after the uploaded ones): the attention projection as
  // W_va [F, HDP] = [value half W_v | key-weighted head sums of the query half W_a | 0] with
  // W_a[:,h] = sum_k W_q[:,h Dk + k] key[h Dk + k], and its bias b_va [HDP];
  // HDP = HD + H rounded up to 16
  int iAV = -1;
  int HDP = 0;
  // with attention: the orbital projections with the H head sums appended as columns
  // ND .. ND + H (up: iORBA, iORBA + 1 bias; down: iORBA + 2, + 3), and with an MLP Jastrow the
  // first Jastrow layer with the value half appended, [W_j1 | W_v] [F, W + HD] (iJV, iJV + 1)
  int iORBA = -1, iJV = -1;
  int ldo_orb = 0;  // row stride of the orbital rows of the last run_orbitals_det
  // workspace
  DevBuf cnt, ptr, nbr, rev, rowflag, rowgrp, AMX_E;
  DevBuf AV0, Z0, F1, PART;  // fused silu epilogue: packed value rows, their pre-activations, silu', slots
  DevBuf AMX_A, AMX_B, AMX_C;  // row maxima (float bits) of XA, XB, XC for the FP16 planes
  const uint32_t* amx_xa = nullptr;  // row maxima of the embedding rows in XA (null: none)
  DevBuf T5, Z2, T5b, M5, XA, XB, XC, O, ENV, ENVC, PF, EF, EOK;
  DevBuf A64, Ainv, AinvT, ldet, dsgn, tr, trip, dgrad;
  DevBuf det_lp, det_sg, det_gr, det_lap;
  DevBuf JR, mlp_ns, mlp_dg, QV, AV, att_ns, att_dg, RIN, RY1, RY2, att_out;
  DevBuf cval, clap, cgrad, epot;
  int chunk = 256;
  bool lu_smem_set = false;
  int64_t* scratch = nullptr;  // this device's static scratch (fire_neighbours_noalloc)
  // pseudopotentials (fire_ecp_upload): radial tables and quadrature directions on the device,
  // and the integration points of the last fire_ecp_energy
  // parameter gradient (fire_grad_logpsi): per-chunk forward intermediates and cotangents
  static constexpr int kGradSlots = 128;
  DevBuf grw[kGradSlots];
  int ecp_n = 0, ecp_nl = 0, ecp_ng = 0, ecp_qmax = 0;
  DevBuf ecp_atom, ecp_grid, ecp_vloc, ecp_vnl, ecp_cut, ecp_nq, ecp_qpts, ecp_pts, ecp_coef, ecp_kq, ecp_npts;
  // optional HIP-event timing of the dense contractions (fire_profile_gemm)
  bool gemm_timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> gemm_events;
  double gemm_flops = 0.0;
  double gemm_bytes = 0.0;  // algorithmic: A read once, C written once, the weight planes once
  // Autotuned tile shape of the BF16 contraction per (split, N, K, floor(log2 M)): the first
  // call of a shape times every candidate on the real operands and keeps the fastest.
  bool autotune = true;
  bool autotune_near = false;
  int force_tile = -1;  // >= 0: every contraction uses this tile form where it applies (fire_autotune)
  int n_cu = 0;          // compute units of the device (read once)  // reuse tuned tiles across neighbouring row buckets
  // silu of the tangent-row MLP layers fused into the contraction epilogue (FIRE_FUSED_ACT=0:
  // the separate element-wise pass instead)
  bool fuse_act = true;
  bool use_graphs = true;  // whole Metropolis sweeps replayed from a captured hipGraph
  // determinant Laplacian's products J Ainv[:, S] in FP64 (FP64 transposed inverse, FP64 matrix
  // cores) instead of FP32 on an FP32 copy of the inverse (fire_det_options)
  bool det_lap64 = false;
  // BF16 planes per operand used by FIRE_MIXED for each contraction class (ElecInit MLP,
  // update MLP, orbitals, MLP Jastrow, attention Q/V, attention readout): 3 = six products
  // (FP32-level), 2 = three products, 1 = plain.  Default: 3 for the embedding and orbital
  // classes, 2 for the Jastrow classes (on equilibrated C16H34 walkers the Jastrow classes at
  // 2 planes leave the FP64 energy error unchanged; any embedding/orbital class at 2 planes
  // raises its rms from 0.4 to 3-4 mHa)
  // 4 = two FP16 planes of operands scaled by powers of two per row of A and column of W
  // (three products, about 2^-22 relative per product); used where the producer of A left
  // its row maxima (tangent rows of the local energy), three BF16 planes elsewhere.  The
  // Jastrow classes stay on a BF16 pair: on C16H34 the FP16 planes there leave the error
  // against FP64 unchanged and are slower (N = 512 joint contraction).
  int planes[6] = {4, 4, 4, 2, 2, 2};
  std::map<std::tuple<int, int, int, int>, int> gemm_tile;
  std::map<std::tuple<int, int, int, int>, std::vector<float>> gemm_tile_ms;
  // Resident walker state for single-electron low-rank updates (fire_state_*, fire_mcmc_steps).
  // Per-electron rows are in group order over all B walkers (up-spin electrons first).
  struct State {
    int B = 0, prec = 0;
    bool ready = false, pending = false;
    int k = -1;
    // static capacities of a proposal: cap changed electrons per walker, tcap = B * cap rows
    int cap = 0;
    int64_t tcap = 0;
    int64_t n_proposals = 0;
    // device: tail = {T_up, T_dn, max changed, T} of the last proposal, stats = {sum of T, max
    // changed} since init, overflow flag, step counter of the Gaussian proposals, accepted
    // moves per walker of the current fire_mcmc_steps call
    DevBuf dstats, doverflow, dstep, nacc;
    // hipGraph of one sweep (n_el steps, electron k = step within the sweep) on `gstream`
    hipGraphExec_t graph = nullptr;
    hipStream_t gstream = nullptr;
    float g_width = 0.f;
    uint64_t g_seed = 0;
    uint64_t g_sig = 0;  // state signature at capture
    bool g_count = false;
    // state: positions [B,N,3]; h0/message rows [G,5F]; embedding [G,F]; raw orbitals and
    // envelopes [G,N*D]; Slater matrices and inverses [B,D,N,N] FP64; log|det|, sign [B,D];
    // MLP-Jastrow rows [G,2]; attention rows [G,H+H*Dk]; log psi, sign [B]
    DevBuf r, M5, hh, orb, env, A, Ainv, ld, sg, jr, av, logpsi, sign;
    // proposal: positions, moved electron, uniform variate, affected sets and row lists
    DevBuf p_r, p_rk, p_u, aff, cu, cd, offu, offd, tail, list, rowof, klist;
    // moved-electron rows (ElecInit, messages, envelope)
    DevBuf kT5, kZ2, kT5b, p_M5k, kcoef, p_envk, k_amx;
    // compact rows of the affected electrons
    DevBuf pcnt, pptr, pnbr, p_hc, p_x, p_h, p_y, p_orb, p_jr, p_qv, p_av, p_amx;
    // determinant and Jastrow of the proposal
    DevBuf phin, sbuf, ldS, sgS, ccusp, cdelta, jsum, rin, ry1, ry2, att_out, cval, clap, cgrad, epot;
    DevBuf p_ld, p_sg, p_logpsi, p_sign, accept, wbscr;
    // state buffers, walker count and precision: a captured sweep stays valid while these hold
    uint64_t signature() {
      uint64_t hsh = 1469598103934665603ull;
      auto mix = [&](uint64_t v) { hsh = (hsh ^ v) * 1099511628211ull; };
      for (DevBuf* x : bufs()) mix(reinterpret_cast<uint64_t>(x->p));
      mix((uint64_t)B);
      mix((uint64_t)prec);
      return hsh;
    }
    std::vector<DevBuf*> bufs() {
      return {&r,     &M5,    &hh,    &orb,     &env,   &A,      &Ainv,  &ld,    &sg,     &jr,    &av,
              &logpsi, &sign, &p_r,   &p_rk,    &p_u,   &aff,    &cu,    &cd,    &offu,   &offd,  &tail,
              &list,  &rowof, &klist, &kT5,     &kZ2,   &kT5b,   &p_M5k, &kcoef, &p_envk, &pcnt,  &pptr,
              &pnbr,  &p_hc,  &p_x,   &p_h,     &p_y,   &p_orb,  &p_jr,  &p_qv,  &p_av,   &phin,  &sbuf, &ccusp, &cdelta,
              &ldS,   &sgS,   &jsum,  &rin,     &ry1,   &ry2,    &att_out, &cval, &clap,  &cgrad, &epot,
              &p_ld,  &p_sg,  &p_logpsi, &p_sign, &accept, &wbscr, &dstats, &doverflow, &dstep, &nacc,
              &k_amx, &p_amx};
    }
    void release() {
      DevBuf* b[] = {&r,     &M5,    &hh,    &orb,     &env,   &A,      &Ainv,  &ld,    &sg,     &jr,    &av,
                     &logpsi, &sign, &p_r,   &p_rk,    &p_u,   &aff,    &cu,    &cd,    &offu,   &offd,  &tail,
                     &list,  &rowof, &klist, &kT5,     &kZ2,   &kT5b,   &p_M5k, &kcoef, &p_envk, &pcnt,  &pptr,
                     &pnbr,  &p_hc,  &p_x,   &p_h,     &p_y,   &p_orb,  &p_jr,  &p_qv,  &p_av,   &phin,  &sbuf, &ccusp, &cdelta,
                     &ldS,   &sgS,   &jsum,  &rin,     &ry1,   &ry2,    &att_out, &cval, &clap,  &cgrad, &epot,
                     &p_ld,  &p_sg,  &p_logpsi, &p_sign, &accept, &wbscr, &dstats, &doverflow, &dstep, &nacc,
              &k_amx, &p_amx};
      if (graph) (void)hipGraphExecDestroy(graph);
      graph = nullptr;
      if (gstream) (void)hipStreamDestroy(gstream);
      gstream = nullptr;
      for (DevBuf* x : b) x->release();
      ready = pending = false;
    }
  } st;
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
    out.push_back({h->iAV, {F, h->HDP}});
    out.push_back({h->iORBA, {F, ND + h->H}});
    out.push_back({h->iORBA + 2, {F, ND + h->H}});
    if (h->has_mlp) out.push_back({h->iJV, {F, W + HD}});
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
  bool fixed = false;  // no autotuning or event timing (stream capture, device-sized launches)
};

// Contraction class of weight tensor wi (index into fire_handle::planes).
int gemm_class(const fire_handle* h, int wi) {
  if (wi < h->iEEs) return 0;
  if (wi < h->iORB) return 1;
  if (wi < h->iJM) return 2;
  if (wi < h->iATT) return 3;
  if (wi == h->iORBA || wi == h->iORBA + 2) return 2;
  if (wi == h->iJV) return 3;  // shared with the attention values: planes of both (gemm())
  return wi == h->iATT || wi == h->iAV ? 4 : 5;
}

// C = A @ W(tensor wi) + bias(tensor bi, optional) on value rows
void gemm(Ctx& c, const float* A, int64_t lda, int wi, int bi, float* C, int64_t ldc, int64_t M, int N, int K,
          const uint8_t* rowflag, int period, DevRows dr = DevRows(), const ActEpi* epp = nullptr) {
  ActEpi ep = epp ? *epp : ActEpi();
  fire_handle* h = c.h;
  const float* bias = bi >= 0 ? h->P[bi] : nullptr;
  const bool bf_ok = h->Pbf[wi] != nullptr && (K % 32) == 0 && (lda % 4) == 0 && N > 16;
  const bool use_bf = (c.prec == FIRE_MIXED || c.prec == FIRE_MIXED_FAST || c.prec == FIRE_BF16) && bf_ok;
  REQUIRE(!use_bf || N == h->Pwn[wi], "contraction of weight " + std::to_string(wi) + " over " + std::to_string(N) +
                                          " of its " + std::to_string(h->Pwn[wi]) + " columns");
  int ns = c.prec == FIRE_MIXED_FAST ? 3 : 1;
  if (c.prec == FIRE_MIXED) {
    int np = h->planes[gemm_class(h, wi)];
    if (wi == h->iJV) np = (np == 4 || h->planes[4] == 4) ? 4 : std::max(np, h->planes[4]);
    if (np == 4 && !(ep.amax && h->Pf16[wi])) np = 3;  // a caller without row maxima of A: three BF16 planes
    ns = np == 4 ? 4 : (np == 3 ? 6 : (np == 2 ? 3 : 1));
  }
  const uint16_t* W0 = h->Pbf[wi];
  const uint16_t* W1 = h->Pbf_lo[wi];
  const uint16_t* W2 = h->Pbf_lo2[wi];
  if (ns == 4) {
    W0 = h->Pf16[wi];
    W1 = h->Pf16_lo[wi];
    W2 = nullptr;
    ep.winv = h->Pwinv[wi];
  } else {
    ep.amax = nullptr;
  }
  int variant = 0;
  int lg = 0;
  while ((int64_t(1) << (lg + 1)) <= M) ++lg;
  if (dr.count) lg += 32;  // rows counted on the device (M is the capacity): tuned apart
  if (ep.rowgrp) lg += 64;  // fused silu epilogue: tuned apart
  if (use_bf && h->force_tile >= 0) {
    // one tile form for every contraction; where that form does not apply, the base tile
    variant = h->force_tile;
    if (variant >= gemm_bf16_variants() || (ep.rowgrp && !gemm_bf16_fusable(variant)) ||
        (ns == 4 && !gemm_bf16_f16_ok(variant)) || gemm_bf16_spills(variant, ns))
      variant = 0;
    h->gemm_tile[std::make_tuple(ns, N, K, lg)] = variant;  // reported by fire_autotune's table
  } else if (use_bf && h->autotune) {
    const auto key = std::make_tuple(ns, N, K, lg);
    const auto gkey = std::make_tuple(h->device, ns, N, K, lg);
    auto it = h->gemm_tile.find(key);
    bool known = it != h->gemm_tile.end();
    if (known) {
      variant = it->second;
    } else {
      std::lock_guard<std::mutex> lk(tile_cache().mu);
      auto g = tile_cache().best.find(gkey);
      if (g != tile_cache().best.end()) {
        variant = g->second;
        h->gemm_tile[key] = variant;
        known = true;
      } else {
        // nearest tuned shape of the same class (device, split, K, epilogue / row-count flags,
        // number of 128-column tiles, and row bucket: the same one, or with near_rows within
        // two): ragged and serving batches bring new N and row counts with every geometry
        int64_t dbest = -1;
        for (const auto& kv : tile_cache().best) {
          const auto& k = kv.first;
          const int lg2 = std::get<4>(k);
          if (std::get<0>(k) != h->device || std::get<1>(k) != ns || std::get<3>(k) != K || (lg2 & ~31) != (lg & ~31))
            continue;
          const int dl = std::abs((lg2 & 31) - (lg & 31));
          if (dl > (h->autotune_near ? 2 : 0)) continue;
          const int n2 = std::get<2>(k);
          if ((n2 + 127) / 128 != (N + 127) / 128) continue;
          const int64_t dist = (int64_t)dl * 100000 + std::abs(n2 - N);
          if (dbest < 0 || dist < dbest) {
            dbest = dist;
            variant = kv.second;
          }
        }
        if (dbest >= 0) {
          h->gemm_tile[key] = variant;
          known = true;
        }
      }
    }
    if (!known && !c.fixed) {
      // time every tile shape on these operands: one untimed launch, then the faster of two
      const int nv = gemm_bf16_variants();
      std::vector<float> ms(nv, 0.f);
      hipEvent_t t0, t1;
      HIPCHK(hipEventCreate(&t0));
      HIPCHK(hipEventCreate(&t1));
      if (h->n_cu == 0) {
        int n = 0;
        h->n_cu = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && n > 0
                      ? n
                      : 256;
      }
      for (int v = 0; v < nv; ++v) {
        int bm = 128, bn = 128;
        gemm_bf16_tile(v, &bm, &bn);
        // a persistent form pays off only with several row tiles per CU
        const bool few_tiles = gemm_bf16_persistent(v) && M < (int64_t)2 * bm * h->n_cu;
        if ((ep.rowgrp && !gemm_bf16_fusable(v)) || (ns == 4 && !gemm_bf16_f16_ok(v)) || gemm_bf16_spills(v, ns) ||
            few_tiles) {
          ms[v] = 1e30f;
          continue;
        }
        launch_gemm_bf16(c.s, A, lda, W0, W1, W2, bias, C, ldc, M, N, K, rowflag, period, ns, v, dr, ep);
        ms[v] = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          HIPCHK(hipEventRecord(t0, c.s));
          launch_gemm_bf16(c.s, A, lda, W0, W1, W2, bias, C, ldc, M, N, K, rowflag, period, ns, v, dr, ep);
          HIPCHK(hipEventRecord(t1, c.s));
          HIPCHK(hipEventSynchronize(t1));
          float t = 0.f;
          HIPCHK(hipEventElapsedTime(&t, t0, t1));
          ms[v] = std::min(ms[v], t);
        }
        if (ms[v] < ms[variant]) variant = v;
      }
      HIPCHK(hipEventDestroy(t0));
      HIPCHK(hipEventDestroy(t1));
      h->gemm_tile[key] = variant;
      h->gemm_tile_ms[key] = ms;
      std::lock_guard<std::mutex> lk(tile_cache().mu);
      tile_cache().best[gkey] = variant;
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = h->gemm_timing && !c.fixed;
  if (timed) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c.s));
  }
  if (use_bf) {
    launch_gemm_bf16(c.s, A, lda, W0, W1, W2, bias, C, ldc, M, N, K, rowflag, period, ns, variant, dr, ep);
  } else
    launch_gemm_f32(c.s, A, lda, h->P[wi], h->Pt32[wi], bias, C, ldc, M, N, K, rowflag, period, dr, ep);
  if (timed) {
    HIPCHK(hipEventRecord(e1, c.s));
    h->gemm_events.push_back({e0, e1});
    h->gemm_flops += 2.0 * (double)M * N * K;
    const double wbytes = use_bf ? 2.0 * (ns == 6 ? 3 : (ns == 1 ? 1 : 2)) : 4.0;
    h->gemm_bytes += 4.0 * (double)M * (K + N) + wbytes * N * K;
  }
}

// C = silu(A @ W + b) over the tangent rows of G groups (forward-Laplacian rule of k_act), the
// activation fused into the contraction's epilogue: the value rows are contracted first (FP32,
// packed) for silu' of every group, the epilogue applies silu / silu' and leaves per-block sums
// of squares of the jacobian rows, k_act_lapfix completes the Laplacian rows.
// With wf >= 0 only the first Nf columns are activated: they are the product with weight
// tensor wf (bias bf), the remaining N - Nf columns are stored as they are (+ bias on value
// rows); requires the fused epilogue.
bool gemm_silu_fusable(const fire_handle* h, int N, int K, int64_t lda) {
  return h->fuse_act && N % 128 == 0 && K % 32 == 0 && lda % 4 == 0;
}

// Whether the contractions of this call use FP16 planes somewhere (row maxima worth producing).
bool f16_wanted(const Ctx& c) {
  if (c.prec != FIRE_MIXED) return false;
  for (int k = 1; k < 5; ++k)
    if (c.h->planes[k] == 4) return true;
  return false;
}

// amax (optional): row maxima of A for FP16 planes; omax (optional): buffer (R entries) that
// receives the row maxima of C over its first Nf columns.  Returns omax when it was filled,
// null otherwise (the unfused fallback leaves none).
const uint32_t* gemm_silu_tan(Ctx& c, const float* A, int64_t lda, int wi, int bi, float* C, int64_t ldc, int64_t R,
                              int N, int K, const uint8_t* flag, int G, const GroupDesc& gd, int wf = -1, int bf = -1,
                              int Nf = 0, const uint32_t* amax = nullptr, DevBuf* omax = nullptr) {
  fire_handle* h = c.h;
  if (wf < 0) {
    wf = wi;
    bf = bi;
    Nf = N;
  }
  if (!gemm_silu_fusable(h, Nf, K, lda) || G == 0) {
    REQUIRE(Nf == N, "partial activation needs the fused epilogue");
    ActEpi ea;
    ea.amax = amax;
    gemm(c, A, lda, wi, bi, C, ldc, R, N, K, flag, 0, DevRows(), &ea);
    launch_act(c.s, C, ldc, N, G, gd, ACT_SILU, 0.f);
    return nullptr;
  }
  uint32_t* om = nullptr;
  if (omax && f16_wanted(c)) {
    omax->ensure(sizeof(uint32_t) * (size_t)R);
    om = omax->as<uint32_t>();
    launch_zero_f32(c.s, reinterpret_cast<float*>(om), R);
  }
  h->AV0.ensure(sizeof(float) * (size_t)G * K);
  h->Z0.ensure(sizeof(float) * (size_t)G * Nf);
  h->F1.ensure(sizeof(float) * (size_t)G * Nf);
  h->PART.ensure(sizeof(float) * (size_t)G * ACT_EPI_SLOTS * Nf);
  float* z0 = h->Z0.as<float>();
  launch_gather_rows(c.s, A, lda, K, G, gd, h->AV0.as<float>());
  launch_gemm_f32(c.s, h->AV0.as<float>(), K, h->P[wf], h->Pt32[wf], bf >= 0 ? h->P[bf] : nullptr, z0, Nf, G, Nf, K,
                  nullptr, 1);
  launch_silu_grad(c.s, z0, h->F1.as<float>(), (int64_t)G * Nf);
  ActEpi ep;
  ep.rowgrp = h->rowgrp.as<int32_t>();
  ep.f1 = h->F1.as<float>();
  ep.part = h->PART.as<float>();
  ep.gd = gd;
  ep.ldf = Nf;
  ep.nfuse = Nf;
  ep.amax = amax;
  ep.omax = om;
  ep.omax_cols = Nf;
  gemm(c, A, lda, wi, bi, C, ldc, R, N, K, flag, 0, DevRows(), &ep);
  launch_act_lapfix(c.s, C, ldc, Nf, G, gd, z0, h->PART.as<float>(), Nf, om);
  return om;
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
  float* ef = nullptr;
  uint8_t* eok = nullptr;
  if (tan) {
    const EnParams ep = en_params(h);
    h->EF.ensure(sizeof(float) * pair_filter_floats(ep.w1) * (size_t)G * h->M);
    h->EOK.ensure((size_t)G * h->M);
    ef = h->EF.as<float>();
    eok = h->EOK.as<uint8_t>();
  }
  launch_elec_init_en(c.s, r, h->R, B, N, h->n_up, h->M, h->cfg.cutoff_1el, F, en_params(h), tan, h->T5.as<float>(),
                      nullptr, 0, ef, eok);
  const int64_t rows5 = (int64_t)G * NR;
  // FP16 planes for the ElecInit MLP (class 0 set to 4): row maxima of each input, taken from
  // the rows themselves (so the sampler's rows of a moved electron get the same scales)
  ActEpi e0;
  const ActEpi* pe0 = nullptr;
  if (c.prec == FIRE_MIXED && h->planes[0] == 4) {
    h->AMX_E.ensure(sizeof(uint32_t) * (size_t)rows5);
    e0.amax = h->AMX_E.as<uint32_t>();
    pe0 = &e0;
  }
  if (pe0) launch_rowmax(c.s, h->T5.as<float>(), F, rows5, F, h->AMX_E.as<uint32_t>());
  gemm(c, h->T5.as<float>(), F, b + 11, -1, h->Z2.as<float>(), 2 * F, rows5, 2 * F, F, nullptr, 0, DevRows(), pe0);
  launch_glu(c.s, h->Z2.as<float>(), G, F, tan, h->T5b.as<float>());
  if (pe0) launch_rowmax(c.s, h->T5b.as<float>(), F, rows5, F, h->AMX_E.as<uint32_t>());
  gemm(c, h->T5b.as<float>(), F, b + 12, b + 13, h->T5.as<float>(), F, rows5, F, F, nullptr, NR, DevRows(), pe0);
  GroupDesc g5{NR, 0, nullptr};
  launch_act(c.s, h->T5.as<float>(), F, F, G, g5, ACT_SILU, 0.f);
  if (pe0) launch_rowmax(c.s, h->T5.as<float>(), F, rows5, F, h->AMX_E.as<uint32_t>());
  gemm(c, h->T5.as<float>(), F, b + 14, b + 15, h->M5.as<float>(), 5 * F, rows5, 5 * F, F, nullptr, NR, DevRows(),
       pe0);

  if (tan) {
    gd = GroupDesc{5, 3, h->ptr.as<int64_t>()};
    R = 5 * (int64_t)G + 3 * gr.n_pairs;
  } else {
    gd = GroupDesc{1, 0, nullptr};
    R = G;
  }
  h->XA.ensure(sizeof(float) * (size_t)R * F);
  h->XB.ensure(sizeof(float) * (size_t)R * F);
  float* pf = nullptr;
  if (tan) {
    h->PF.ensure(sizeof(float) * pair_filter_floats(edge_params(h, h->iEEs).filt.w1) *
                 (size_t)std::max<int64_t>(gr.n_pairs, 1));
    pf = h->PF.as<float>();
  }
  uint32_t* amx_msg = nullptr;  // row maxima (bounds) of the message rows, set per group by k_messages
  if (tan && f16_wanted(c)) {
    h->AMX_A.ensure(sizeof(uint32_t) * (size_t)R);
    amx_msg = h->AMX_A.as<uint32_t>();
  }
  launch_messages(c.s, r, B, N, h->n_up, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), h->M5.as<float>(),
                  edge_params(h, h->iEEs), edge_params(h, h->iEEd), h->cut_learned, F, tan, h->XA.as<float>(), nullptr,
                  0, -1, nullptr, nullptr, pf, amx_msg);
  h->amx_xa = nullptr;
  const uint8_t* flag = nullptr;
  int period = 1;
  if (tan) {
    h->rowflag.ensure((size_t)R);
    launch_rowflag(c.s, G, gd, R, h->rowflag.as<uint8_t>());
    flag = h->rowflag.as<uint8_t>();
    period = 0;
    h->rowgrp.ensure(sizeof(int32_t) * (size_t)R);
    launch_rowgrp(c.s, G, gd, h->rowgrp.as<int32_t>());
  }
  const int m = h->iMLP;
  if (tan) {
    const uint32_t* amx_b = gemm_silu_tan(c, h->XA.as<float>(), F, m + 0, m + 1, h->XB.as<float>(), F, R, F, F, flag,
                                          G, gd, -1, -1, 0, amx_msg, &h->AMX_B);
    h->amx_xa = gemm_silu_tan(c, h->XB.as<float>(), F, m + 2, m + 3, h->XA.as<float>(), F, R, F, F, flag, G, gd, -1,
                              -1, 0, amx_b, &h->AMX_A);
    return;
  }
  // value rows: FP16 planes (classes 1 and 2 set to 4) with the row maxima of each input
  ActEpi em;
  const bool f16m = c.prec == FIRE_MIXED && h->planes[1] == 4;
  if (f16m || (c.prec == FIRE_MIXED && h->planes[2] == 4)) {
    h->AMX_A.ensure(sizeof(uint32_t) * (size_t)R);
    h->AMX_B.ensure(sizeof(uint32_t) * (size_t)R);
  }
  if (f16m) {
    launch_rowmax(c.s, h->XA.as<float>(), F, R, F, h->AMX_A.as<uint32_t>());
    em.amax = h->AMX_A.as<uint32_t>();
  }
  gemm(c, h->XA.as<float>(), F, m + 0, m + 1, h->XB.as<float>(), F, R, F, F, flag, period, DevRows(), &em);
  launch_act(c.s, h->XB.as<float>(), F, F, G, gd, ACT_SILU, 0.f);
  if (f16m) {
    launch_rowmax(c.s, h->XB.as<float>(), F, R, F, h->AMX_B.as<uint32_t>());
    em.amax = h->AMX_B.as<uint32_t>();
  }
  gemm(c, h->XB.as<float>(), F, m + 2, m + 3, h->XA.as<float>(), F, R, F, F, flag, period, DevRows(), &em);
  launch_act(c.s, h->XA.as<float>(), F, F, G, gd, ACT_SILU, 0.f);
  if (c.prec == FIRE_MIXED && h->planes[2] == 4) {  // the orbital contraction's row maxima
    launch_rowmax(c.s, h->XA.as<float>(), F, R, F, h->AMX_A.as<uint32_t>());
    h->amx_xa = h->AMX_A.as<uint32_t>();
  }
}

struct DetOut {
  double *logpsi, *sign, *grad, *lap;
};

// Copies kept by the state initialisation (value rows only): raw orbital rows, envelope
// rows, Slater matrices, inverses, log|det| and signs.
struct Keep {
  float *orb, *env;
  double *A, *Ainv, *ld, *sg;
};

void run_orbitals_det(Ctx& c, const float* r, int B, const Graph& gr, bool tan, const GroupDesc& gd, int64_t R,
                      DetOut& out, const Keep* keep = nullptr) {
  fire_handle* h = c.h;
  const int N = h->N, F = h->F, G = B * N, ND = h->ND, D = h->D;
  const int NR = tan ? 5 : 1;
  // orbital rows at a stride of a multiple of 32 floats (128-byte lines); the padding columns
  // are read (against zero inverse columns) by the determinant Laplacian's vector loads and are
  // zeroed whenever the buffer is allocated (every later writer stores finite values)
  // tangent rows with attention: the orbital contraction also yields the H head sums (columns
  // ND .. ND + H; the value path keeps the orbital-only form the Metropolis updates share)
  const bool aug = tan && h->has_att;
  const int NDA = ND + (aug ? h->H : 0);
  const int ldo = (NDA + 31) & ~31;
  h->ldo_orb = ldo;
  const size_t ocols = std::max(ldo, h->has_att ? h->HDP : 0);
  const void* o_prev = h->O.p;
  h->O.ensure(sizeof(float) * (size_t)R * ocols);
  if (h->O.p != o_prev) launch_zero_f32(c.s, h->O.as<float>(), (int64_t)(h->O.bytes / sizeof(float)));
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
  const int wup = aug ? h->iORBA : o, wdn = aug ? h->iORBA + 2 : o + 2;
  ActEpi eu, ed;  // row maxima of the embedding rows for FP16 planes (tangent rows)
  eu.amax = h->amx_xa;  // null when no FP16 planes are wanted
  ed.amax = eu.amax ? eu.amax + R_up_rows : nullptr;
  gemm(c, h->XA.as<float>(), F, wup, wup + 1, Ob, ldo, R_up_rows, NDA, F, flag, period, DevRows(), &eu);
  if (R > R_up_rows)
    gemm(c, h->XA.as<float>() + (size_t)R_up_rows * F, F, wdn, wdn + 1, Ob + (size_t)R_up_rows * ldo, ldo,
         R - R_up_rows, NDA, F, flag ? flag + R_up_rows : nullptr, period, DevRows(), &ed);
  if (keep)
    HIPCHK(hipMemcpy2DAsync(keep->orb, sizeof(float) * ND, Ob, sizeof(float) * ldo, sizeof(float) * ND, (size_t)R,
                            hipMemcpyDeviceToDevice, c.s));
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
      const float* piT = h->Pt32[o + (sp == 0 ? 5 : 7)];
      launch_gemm_f32_batched(c.s, h->ENVC.as<float>() + (size_t)r0 * ME, ME, rows_all * ME, piT, (int64_t)ME * N,
                              h->ENV.as<float>() + (size_t)r0 * ND, ND, N, nr, N, ME, D);
    }
  }
  if (keep)
    HIPCHK(hipMemcpyAsync(keep->env, h->ENV.p, sizeof(float) * (size_t)G * ND, hipMemcpyDeviceToDevice, c.s));
  launch_orb_product(c.s, Ob, ldo, ND, G, gd, h->ENV.as<float>(), tan);
  const size_t nmat = (size_t)B * D;
  h->A64.ensure(sizeof(double) * nmat * N * N);
  h->ldet.ensure(sizeof(double) * nmat);
  h->dsgn.ensure(sizeof(double) * nmat);
  launch_gather_A(c.s, Ob, ldo, ND, gd, B, N, h->n_up, D, h->A64.as<double>());
  if (keep)
    HIPCHK(hipMemcpyAsync(keep->A, h->A64.p, sizeof(double) * nmat * N * N, hipMemcpyDeviceToDevice, c.s));
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
    const bool lap64 = h->det_lap64 && det_lap64_supported(N, ND, ldo);
    h->AinvT.ensure((lap64 ? sizeof(double) : sizeof(float)) * nmat * N * N);
    launch_lu_inv(c.s, h->A64.as<double>(), (int)nmat, N, h->Ainv.as<double>(), lap64 ? nullptr : h->AinvT.as<float>(),
                  h->ldet.as<double>(), h->dsgn.as<double>(), true, lu_lds);
    launch_det_trace(c.s, Ob, ldo, gd, B, N, h->n_up, D, h->Ainv.as<double>(), h->tr.as<double>());
    const int smax = gr.max_nb + 1;
    REQUIRE(smax <= FIRE_SMAX, "an electron has more than 63 neighbours within the cutoff (determinant Laplacian limit)");
    if (lap64) {
      launch_transpose_f64(c.s, h->Ainv.as<double>(), (int)nmat, N, h->AinvT.as<double>());
      launch_det_lap64(c.s, Ob, ldo, ND, gd, h->ENV.as<float>(), h->ptr.as<int64_t>(), h->nbr.as<int32_t>(),
                       h->rev.as<int32_t>(), B, N, h->n_up, D, smax, h->AinvT.as<double>(), h->trip.as<double>(),
                       h->dgrad.as<double>());
    } else {
      launch_det_lap(c.s, Ob, ldo, ND, gd, h->ENV.as<float>(), h->ptr.as<int64_t>(), h->nbr.as<int32_t>(),
                     h->rev.as<int32_t>(), B, N, h->n_up, D, smax, h->AinvT.as<float>(), h->trip.as<double>(),
                     h->dgrad.as<double>());
    }
    out.grad = h->det_gr.as<double>();
    out.lap = h->det_lap.as<double>();
  } else {
    launch_lu_inv(c.s, h->A64.as<double>(), (int)nmat, N, keep ? keep->Ainv : nullptr, nullptr, h->ldet.as<double>(),
                  h->dsgn.as<double>(), keep != nullptr, lu_lds);
    out.grad = nullptr;
    out.lap = nullptr;
  }
  launch_det_combine(c.s, B, N, D, h->ldet.as<double>(), h->dsgn.as<double>(), tan ? h->tr.as<double>() : nullptr,
                     tan ? h->trip.as<double>() : nullptr, tan ? h->dgrad.as<double>() : nullptr, out.logpsi, out.sign,
                     out.grad, out.lap, tan);
  if (keep) {
    HIPCHK(hipMemcpyAsync(keep->ld, h->ldet.p, sizeof(double) * nmat, hipMemcpyDeviceToDevice, c.s));
    HIPCHK(hipMemcpyAsync(keep->sg, h->dsgn.p, sizeof(double) * nmat, hipMemcpyDeviceToDevice, c.s));
  }
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
  // tangent rows, two Jastrow layers and attention: the first Jastrow layer and the attention
  // values are one contraction of the features ([W_j1 | W_v], silu on the first W columns
  // only), its rows [h1 (W) | v (HD)] in XC; the head sums come with the orbital rows
  const bool joint = tan && h->has_mlp && h->has_att && h->depth == 2 && h->iJV >= 0 &&
                     gemm_silu_fusable(h, W, F, F) && (W + h->HD) % 4 == 0;
  if (h->has_mlp) {
    h->XC.ensure(sizeof(float) * (size_t)R * (joint ? W + h->HD : W));
    h->XB.ensure(sizeof(float) * (size_t)R * std::max(W, F));
    h->JR.ensure(sizeof(float) * (size_t)R * 2);
    h->mlp_ns.ensure(sizeof(double) * (size_t)B * 8);
    const int j = h->iJM;
    float* cur = h->XB.as<float>();
    float* nxt = h->XC.as<float>();
    if (joint) {
      const int NJ = W + h->HD;
      ActEpi e2;
      // the second Jastrow layer reads FP16 planes of XC only with class 3 on them
      e2.amax = gemm_silu_tan(c, h->XA.as<float>(), F, h->iJV, h->iJV + 1, h->XC.as<float>(), NJ, R, NJ, F, flag, G,
                              gd, j + 0, j + 1, W, h->amx_xa, h->planes[3] == 4 ? &h->AMX_C : nullptr);
      gemm(c, h->XC.as<float>(), NJ, j + 2, j + 3, h->XB.as<float>(), W, R, W, W, flag, period, DevRows(), &e2);
    } else {
      // the silu of the last hidden layer is applied inside k_jmlp_last, the others in the
      // epilogue of their contraction (tangent rows) or by k_act
      for (int l = 0; l < h->depth; ++l) {
        const float* in = l == 0 ? h->XA.as<float>() : nxt;
        const int fan = l == 0 ? F : W;
        if (l + 1 < h->depth && tan) {
          gemm_silu_tan(c, in, fan, j + 2 * l, j + 2 * l + 1, cur, W, R, W, fan, flag, G, gd, -1, -1, 0,
                        l == 0 ? h->amx_xa : nullptr);
        } else {
          ActEpi e1;
          e1.amax = (l == 0 && tan) ? h->amx_xa : nullptr;
          gemm(c, in, fan, j + 2 * l, j + 2 * l + 1, cur, W, R, W, fan, flag, period, DevRows(), &e1);
          if (l + 1 < h->depth) launch_act(c.s, cur, W, W, G, gd, ACT_SILU, 0.f);
        }
        std::swap(cur, nxt);
      }
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
    h->AV.ensure(sizeof(float) * (size_t)R * nc);
    h->att_ns.ensure(sizeof(double) * (size_t)B * 2 * 2 * nc);
    if (joint) {  // values behind the first Jastrow layer, head sums behind the orbitals
      launch_att_group(c.s, h->XC.as<float>() + W, W + HD, h->O.as<float>() + h->ND, h->ldo_orb, H, h->Dk, G, gd,
                       tan, h->AV.as<float>());
    } else {
      // one contraction gives the value half and the H key-weighted head sums of the query half
      // (W_a = W_q key per head): rows [v (HD) | a (H) | 0]; the orbital rows are no longer needed
      float* QV = h->O.as<float>();
      const int HDP = h->HDP;
      ActEpi ea;
      ea.amax = tan ? h->amx_xa : nullptr;
      gemm(c, h->XA.as<float>(), F, h->iAV, h->iAV + 1, QV, HDP, R, HDP, F, flag, period, DevRows(), &ea);
      launch_att_group(c.s, QV, HDP, QV + HD, HDP, H, h->Dk, G, gd, tan, h->AV.as<float>());
    }
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

// ---------------------------------------------------------------------------------------
// Resident walker state and single-electron low-rank proposals
// (log_psi_with_state wave_function.py:226-249, log_psi_low_rank_update :251-287,
// mcmc_steps_low_rank mcmc.py:196-275).
// ---------------------------------------------------------------------------------------

// Value pass over all B walkers at once (one group order for the whole state) keeping every
// per-electron quantity a later single-electron update reuses.
void state_init(Ctx& c, const float* r_dev, int B) {
  fire_handle* h = c.h;
  auto& S = h->st;
  const int N = h->N, F = h->F, G = B * N, ND = h->ND, D = h->D, W = h->W, HD = h->HD;
  const int nc = h->H + HD;
  S.ready = S.pending = false;  // a captured sweep is kept if the buffers stay in place
  S.B = B;
  S.prec = c.prec;
  S.n_proposals = 0;
  S.r.ensure(sizeof(float) * (size_t)G * 3);
  S.p_r.ensure(sizeof(float) * (size_t)G * 3);
  HIPCHK(hipMemcpyAsync(S.r.p, r_dev, sizeof(float) * (size_t)G * 3, hipMemcpyDeviceToDevice, c.s));
  HIPCHK(hipMemcpyAsync(S.p_r.p, r_dev, sizeof(float) * (size_t)G * 3, hipMemcpyDeviceToDevice, c.s));
  Graph g = build_graph(c, S.r.as<float>(), B);
  GroupDesc gd;
  int64_t R;
  run_embedding(c, S.r.as<float>(), B, g, false, gd, R);
  S.M5.ensure(sizeof(float) * (size_t)G * 5 * F);
  S.hh.ensure(sizeof(float) * (size_t)G * F);
  HIPCHK(hipMemcpyAsync(S.M5.p, h->M5.p, sizeof(float) * (size_t)G * 5 * F, hipMemcpyDeviceToDevice, c.s));
  HIPCHK(hipMemcpyAsync(S.hh.p, h->XA.p, sizeof(float) * (size_t)G * F, hipMemcpyDeviceToDevice, c.s));
  const size_t nmat = (size_t)B * D;
  S.orb.ensure(sizeof(float) * (size_t)G * ND);
  S.env.ensure(sizeof(float) * (size_t)G * ND);
  S.A.ensure(sizeof(double) * nmat * N * N);
  S.Ainv.ensure(sizeof(double) * nmat * N * N);
  S.ld.ensure(sizeof(double) * nmat);
  S.sg.ensure(sizeof(double) * nmat);
  Keep keep{S.orb.as<float>(), S.env.as<float>(), S.A.as<double>(), S.Ainv.as<double>(), S.ld.as<double>(),
            S.sg.as<double>()};
  DetOut det;
  run_orbitals_det(c, S.r.as<float>(), B, g, false, gd, R, det, &keep);
  S.logpsi.ensure(sizeof(double) * B);
  S.sign.ensure(sizeof(double) * B);
  run_jastrow_and_assemble(c, S.r.as<float>(), B, false, gd, R, det, nullptr, nullptr, nullptr, S.logpsi.as<double>(),
                           S.sign.as<double>(), nullptr, nullptr);
  if (h->has_mlp) {
    S.jr.ensure(sizeof(float) * (size_t)G * 2);
    HIPCHK(hipMemcpyAsync(S.jr.p, h->JR.p, sizeof(float) * (size_t)G * 2, hipMemcpyDeviceToDevice, c.s));
  }
  if (h->has_att) {
    S.av.ensure(sizeof(float) * (size_t)G * nc);
    HIPCHK(hipMemcpyAsync(S.av.p, h->AV.p, sizeof(float) * (size_t)G * nc, hipMemcpyDeviceToDevice, c.s));
  }
  // proposal buffers at their static capacities (no allocation while proposing, so a sweep
  // can be captured into a graph): at most cap changed electrons per walker and move
  S.cap = std::min(N, 64);
  // a smaller capacity (FIRE_STATE_CAP) lets a test drive a move past it
  if (const char* e = getenv("FIRE_STATE_CAP")) S.cap = std::max(1, std::min(S.cap, atoi(e)));
  S.tcap = (int64_t)B * S.cap;
  const int64_t T = S.tcap;
  const int ME = h->M * h->Eiso, XW = std::max(F, std::max(W, 2 * HD)), amax = S.cap;
  S.accept.ensure(sizeof(int32_t) * B);
  S.nacc.ensure(sizeof(int32_t) * B);
  S.p_rk.ensure(sizeof(float) * B * 3);
  S.p_u.ensure(sizeof(double) * B);
  S.aff.ensure(sizeof(int32_t) * (size_t)B * N);
  for (DevBuf* b : {&S.cu, &S.cd, &S.offu, &S.offd}) b->ensure(sizeof(int32_t) * B);
  S.tail.ensure(sizeof(int32_t) * 4);
  S.dstats.ensure(sizeof(int64_t) * 2);
  S.doverflow.ensure(sizeof(int32_t) * 2);
  S.dstep.ensure(sizeof(int64_t));
  S.list.ensure(sizeof(int2) * (size_t)T);
  S.rowof.ensure(sizeof(int32_t) * (size_t)B * N);
  S.klist.ensure(sizeof(int2) * B);
  S.kT5.ensure(sizeof(float) * (size_t)B * F);
  S.kT5b.ensure(sizeof(float) * (size_t)B * F);
  S.kZ2.ensure(sizeof(float) * (size_t)B * 2 * F);
  S.p_M5k.ensure(sizeof(float) * (size_t)B * 5 * F);
  S.pcnt.ensure(sizeof(int32_t) * (size_t)T);
  S.pptr.ensure(sizeof(int64_t) * (size_t)(T + 2));
  S.pnbr.ensure(sizeof(int32_t) * (size_t)T * std::max(N - 1, 1));
  S.p_hc.ensure(sizeof(float) * (size_t)T * F);
  S.p_amx.ensure(sizeof(uint32_t) * (size_t)T);
  S.k_amx.ensure(sizeof(uint32_t) * (size_t)B);
  S.p_x.ensure(sizeof(float) * (size_t)T * XW);
  S.p_y.ensure(sizeof(float) * (size_t)T * XW);
  S.p_h.ensure(sizeof(float) * (size_t)T * F);
  S.p_orb.ensure(sizeof(float) * (size_t)T * ND);
  S.kcoef.ensure(sizeof(float) * (size_t)D * B * ME);
  S.p_envk.ensure(sizeof(float) * (size_t)B * ND);
  S.phin.ensure(sizeof(double) * (size_t)B * D * wb_vs_doubles(N));
  S.sbuf.ensure(sizeof(double) * (size_t)B * D * 64 * 64);
  S.ldS.ensure(sizeof(double) * (size_t)B * D);
  S.sgS.ensure(sizeof(double) * (size_t)B * D);
  S.wbscr.ensure((size_t)B * D * wb_update_scratch(N));
  S.jsum.ensure(sizeof(double) * (size_t)B * 2);
  S.p_jr.ensure(sizeof(float) * (size_t)T * 2);
  S.p_qv.ensure(sizeof(float) * (size_t)T * std::max(h->HDP, 1));
  S.p_av.ensure(sizeof(float) * (size_t)T * nc);
  S.rin.ensure(sizeof(float) * (size_t)B * 2 * std::max(HD, 1));
  S.ry1.ensure(sizeof(float) * (size_t)B * W);
  S.ry2.ensure(sizeof(float) * (size_t)B * W);
  S.att_out.ensure(sizeof(float) * (size_t)B * 2);
  S.cval.ensure(sizeof(double) * (size_t)B * N);
  S.clap.ensure(sizeof(double) * (size_t)B * N);
  S.cgrad.ensure(sizeof(double) * (size_t)B * N * 3);
  S.epot.ensure(sizeof(double) * (size_t)B * N);
  S.ccusp.ensure(sizeof(double) * (size_t)B);
  S.cdelta.ensure(sizeof(double) * (size_t)B);
  if (h->cusp) {
    // the walkers' e-e cusp totals; proposals add the change of the moved electron's pairs
    launch_cusp_coulomb(c.s, S.r.as<float>(), h->R, h->Z, B, N, h->n_up, h->M, h->cusp, h->P[h->iCUSP],
                        S.cval.as<double>(), S.clap.as<double>(), S.cgrad.as<double>(), S.epot.as<double>());
    launch_mc_cusp_total(c.s, S.cval.as<double>(), B, N, S.ccusp.as<double>());
  }
  for (DevBuf* x : {&S.p_ld, &S.p_sg}) x->ensure(sizeof(double) * (size_t)B * D);
  for (DevBuf* x : {&S.p_logpsi, &S.p_sign}) x->ensure(sizeof(double) * B);
  launch_fill_i32(c.s, S.doverflow.as<int32_t>(), 2, 0);
  launch_fill_i32(c.s, reinterpret_cast<int32_t*>(S.dstats.p), 4, 0);
  launch_set_i64(c.s, S.dstep.as<int64_t>(), 0);
  S.ready = true;
}

void state_commit(Ctx& c);

// Put the positions of a pending, uncommitted proposal back to the state's.
void state_drop_pending(Ctx& c) {
  auto& S = c.h->st;
  if (!S.pending) return;
  launch_fill_i32(c.s, S.accept.as<int32_t>(), S.B, 0);
  state_commit(c);
}

// Proposal for electron k of every walker: mode 0 Gaussian step (width, seed, step index
// *dstep + step), mode 1 new positions rk_in [B,3].  Leaves log psi / sign of the proposals in
// p_logpsi / p_sign and, in mode 0, the Metropolis decisions in accept (n_acc counts them).
// Every launch is sized by the static capacities and reads the actual row counts from the
// device, so the sequence has no host synchronisation and can be captured into a graph.
void state_propose(Ctx& c, int k, int mode, const float* rk_in, float width, uint64_t seed, uint64_t step,
                   int32_t* n_acc, bool warmup = false) {
  fire_handle* h = c.h;
  auto& S = h->st;
  REQUIRE(S.ready, "no walker state (call fire_state_init first)");
  REQUIRE(k >= 0 && k < h->N, "electron index out of range");
  state_drop_pending(c);
  const int B = S.B, N = h->N, F = h->F, ND = h->ND, D = h->D, W = h->W, nu = h->n_up;
  const int H = h->H, Dk = h->Dk, HD = h->HD;
  const int64_t T = S.tcap;
  const int amax = S.cap;
  Ctx cc{h, c.s, S.prec, c.fixed};
  int32_t* tail = S.tail.as<int32_t>();
  const int32_t* dT_up = tail + 0;
  const int32_t* dT_dn = tail + 1;
  const int32_t* dT = tail + 3;
  launch_mc_propose(c.s, S.r.as<float>(), B, N, k, width, seed, step, S.dstep.as<int64_t>(), mode, rk_in,
                    S.p_r.as<float>(), S.p_rk.as<float>(), S.p_u.as<double>());
  launch_mc_affected(c.s, S.r.as<float>(), S.p_r.as<float>(), B, N, nu, k, h->cfg.cutoff, S.cap, S.aff.as<int32_t>(),
                     S.cu.as<int32_t>(), S.cd.as<int32_t>(), S.doverflow.as<int32_t>());
  launch_mc_scan(c.s, S.cu.as<int32_t>(), S.cd.as<int32_t>(), B, S.offu.as<int32_t>(), S.offd.as<int32_t>(), tail,
                 warmup ? nullptr : S.dstats.as<int64_t>());
  launch_mc_rowlist(c.s, S.aff.as<int32_t>(), S.cu.as<int32_t>(), S.cd.as<int32_t>(), S.offu.as<int32_t>(),
                    S.offd.as<int32_t>(), tail, B, N, nu, S.list.as<int2>(), S.rowof.as<int32_t>());
  S.k = k;

  // h0 and messages of the moved electron (ElecInit on one electron per walker)
  const int b = h->iEN;
  launch_mc_klist(c.s, B, k, S.klist.as<int2>());
  launch_elec_init_en(c.s, S.p_r.as<float>(), h->R, B, N, nu, h->M, h->cfg.cutoff_1el, F, en_params(h), false,
                      S.kT5.as<float>(), S.klist.as<int2>(), B);
  // FP16 planes of the classes set to 4 take the row maxima of their inputs, as the value path
  // does, so the moved electron's rows match a fresh evaluation
  const bool f16e = c.prec == FIRE_MIXED && h->planes[0] == 4;
  const bool f16m = c.prec == FIRE_MIXED && h->planes[1] == 4;
  const bool f16o = c.prec == FIRE_MIXED && h->planes[2] == 4;
  ActEpi ek;
  if (f16e) ek.amax = S.k_amx.as<uint32_t>();
  if (f16e) launch_rowmax(c.s, S.kT5.as<float>(), F, B, F, S.k_amx.as<uint32_t>());
  gemm(cc, S.kT5.as<float>(), F, b + 11, -1, S.kZ2.as<float>(), 2 * F, B, 2 * F, F, nullptr, 0, DevRows(), &ek);
  launch_glu(c.s, S.kZ2.as<float>(), B, F, false, S.kT5b.as<float>());
  if (f16e) launch_rowmax(c.s, S.kT5b.as<float>(), F, B, F, S.k_amx.as<uint32_t>());
  gemm(cc, S.kT5b.as<float>(), F, b + 12, b + 13, S.kT5.as<float>(), F, B, F, F, nullptr, 1, DevRows(), &ek);
  const GroupDesc g1{1, 0, nullptr};
  launch_act(c.s, S.kT5.as<float>(), F, F, B, g1, ACT_SILU, 0.f);
  if (f16e) launch_rowmax(c.s, S.kT5.as<float>(), F, B, F, S.k_amx.as<uint32_t>());
  gemm(cc, S.kT5.as<float>(), F, b + 14, b + 15, S.p_M5k.as<float>(), 5 * F, B, 5 * F, F, nullptr, 1, DevRows(), &ek);

  // neighbour lists of the affected centres at the new positions, messages, update MLP
  const DevRows rT{dT, nullptr};
  launch_nbr_count(c.s, S.p_r.as<float>(), B, N, nu, h->cfg.cutoff, S.pcnt.as<int32_t>(), S.list.as<int2>(), (int)T,
                   dT);
  launch_scan(c.s, S.pcnt.as<int32_t>(), (int)T, S.pptr.as<int64_t>(), dT);
  launch_nbr_fill(c.s, S.p_r.as<float>(), B, N, nu, h->cfg.cutoff, S.pptr.as<int64_t>(), S.pnbr.as<int32_t>(),
                  S.list.as<int2>(), (int)T, dT);
  launch_messages(c.s, S.p_r.as<float>(), B, N, nu, S.pptr.as<int64_t>(), S.pnbr.as<int32_t>(), S.M5.as<float>(),
                  edge_params(h, h->iEEs), edge_params(h, h->iEEd), h->cut_learned, F, false, S.p_hc.as<float>(),
                  S.list.as<int2>(), (int)T, k, S.p_M5k.as<float>(), dT);
  const int m = h->iMLP;
  ActEpi et;
  if (f16m) et.amax = S.p_amx.as<uint32_t>();
  if (f16m) launch_rowmax(c.s, S.p_hc.as<float>(), F, T, F, S.p_amx.as<uint32_t>(), dT);
  gemm(cc, S.p_hc.as<float>(), F, m + 0, m + 1, S.p_x.as<float>(), F, T, F, F, nullptr, 1, rT, &et);
  launch_act(c.s, S.p_x.as<float>(), F, F, (int)T, g1, ACT_SILU, 0.f, dT);
  if (f16m) launch_rowmax(c.s, S.p_x.as<float>(), F, T, F, S.p_amx.as<uint32_t>(), dT);
  gemm(cc, S.p_x.as<float>(), F, m + 2, m + 3, S.p_h.as<float>(), F, T, F, F, nullptr, 1, rT, &et);
  launch_act(c.s, S.p_h.as<float>(), F, F, (int)T, g1, ACT_SILU, 0.f, dT);

  // orbital rows of the affected electrons (up-spin rows first), envelope row of electron k
  const int o = h->iORB;
  ActEpi eo;
  if (f16o) {
    launch_rowmax(c.s, S.p_h.as<float>(), F, T, F, S.p_amx.as<uint32_t>(), dT);
    eo.amax = S.p_amx.as<uint32_t>();  // the down-spin contraction offsets it with its rows
  }
  gemm(cc, S.p_h.as<float>(), F, o + 0, o + 1, S.p_orb.as<float>(), ND, T, ND, F, nullptr, 1, DevRows{dT_up, nullptr},
       &eo);
  gemm(cc, S.p_h.as<float>(), F, o + 2, o + 3, S.p_orb.as<float>(), ND, T, ND, F, nullptr, 1, DevRows{dT_dn, dT_up},
       &eo);
  const int ME = h->M * h->Eiso;
  launch_env_coef(c.s, S.p_r.as<float>(), h->R, B, N, nu, h->M, D, h->Eiso, h->P[o + 4], h->P[o + 6], false,
                  S.kcoef.as<float>(), S.klist.as<int2>(), B);
  {
    const int pi_i = o + (k < nu ? 5 : 7);
    launch_gemm_f32_batched(c.s, S.kcoef.as<float>(), ME, (int64_t)B * ME, h->Pt32[pi_i], (int64_t)ME * N,
                            S.p_envk.as<float>(), ND, N, B, N, ME, D);
  }

  // determinant ratios (Woodbury, FP64)
  launch_wb_prep(c.s, S.p_orb.as<float>(), S.p_envk.as<float>(), S.env.as<float>(), S.A.as<double>(),
                 S.Ainv.as<double>(), S.aff.as<int32_t>(), S.cu.as<int32_t>(), S.cd.as<int32_t>(),
                 S.rowof.as<int32_t>(), B, N, nu, D, k, S.phin.as<double>(), S.sbuf.as<double>(), S.ldS.as<double>(),
                 S.sgS.as<double>());

  // Jastrow rows of the affected electrons and the per-walker sums
  if (h->has_mlp) {
    const int j = h->iJM;
    float* cur = S.p_x.as<float>();
    float* nxt = S.p_y.as<float>();
    // the silu of the last hidden layer is applied inside k_jmlp_last
    gemm(cc, S.p_h.as<float>(), F, j + 0, j + 1, cur, W, T, W, F, nullptr, 1, rT);
    for (int l = 1; l < h->depth; ++l) {
      launch_act(c.s, cur, W, W, (int)T, g1, ACT_SILU, 0.f, dT);
      gemm(cc, cur, W, j + 2 * l, j + 2 * l + 1, nxt, W, T, W, W, nullptr, 1, rT);
      std::swap(cur, nxt);
    }
    const int jl = j + 2 * h->depth;
    launch_jmlp_last(c.s, cur, W, (int)T, g1, h->P[jl], h->P[jl + 1], h->P[jl + 2], h->P[jl + 3], S.p_jr.as<float>(),
                     dT);
  }
  const float* att_scale = nullptr;
  const float* att_zb = nullptr;
  if (h->has_att) {
    const int HDP = h->HDP;
    gemm(cc, S.p_h.as<float>(), F, h->iAV, h->iAV + 1, S.p_qv.as<float>(), HDP, T, HDP, F, nullptr, 1, rT);
    launch_att_group(c.s, S.p_qv.as<float>(), HDP, S.p_qv.as<float>() + HD, HDP, H, Dk, (int)T, g1, false,
                     S.p_av.as<float>(), dT);
  }
  if (h->has_mlp || h->has_att)
    launch_mc_jsum(c.s, S.jr.as<float>(), S.p_jr.as<float>(), S.av.as<float>(), S.p_av.as<float>(),
                   S.rowof.as<int32_t>(), B, N, nu, H, Dk, h->has_mlp, h->has_att, S.jsum.as<double>(),
                   S.rin.as<float>());
  if (h->has_att) {
    const int ac = h->iATT;
    float* cur = S.ry1.as<float>();
    float* nxt = S.ry2.as<float>();
    gemm(cc, S.rin.as<float>(), 2 * HD, ac + 3, ac + 4, cur, W, B, W, 2 * HD, nullptr, 1);
    launch_act(c.s, cur, W, W, B, g1, ACT_SILU, 0.f);
    for (int l = 1; l < h->depth; ++l) {
      gemm(cc, cur, W, ac + 3 + 2 * l, ac + 4 + 2 * l, nxt, W, B, W, W, nullptr, 1);
      launch_act(c.s, nxt, W, W, B, g1, ACT_SILU, 0.f);
      std::swap(cur, nxt);
    }
    const int al = ac + 3 + 2 * h->depth;
    launch_gemm_f32(c.s, cur, W, h->P[al], nullptr, h->P[al + 1], S.att_out.as<float>(), 2, B, 2, W, nullptr, 1);
    att_scale = h->P[al + 2];
    att_zb = h->P[al + 3];
  }
  // e-e cusp of the new positions: the state's total plus the change of the pairs (k, j)
  // (the reference recomputes the whole pair sum; equal up to FP64 rounding)
  if (h->cusp)
    launch_mc_cusp_delta(c.s, S.r.as<float>(), S.p_r.as<float>(), B, N, nu, k, h->P[h->iCUSP], S.cdelta.as<double>());
  launch_mc_accept(c.s, B, D, h->cfg.use_mlp_jastrow, h->cfg.use_log_jastrow, h->has_att, h->cusp, S.ld.as<double>(),
                   S.sg.as<double>(), S.ldS.as<double>(), S.sgS.as<double>(), S.jsum.as<double>(),
                   S.att_out.as<float>(), att_scale, att_zb, S.ccusp.as<double>(), S.cdelta.as<double>(),
                   S.logpsi.as<double>(),
                   S.p_u.as<double>(), mode, S.p_ld.as<double>(), S.p_sg.as<double>(), S.p_logpsi.as<double>(),
                   S.p_sign.as<double>(), S.accept.as<int32_t>(), n_acc);
  S.pending = true;
}

// Apply the pending proposal to the walkers whose accept flag is set.
void state_commit(Ctx& c) {
  fire_handle* h = c.h;
  auto& S = h->st;
  REQUIRE(S.pending, "no pending proposal to commit");
  const int B = S.B, N = h->N, nu = h->n_up, D = h->D;
  const int nc = h->H + h->HD;
  launch_wb_update(c.s, S.accept.as<int32_t>(), S.aff.as<int32_t>(), S.cu.as<int32_t>(), S.cd.as<int32_t>(), B, N, D,
                   S.phin.as<double>(), S.sbuf.as<double>(), S.wbscr.as<double>(), S.A.as<double>(),
                   S.Ainv.as<double>());
  launch_mc_commit(c.s, S.accept.as<int32_t>(), S.list.as<int2>(), (int)S.tcap, S.tail.as<int32_t>() + 3, B, N, nu, D,
                   h->F, h->ND, nc, S.k, S.p_h.as<float>(), S.p_orb.as<float>(), S.p_jr.as<float>(),
                   S.p_av.as<float>(), S.p_rk.as<float>(), S.p_M5k.as<float>(), S.p_envk.as<float>(),
                   S.p_ld.as<double>(), S.p_sg.as<double>(), S.p_logpsi.as<double>(), S.p_sign.as<double>(),
                   S.hh.as<float>(), S.orb.as<float>(), h->has_mlp ? S.jr.as<float>() : nullptr,
                   h->has_att ? S.av.as<float>() : nullptr, S.r.as<float>(), S.M5.as<float>(), S.env.as<float>(),
                   S.ld.as<double>(), S.sg.as<double>(), S.logpsi.as<double>(), S.sign.as<double>(),
                   S.p_r.as<float>(), S.cdelta.as<double>(), h->cusp ? S.ccusp.as<double>() : nullptr);
  S.pending = false;
}

// Raise if a proposal since init changed more electrons than the static capacity (the
// affected set was truncated and the state is invalid).  Synchronises the stream.
void state_check_overflow(Ctx& c) {
  auto& S = c.h->st;
  int32_t ov[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(ov, S.doverflow.p, sizeof(ov), hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  if (ov[0] != 0) {
    S.ready = S.pending = false;
    throw FireError("a move changed " + std::to_string(ov[1]) + " electrons, more than the " +
                    std::to_string(S.cap) +
                    " of the low-rank update's static capacity; the walker state is invalid (call fire_state_init)");
  }
}

// n_steps Metropolis steps from step0.  Whole sweeps starting at a multiple of n_el run as
// replays of one captured hipGraph (the step counter advances on the device); other steps
// are issued one by one.
void state_steps(Ctx& c, int64_t n_steps, float width, uint64_t seed, int64_t step0, bool use_graph) {
  fire_handle* h = c.h;
  auto& S = h->st;
  const int N = h->N;
  state_drop_pending(c);
  launch_set_i64(c.s, S.dstep.as<int64_t>(), 0);
  int64_t s = step0;
  const int64_t end = step0 + n_steps;
  // leading partial sweep
  while (s < end && (!use_graph || s % N != 0)) {
    state_propose(c, (int)(s % N), 0, nullptr, width, seed, (uint64_t)s, S.nacc.as<int32_t>());
    state_commit(c);
    S.n_proposals += S.B;
    ++s;
  }
  const int64_t n_sweeps = (end - s) / N;
  if (use_graph && n_sweeps > 0) {
    // (re)capture when the sweep parameters change; nodes take the step counter from the device
    if (!S.graph || S.g_width != width || S.g_seed != seed || S.g_sig != S.signature()) {
      if (S.graph) HIPCHK(hipGraphExecDestroy(S.graph));
      S.graph = nullptr;
      if (!S.gstream) HIPCHK(hipStreamCreateWithFlags(&S.gstream, hipStreamNonBlocking));
      // one eager zero-width proposal first (dropped, the state is unchanged), so that every
      // kernel attribute is set before the capture starts
      state_propose(c, 0, 0, nullptr, 0.f, seed, 0, nullptr, true);
      state_drop_pending(c);
      Ctx ce{h, S.gstream, c.prec, true};  // captured: no tuning or event timing
      HIPCHK(hipStreamSynchronize(c.s));
      HIPCHK(hipStreamBeginCapture(S.gstream, hipStreamCaptureModeRelaxed));
      for (int kk = 0; kk < N; ++kk) {
        state_propose(ce, kk, 0, nullptr, width, seed, (uint64_t)kk, S.nacc.as<int32_t>());
        state_commit(ce);
      }
      launch_add_i64(S.gstream, S.dstep.as<int64_t>(), N);
      hipGraph_t gr = nullptr;
      HIPCHK(hipStreamEndCapture(S.gstream, &gr));
      HIPCHK(hipGraphInstantiate(&S.graph, gr, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(gr));
      S.g_width = width;
      S.g_seed = seed;
      S.g_sig = S.signature();
      if (getenv("FIRE_LOG_CAPTURE")) fprintf(stderr, "fire: captured a sweep graph (B=%d, N=%d)\n", S.B, N);
    }
    // the graph's proposals use step = *dstep + k with *dstep = first step of the sweep
    launch_set_i64(c.s, S.dstep.as<int64_t>(), s);
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, c.s));
    HIPCHK(hipStreamWaitEvent(S.gstream, ev, 0));
    for (int64_t q = 0; q < n_sweeps; ++q) HIPCHK(hipGraphLaunch(S.graph, S.gstream));
    HIPCHK(hipEventRecord(ev, S.gstream));
    HIPCHK(hipStreamWaitEvent(c.s, ev, 0));
    HIPCHK(hipEventDestroy(ev));
    S.n_proposals += (int64_t)S.B * N * n_sweeps;
    s += n_sweeps * N;
    launch_set_i64(c.s, S.dstep.as<int64_t>(), 0);
  }
  while (s < end) {
    state_propose(c, (int)(s % N), 0, nullptr, width, seed, (uint64_t)s, S.nacc.as<int32_t>());
    state_commit(c);
    S.n_proposals += S.B;
    ++s;
  }
}

// CSR neighbour lists of B walkers in walker-major order (centre w * N + i) written straight
// into the caller's counts [B * N] and, when nbr != NULL, nbr [cap]; ptr (int64 [B * N + 2])
// receives the row pointer, the total and the largest count.  Returns the total, read back with
// one stream synchronisation (the capacity check needs it before the fill pass).
int64_t neighbours_walker_major(fire_handle* h, hipStream_t s, const float* r, int B, int32_t* counts, int32_t* nbr,
                                int64_t cap, int64_t* ptr) {
  const int N = h->N;
  const int G = B * N;
  launch_nbr_count(s, r, B, N, h->n_up, h->cfg.cutoff, counts, nullptr, 0, nullptr, true);
  launch_scan(s, counts, G, ptr);
  int64_t tail[2];
  HIPCHK(hipMemcpyAsync(tail, ptr + G, sizeof(tail), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (nbr) {
    REQUIRE(cap >= tail[0], "neighbour capacity too small: need " + std::to_string(tail[0]));
    launch_nbr_fill(s, r, B, N, h->n_up, h->cfg.cutoff, ptr, nbr, nullptr, 0, nullptr, true);
  }
  return tail[0];
}

int fail(const std::exception& e) {
  g_last_error = e.what();
  return 1;
}

// Walkers per chunk of the forward-Laplacian path: up to 256, fewer when the tangent-row
// workspace of one walker (about 5N + 3N * min(N - 1, 40) rows of the widest contraction)
// would put a single buffer above 12 GB.
int tangent_chunk(const fire_handle* h) {
  const int64_t N = h->N;
  const int64_t rows = 5 * N + 3 * N * std::min<int64_t>(N - 1, 40);
  const int64_t width =
      std::max<int64_t>(std::max<int64_t>(h->F, h->ND), std::max<int64_t>(h->W, std::max(2 * h->HD, h->HDP)));
  const int64_t per_walker = rows * width * 4;
  return (int)std::max<int64_t>(8, std::min<int64_t>(h->chunk, (int64_t)12000000000LL / per_walker));
}

bool valid_precision(int p) { return p == FIRE_FP32 || p == FIRE_MIXED || p == FIRE_BF16 || p == FIRE_MIXED_FAST; }

void check_ready(fire_handle* h) {
  REQUIRE(h != nullptr, "null handle");
  REQUIRE(h->params_ok, "parameters not uploaded (call fire_upload_params first)");
}

void check_config(const fire_config& c) {
  REQUIRE(c.n_el > 0 && c.n_up >= 0 && c.n_up <= c.n_el && c.n_nuc > 0, "invalid electron / nucleus counts");
  REQUIRE(c.feature_dim % 64 == 0 && c.feature_dim <= 256,
          "feature_dim must be a multiple of 64, at most 256 (one feature column per thread of a 256-thread "
          "workgroup in the embedding kernels)");
  REQUIRE(!c.use_attention || (c.att_dim <= 64 && (c.att_dim & (c.att_dim - 1)) == 0 && c.att_heads * c.att_dim <= 256),
          "attention_dim must be a power of two <= 64 and attention_heads * attention_dim <= 256");
  REQUIRE(c.filter_w0 > 0 && c.filter_w0 <= FIRE_W0MAX && (c.filter_w1 == 4 || c.filter_w1 == 8 || c.filter_w1 == 16),
          "pair_mlp_widths out of range (w0 <= 32; w1 zero-padded to 4, 8 or 16)");
  REQUIRE(c.n_env_en > 0 && c.n_env_en <= FIRE_EEMAX, "pair_n_envelopes out of range (<= 64)");
  REQUIRE(c.n_det > 0 && c.n_det <= 16 && c.n_el * c.n_det <= 2048,
          "n_determinants must be <= 16 and n_el * n_determinants <= 2048");
  REQUIRE(c.n_det * c.n_env_iso <= 1600, "n_determinants * n_envelopes must be <= 1600");
  REQUIRE(c.e_e_cusp == 0 || c.e_e_cusp == 1, "e_e_cusp must be 0 (none) or 1 (psiformer)");
  REQUIRE(c.mlp_depth >= 1 && c.mlp_width > 0, "mlp_depth >= 1 and mlp_width > 0 required");
}

// Every workspace buffer of the handle outside the walker state.
std::vector<DevBuf*> handle_bufs(fire_handle* h) {
  std::vector<DevBuf*> v{&h->AMX_E, &h->cnt,    &h->ptr,     &h->nbr,    &h->rev,     &h->rowflag, &h->T5,     &h->Z2,      &h->T5b,
          &h->M5,     &h->XA,      &h->XB,     &h->XC,      &h->O,       &h->ENV,    &h->ENVC,    &h->A64,
          &h->Ainv,   &h->AinvT,   &h->ldet,   &h->dsgn,    &h->tr,      &h->trip,   &h->dgrad,   &h->det_lp,
          &h->det_sg, &h->det_gr,  &h->det_lap, &h->JR,     &h->mlp_ns,  &h->mlp_dg, &h->QV,      &h->AV,
          &h->att_ns, &h->att_dg,  &h->RIN,    &h->RY1,     &h->RY2,     &h->att_out, &h->cval,   &h->clap,
          &h->cgrad,  &h->epot,    &h->PF,     &h->EF,      &h->EOK,     &h->rowgrp, &h->AV0,     &h->Z0,
          &h->F1,     &h->PART,    &h->AMX_A,  &h->AMX_B,   &h->AMX_C,  &h->ecp_atom, &h->ecp_grid, &h->ecp_vloc,
          &h->ecp_vnl, &h->ecp_cut, &h->ecp_nq, &h->ecp_qpts, &h->ecp_pts, &h->ecp_coef, &h->ecp_kq, &h->ecp_npts};
  for (DevBuf& b : h->grw) v.push_back(&b);
  return v;
}

void free_params(fire_handle* h) {
  for (auto* v : {&h->Pbf, &h->Pbf_lo, &h->Pbf_lo2, &h->Pf16, &h->Pf16_lo})
    for (uint16_t* p : *v)
      if (p) (void)hipFree(p);
  for (auto* v : {&h->P, &h->Pt32, &h->Pwinv})
    for (float* p : *v)
      if (p) (void)hipFree(p);
  for (auto* v : {&h->Pbf, &h->Pbf_lo, &h->Pbf_lo2, &h->Pf16, &h->Pf16_lo}) v->clear();
  for (auto* v : {&h->P, &h->Pt32, &h->Pwinv}) v->clear();
  h->params_ok = false;
}

// Molecule and model of the handle (fire_create, fire_reconfigure): the workspace buffers are
// kept (they grow on demand), parameters and the walker state belong to the old model and go.
void apply_config(fire_handle* h, const fire_config& c, const float* R, const float* Z) {
  free_params(h);
  if (h->st.graph) (void)hipGraphExecDestroy(h->st.graph);
  h->st.graph = nullptr;
  h->st.ready = h->st.pending = false;
  h->ecp_n = 0;  // pseudopotentials belong to the previous molecule
  h->cfg = c;
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
  h->HDP = h->HD + (c.att_heads + 15) / 16 * 16;
  h->depth = c.mlp_depth;
  h->has_mlp = c.use_mlp_jastrow || c.use_log_jastrow;
  h->has_att = c.use_attention != 0;
  h->cusp = c.e_e_cusp;
  h->iAV = h->iORBA = h->iJV = -1;
  h->amx_xa = nullptr;
  double enn = 0.0;
  for (int I = 0; I < c.n_nuc; ++I)
    for (int J = I + 1; J < c.n_nuc; ++J) {
      const double dx = (double)R[3 * I] - R[3 * J], dy = (double)R[3 * I + 1] - R[3 * J + 1],
                   dz = (double)R[3 * I + 2] - R[3 * J + 2];
      enn += (double)Z[I] * Z[J] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  h->e_nn = enn;
  if (h->R) HIPCHK(hipFree(h->R));
  if (h->Z) HIPCHK(hipFree(h->Z));
  h->R = h->Z = nullptr;
  HIPCHK(hipMalloc(&h->R, sizeof(float) * 3 * c.n_nuc));
  HIPCHK(hipMalloc(&h->Z, sizeof(float) * c.n_nuc));
  HIPCHK(hipMemcpy(h->R, R, sizeof(float) * 3 * c.n_nuc, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->Z, Z, sizeof(float) * c.n_nuc, hipMemcpyHostToDevice));
  build_expect(h);
}


// Parameter gradient of log|psi| for a chunk of Bc walkers (value path, FP32 forward with FP64
// determinants): gout [Bc][n_total] w.r.t. the packed tensors, logpsi [Bc] (optional).
void run_grad(Ctx& c, const float* r, int Bc, float* gout, int64_t gstride, const std::vector<int64_t>& off,
              double* logpsi_out) {
  fire_handle* h = c.h;
  hipStream_t s = c.s;
  const int N = h->N, nu = h->n_up, nd = h->n_dn, F = h->F, M = h->M, D = h->D, Ei = h->Eiso, ND = h->ND;
  const int W = h->W, depth = h->depth, H = h->H, Dk = h->Dk, HD = h->HD, w1 = h->w1, w0 = h->w0;
  const int64_t G = (int64_t)Bc * N, rowsEN = G * M;
  // one thread per feature in the LayerNorm kernels, per attention channel in the head kernels
  REQUIRE(F <= 256, "gradient: feature_dim above 256");
  REQUIRE(!h->has_att || (HD <= 256 && Dk <= 64 && (Dk & (Dk - 1)) == 0),
          "gradient: attention heads x dim above 256 or a dim that is not a power of two");
  int nb = 0;
  auto buf = [&](size_t bytes) -> void* {
    REQUIRE(nb < fire_handle::kGradSlots, "gradient workspace slots exhausted");
    DevBuf& b = h->grw[nb++];
    b.ensure(std::max<size_t>(bytes, 256));
    return b.p;
  };
  auto fb = [&](int64_t n) { return (float*)buf(sizeof(float) * (size_t)n); };
  auto g32 = [&](const float* A, int64_t lda, int wi, int bi, float* C, int64_t ldc, int64_t Mr, int Nc, int K) {
    launch_gemm_f32(s, A, lda, h->P[wi], h->Pt32[wi], bi >= 0 ? h->P[bi] : nullptr, C, ldc, Mr, Nc, K, nullptr, 1);
  };
  // C = A W^T for a weight W stored [K_w][N_w] row-major: the transposed product takes W itself
  // as its [N][K] operand
  // (products with at most 16 columns run on the small-N kernel, which reads the [K][N] operand:
  // the weight is transposed into a workspace slot of its own first)
  float* wsmall = nullptr;
  int64_t wsmall_n = 0;
  auto g32t = [&](const float* A, int64_t lda, const float* Wrow, float* C, int64_t ldc, int64_t Mr, int Nc, int K) {
    const float* Wk = nullptr;
    if (Nc <= 16) {
      REQUIRE((int64_t)Nc * K <= wsmall_n, "gradient: small-product workspace too small");
      launch_transpose2d(s, Wrow, Nc, K, wsmall);
      Wk = wsmall;
    }
    launch_gemm_f32(s, A, lda, Wk, Wrow, nullptr, C, ldc, Mr, Nc, K, nullptr, 1);
  };
  wsmall_n = std::max<int64_t>({(int64_t)M * Ei * N, (int64_t)F * W, (int64_t)W * W, 2 * (int64_t)W, 16 * (int64_t)F});
  wsmall = fb(wsmall_n);
  launch_fill_f32(s, gout, (int64_t)Bc * gstride, 0.f);
  const WRows both{nullptr, 0, nu, nu, (int64_t)Bc * nu, nd, nd};
  const WRows up{nullptr, 0, nu, nu, 0, 0, 0};
  const WRows dn{nullptr, (int64_t)Bc * nu, nd, nd, 0, 0, 0};
  // partial products of walkers whose rows are split over several workgroups
  const int64_t part_cap = (int64_t)16 << 20;
  float* wpart = fb(part_cap);
  auto wg = [&](const float* X, int64_t ldx, const float* Y, int64_t ldy, int K, int J, const WRows& wr, int64_t o,
                const uint8_t* mask = nullptr, int want = 1) {
    launch_wgrad(s, X, ldx, Y, ldy, K, J, Bc, wr, gout + o, gstride, false, mask, want, wpart, part_cap);
  };

  // ---- forward ---------------------------------------------------------------------------
  Graph gr = build_graph(c, r, Bc);
  const int64_t P = gr.n_pairs;
  const EnParams ep = en_params(h);
  float* en_rec = fb(rowsEN * PR_FLOATS);
  float* en_beta = fb(rowsEN * w1);
  float* en_lsf = fb(rowsEN * 4);
  uint8_t* en_mask = (uint8_t*)buf((size_t)rowsEN);
  launch_gr_en_fwd(s, r, h->R, Bc, N, nu, M, h->cfg.cutoff_1el, ep, rowsEN, en_rec, en_beta, en_lsf, en_mask);
  float* hen = fb(G * F);
  launch_gr_en_sum(s, en_beta, en_lsf, ep, en_mask, G, M, F, hen);
  float* y1 = fb(G * F);
  float* lnn = fb(G * F);
  float* lninv = fb(G);
  launch_gr_ln_fwd(s, hen, G, F, ep.ln_g, ep.ln_b, y1, lnn, lninv);
  const int b = h->iEN;
  float* z2 = fb(G * 2 * F);
  g32(y1, F, b + 11, -1, z2, 2 * F, G, 2 * F, F);
  float* y2 = fb(G * F);
  launch_gr_glu_fwd(s, z2, G, F, y2);
  float* z3 = fb(G * F);
  g32(y2, F, b + 12, b + 13, z3, F, G, F, F);
  float* y3 = fb(G * F);
  launch_silu_fwd(s, y3, z3, G * F);
  float* m5 = fb(G * 5 * F);
  g32(y3, F, b + 14, b + 15, m5, 5 * F, G, 5 * F, F);
  const EdgeParams eps = edge_params(h, h->iEEs), epd = edge_params(h, h->iEEd);
  float* hc = fb(G * F);
  launch_messages(s, r, Bc, N, nu, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), m5, eps, epd, h->cut_learned, F, false,
                  hc);
  const int m = h->iMLP;
  float* zu1 = fb(G * F);
  g32(hc, F, m + 0, m + 1, zu1, F, G, F, F);
  float* au1 = fb(G * F);
  launch_silu_fwd(s, au1, zu1, G * F);
  float* zu2 = fb(G * F);
  g32(au1, F, m + 2, m + 3, zu2, F, G, F, F);
  float* x = fb(G * F);
  launch_silu_fwd(s, x, zu2, G * F);
  // orbitals, envelopes, determinants
  const int o = h->iORB;
  const int64_t Gu = (int64_t)Bc * nu, Gd = (int64_t)Bc * nd;
  float* orb = fb(G * ND);
  g32(x, F, o + 0, o + 1, orb, ND, Gu, ND, F);
  if (Gd) g32(x + Gu * F, F, o + 2, o + 3, orb + Gu * ND, ND, Gd, ND, F);
  const int DME = D * M * Ei;
  float* ex = fb(G * DME);
  float* env = fb(G * ND);
  float* dist = fb(G * M);
  // exp(-sigma |r - R|) per (row, det, nucleus, envelope); the envelope sums as FP32 matrix-core
  // contractions per spin block, batched over determinants (pi transposed per determinant)
  launch_gr_env(s, r, h->R, Bc, N, nu, M, D, Ei, h->P[o + 4], h->P[o + 5], h->P[o + 6], h->P[o + 7], ex, nullptr,
                dist);
  launch_gemm_f32_batched(s, ex, DME, M * Ei, h->Pt32[o + 5], (int64_t)M * Ei * N, env, ND, N, Gu, N, M * Ei, D);
  if (Gd)
    launch_gemm_f32_batched(s, ex + Gu * DME, DME, M * Ei, h->Pt32[o + 7], (int64_t)M * Ei * N, env + Gu * ND, ND, N,
                            Gd, N, M * Ei, D);
  const int64_t nmat = (int64_t)Bc * D;
  double* A64 = (double*)buf(sizeof(double) * (size_t)nmat * N * N);
  double* Ainv = (double*)buf(sizeof(double) * (size_t)nmat * N * N);
  double* ldet = (double*)buf(sizeof(double) * (size_t)nmat);
  double* dsgn = (double*)buf(sizeof(double) * (size_t)nmat);
  launch_gr_phi(s, orb, env, Bc, N, nu, D, A64);
  const bool lu_lds = lu_lds_bytes(N) <= 160 * 1024;
  if (lu_lds && lu_lds_bytes(N) > 64 * 1024 && !h->lu_smem_set) {
    HIPCHK(lu_set_smem_limit(lu_lds_bytes(N)));
    h->lu_smem_set = true;
  }
  launch_lu_inv(s, A64, (int)nmat, N, Ainv, nullptr, ldet, dsgn, true, lu_lds);
  // Jastrow MLP
  const int jm = h->iJM;
  std::vector<float*> jz(depth + 1, nullptr), ja(depth, nullptr);
  if (h->has_mlp) {
    const float* in = x;
    int fan = F;
    for (int l = 0; l < depth; ++l) {
      jz[l] = fb(G * W);
      ja[l] = fb(G * W);
      g32(in, fan, jm + 2 * l, jm + 2 * l + 1, jz[l], W, G, W, fan);
      launch_silu_fwd(s, ja[l], jz[l], G * W);
      in = ja[l];
      fan = W;
    }
    jz[depth] = fb(G * 2);
    g32(in, W, jm + 2 * depth, jm + 2 * depth + 1, jz[depth], 2, G, 2, W);
  }
  // attention projections [q | v] from the packed [W_q | W_v]
  const int at = h->iATT;
  float* qv = nullptr;
  if (h->has_att) {
    float* qvT = fb((int64_t)2 * HD * F);
    launch_transpose2d(s, h->P[at], F, 2 * HD, qvT);
    qv = fb(G * 2 * HD);
    launch_gemm_f32(s, x, F, h->P[at], qvT, h->P[at + 1], qv, 2 * HD, G, 2 * HD, F, nullptr, 1);
  }

  // ---- per walker: determinants, Jastrow, attention readout, cusp -----------------------------
  double* wdet = (double*)buf(sizeof(double) * (size_t)nmat);
  double* lp = logpsi_out ? logpsi_out : (double*)buf(sizeof(double) * (size_t)Bc);
  float* dJ = fb((int64_t)Bc * 2);
  float* dval = fb((int64_t)Bc * 2 * std::max(HD, 1));
  float* dnorm = fb((int64_t)Bc * 2 * std::max(H, 1));
  GradHead ha{};
  ha.B = Bc;
  ha.N = N;
  ha.nu = nu;
  ha.D = D;
  ha.W = W;
  ha.depth = depth;
  ha.H = H;
  ha.Dk = Dk;
  ha.HD = HD;
  ha.lwmax = std::max(2 * HD, W);
  ha.has_mlp = h->has_mlp;
  ha.use_mlp = h->cfg.use_mlp_jastrow;
  ha.use_log = h->cfg.use_log_jastrow;
  ha.has_att = h->has_att;
  ha.cusp = h->cusp;
  ha.ldet = ldet;
  ha.dsgn = dsgn;
  ha.z3 = jz[depth];
  if (h->has_mlp) {
    ha.mlp_scale = h->P[jm + 2 * depth + 2];
    ha.mlp_zb = h->P[jm + 2 * depth + 3];
    ha.o_mlp_scale = off[jm + 2 * depth + 2];
    ha.o_mlp_zb = off[jm + 2 * depth + 3];
  }
  if (h->has_att) {
    REQUIRE(depth + 1 <= 8, "gradient: at most 7 hidden readout layers");
    ha.qv = qv;
    ha.keys = h->P[at + 2];
    for (int l = 0; l <= depth; ++l) {
      ha.ro_W[l] = h->P[at + 3 + 2 * l];
      ha.ro_b[l] = h->P[at + 4 + 2 * l];
      ha.o_ro_W[l] = off[at + 3 + 2 * l];
      ha.o_ro_b[l] = off[at + 4 + 2 * l];
    }
    ha.att_scale = h->P[at + 5 + 2 * depth];
    ha.att_zb = h->P[at + 6 + 2 * depth];
    ha.o_att_scale = off[at + 5 + 2 * depth];
    ha.o_att_zb = off[at + 6 + 2 * depth];
  }
  if (h->cusp) {
    ha.cusp_p = h->P[h->iCUSP];
    ha.o_cusp = off[h->iCUSP];
  }
  ha.r = r;
  ha.wdet = wdet;
  ha.logpsi = lp;
  ha.dJ = dJ;
  ha.dval = dval;
  ha.dnorm = dnorm;
  ha.gout = gout;
  ha.gstride = gstride;
  REQUIRE(gr_head_lds(ha) <= 64 * 1024, "gradient: readout layers too wide for one workgroup's LDS");
  launch_gr_head(s, ha);
  // ---- per row cotangents --------------------------------------------------------------------
  float* dorb = fb(G * ND);
  float* denv = fb(G * ND);
  float* dz3 = h->has_mlp ? fb(G * 2) : nullptr;
  float* dqv = h->has_att ? fb(G * 2 * HD) : nullptr;
  float* dkey = h->has_att ? fb(G * HD) : nullptr;
  GradRows ra{};
  ra.B = Bc;
  ra.N = N;
  ra.nu = nu;
  ra.D = D;
  ra.H = H;
  ra.Dk = Dk;
  ra.HD = HD;
  ra.has_mlp = h->has_mlp;
  ra.has_att = h->has_att;
  ra.wdet = wdet;
  ra.Ainv = Ainv;
  ra.orb = orb;
  ra.env = env;
  ra.z3 = jz[depth];
  ra.dJ = dJ;
  ra.mlp_scale = ha.mlp_scale;
  ra.qv = qv;
  ra.keys = ha.keys;
  ra.dval = dval;
  ra.dnorm = dnorm;
  ra.dorb = dorb;
  ra.denv = denv;
  ra.dz3 = dz3;
  ra.dqv = dqv;
  ra.dkey = dkey;
  launch_gr_rows(s, ra, G);

  // ---- orbitals and envelopes --------------------------------------------------------------
  wg(x, F, dorb, ND, F, ND, up, off[o + 0]);
  wg(nullptr, 0, dorb, ND, 1, ND, up, off[o + 1]);
  if (nd) {
    wg(x, F, dorb, ND, F, ND, dn, off[o + 2]);
    wg(nullptr, 0, dorb, ND, 1, ND, dn, off[o + 3]);
  }
  const int ME = M * Ei;
  float* tt = fb(G * DME);
  for (int d = 0; d < D; ++d) {
    wg(ex + (int64_t)d * ME, DME, denv + (int64_t)d * N, ND, ME, N, up, off[o + 5] + (int64_t)d * ME * N);
    g32t(denv + (int64_t)d * N, ND, h->P[o + 5] + (int64_t)d * ME * N, tt + (int64_t)d * ME, DME, Gu, ME, N);
    if (nd) {
      wg(ex + (int64_t)d * ME, DME, denv + (int64_t)d * N, ND, ME, N, dn, off[o + 7] + (int64_t)d * ME * N);
      g32t(denv + Gu * ND + (int64_t)d * N, ND, h->P[o + 7] + (int64_t)d * ME * N, tt + Gu * DME + (int64_t)d * ME,
           DME, Gd, ME, N);
    }
  }
  launch_gr_sp(s, ex, tt, dist, Bc, N, nu, M, D, Ei, gout, gstride, off[o + 4], off[o + 6]);
  // d x from the orbitals
  float* dx = fb(G * F);
  g32t(dorb, ND, h->P[o + 0], dx, F, Gu, F, ND);
  if (nd) g32t(dorb + Gu * ND, ND, h->P[o + 2], dx + Gu * F, F, Gd, F, ND);
  float* tmpF = fb(G * std::max(F, W));
  float* tmpF2 = fb(G * std::max(F, W));
  // ---- MLP Jastrow -----------------------------------------------------------------------
  if (h->has_mlp) {
    wg(ja[depth - 1], W, dz3, 2, W, 2, both, off[jm + 2 * depth]);
    wg(nullptr, 0, dz3, 2, 1, 2, both, off[jm + 2 * depth + 1]);
    float* da = tmpF;
    g32t(dz3, 2, h->P[jm + 2 * depth], da, W, G, W, 2);
    for (int l = depth - 1; l >= 0; --l) {
      const float* in = l == 0 ? x : ja[l - 1];
      const int fan = l == 0 ? F : W;
      float* dz = tmpF2;
      launch_silu_bwd(s, dz, W, da, W, jz[l], W, G, W);
      wg(in, fan, dz, W, fan, W, both, off[jm + 2 * l]);
      wg(nullptr, 0, dz, W, 1, W, both, off[jm + 2 * l + 1]);
      g32t(dz, W, h->P[jm + 2 * l], da, fan, G, fan, W);
    }
    launch_axpy(s, dx, da, G * F);
  }
  // ---- attention -----------------------------------------------------------------------
  if (h->has_att) {
    wg(x, F, dqv, 2 * HD, F, 2 * HD, both, off[at + 0]);
    wg(nullptr, 0, dqv, 2 * HD, 1, 2 * HD, both, off[at + 1]);
    wg(nullptr, 0, dkey, HD, 1, HD, both, off[at + 2]);
    g32t(dqv, 2 * HD, h->P[at], tmpF, F, G, F, 2 * HD);
    launch_axpy(s, dx, tmpF, G * F);
  }
  // ---- update MLP --------------------------------------------------------------------------
  float* dz = tmpF;
  launch_silu_bwd(s, dz, F, dx, F, zu2, F, G, F);
  wg(au1, F, dz, F, F, F, both, off[m + 2]);
  wg(nullptr, 0, dz, F, 1, F, both, off[m + 3]);
  g32t(dz, F, h->P[m + 2], tmpF2, F, G, F, F);
  launch_silu_bwd(s, dz, F, tmpF2, F, zu1, F, G, F);
  wg(hc, F, dz, F, F, F, both, off[m + 0]);
  wg(nullptr, 0, dz, F, 1, F, both, off[m + 1]);
  float* dhc = fb(G * F);
  g32t(dz, F, h->P[m + 0], dhc, F, G, F, F);

  // ---- messages and e-e edges -------------------------------------------------------------------
  float* dm5 = fb(G * 5 * F);
  HIPCHK(hipMemcpy2DAsync(dm5, sizeof(float) * 5 * F, dhc, sizeof(float) * F, sizeof(float) * F, (size_t)G,
                          hipMemcpyDeviceToDevice, s));
  if (P > 0) {
    float* prec = fb(P * PR_FLOATS);
    float* pbeta = fb(P * w1);
    float* plsf = fb(P * 4);
    uint8_t* psame = (uint8_t*)buf((size_t)P);
    int32_t* pctr = (int32_t*)buf(sizeof(int32_t) * (size_t)P);
    launch_gr_pair_fwd(s, r, Bc, N, nu, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), eps, epd, h->cut_learned, P, prec,
                       pbeta, plsf, psame, pctr);
    // per edge class: Gamma / e cotangents, d beta and the filter backward in one pass (du kept
    // for the message collection), then the messages' rows of the centres and neighbours
    float* pdu = fb(P * F);
    const int64_t nblk = (N + 3) / 4;
    float* ppart = fb((int64_t)Bc * nblk * ((int64_t)(w1 + 5) * F + 6 * w0 + 2 * w0 * w1 + w1 + 1));
    for (int k = 0; k < 2; ++k) {
      const int base = k == 0 ? h->iEEs : h->iEEd;
      launch_gr_pair_cot(s, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), psame, k == 0 ? 1 : 0, pbeta, plsf, prec, m5,
                         dhc, k == 0 ? eps : epd, h->cut_learned, Bc, N, nu, F, pdu, off[base + 0], off[base + 1],
                         off[base + 2], off[base + 3], off[base + 4], off[base + 5], off[base + 6], off[base + 7],
                         off[base + 8], off[h->iSCAL], k == 1, ppart, gout, gstride);
    }
    launch_gr_msg_collect(s, pdu, h->ptr.as<int64_t>(), h->nbr.as<int32_t>(), h->rev.as<int32_t>(), psame, Bc, N, nu, F,
                          dm5);
  }
  // ---- ElecInit dense stage ------------------------------------------------------------------
  wg(y3, F, dm5, 5 * F, F, 5 * F, both, off[b + 14]);
  wg(nullptr, 0, dm5, 5 * F, 1, 5 * F, both, off[b + 15]);
  float* dy = tmpF;
  g32t(dm5, 5 * F, h->P[b + 14], dy, F, G, F, 5 * F);
  float* dz3e = tmpF2;
  launch_silu_bwd(s, dz3e, F, dy, F, z3, F, G, F);
  wg(y2, F, dz3e, F, F, F, both, off[b + 12]);
  wg(nullptr, 0, dz3e, F, 1, F, both, off[b + 13]);
  g32t(dz3e, F, h->P[b + 12], dy, F, G, F, F);
  float* dz2 = fb(G * 2 * F);
  launch_gr_glu_bwd(s, dy, z2, G, F, dz2);
  wg(y1, F, dz2, 2 * F, F, 2 * F, both, off[b + 11]);
  float* dy1 = tmpF2;
  g32t(dz2, 2 * F, h->P[b + 11], dy1, F, G, F, 2 * F);
  launch_mul(s, tmpF, dy1, lnn, G * F);
  wg(nullptr, 0, tmpF, F, 1, F, both, off[b + 9]);
  wg(nullptr, 0, dy1, F, 1, F, both, off[b + 10]);
  float* dhen = fb(G * F);
  launch_gr_ln_bwd(s, dy1, lnn, lninv, ep.ln_g, G, F, dhen);
  // ---- e-n edges -----------------------------------------------------------------------------
  // the e-n edge backward in one pass: Gamma / e cotangents and d beta in registers, then the
  // filter backward per edge; WG, WF, Wd, bd, We per walker, K, b, sp, h_nuc per nucleus
  float* enpart = fb((int64_t)Bc * M * (w1 + 4) * F);
  float* enpart2 = fb((int64_t)Bc * M * w1 * (w0 + ep.E + 1));
  launch_gr_en_cot(s, en_beta, en_lsf, en_mask, dhen, en_rec, h->cfg.cutoff_1el, ep, Bc, N, nu, M, F, enpart, enpart2,
                   gout, gstride, off[b + 7], off[b + 8], off[b + 3], off[b + 0], off[b + 1], off[b + 2], off[b + 4],
                   off[b + 5], off[b + 6]);
  HIPCHK(hipGetLastError());
}
}  // namespace

extern "C" {

const char* fire_last_error(void) { return g_last_error.c_str(); }

int fire_create(const fire_config* cfg, const float* R, const float* Z, int device, fire_handle** out) {
  try {
    REQUIRE(cfg && R && Z && out, "null argument");
    check_config(*cfg);
    HIPCHK(hipSetDevice(device));
    int64_t* scr = scratch_address(device);
    fire_handle* h = new fire_handle();
    h->scratch = scr;
    if (const char* e = getenv("FIRE_FUSED_ACT")) h->fuse_act = e[0] != '0';
    h->device = device;
    try {
      apply_config(h, *cfg, R, Z);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_reconfigure(fire_handle* h, const fire_config* cfg, const float* R, const float* Z) {
  try {
    REQUIRE(h && cfg && R && Z, "null argument");
    check_config(*cfg);
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipDeviceSynchronize());  // work still queued on the old molecule must finish first
    apply_config(h, *cfg, R, Z);
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_workspace_bytes(fire_handle* h, int64_t* bytes) {
  try {
    REQUIRE(h && bytes, "null argument");
    int64_t n = 0;
    for (DevBuf* b : handle_bufs(h)) n += (int64_t)b->bytes;
    for (DevBuf* b : h->st.bufs()) n += (int64_t)b->bytes;
    for (size_t k = 0; k < h->P.size(); ++k) n += 4 * (k < h->expect.size() ? h->expect[k] : 0);
    *bytes = n;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_destroy(fire_handle* h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  free_params(h);
  if (h->R) (void)hipFree(h->R);
  if (h->Z) (void)hipFree(h->Z);
  for (DevBuf* b : handle_bufs(h)) b->release();
  h->st.release();
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
    free_params(h);
    const int n_all = n_tensors + (h->has_att ? (h->has_mlp ? 8 : 6) : 0);
    h->P.assign(n_all, nullptr);
    h->Pbf.assign(n_all, nullptr);
    h->Pbf_lo.assign(n_all, nullptr);
    h->Pbf_lo2.assign(n_all, nullptr);
    h->Pt32.assign(n_all, nullptr);
    h->Pf16.assign(n_all, nullptr);
    h->Pf16_lo.assign(n_all, nullptr);
    h->Pwinv.assign(n_all, nullptr);
    h->Pwn.assign(n_all, 0);
    // host copies of every tensor (uploaded, then derived)
    std::vector<const float*> host(n_all, nullptr);
    std::vector<int64_t> hsize(n_all, 0);
    size_t off = 0;
    for (int k = 0; k < n_tensors; ++k) {
      host[k] = packed + off;
      hsize[k] = sizes[k];
      off += sizes[k];
    }
    std::vector<float> derived[8];
    if (h->has_att) {
      const int F = h->F, H = h->H, Dk = h->Dk, HD = h->HD, HDP = h->HDP;
      const float* wq = host[h->iATT];  // [F, 2 HD]: query columns, then value columns
      const float* bq = host[h->iATT + 1];
      const float* key = host[h->iATT + 2];
      derived[0].assign((size_t)F * HDP, 0.f);
      derived[1].assign(HDP, 0.f);
      for (int f = 0; f < F; ++f) {
        for (int n = 0; n < HD; ++n) derived[0][(size_t)f * HDP + n] = wq[(size_t)f * 2 * HD + HD + n];
        for (int hh = 0; hh < H; ++hh) {
          double a = 0.0;
          for (int k = 0; k < Dk; ++k) a += (double)wq[(size_t)f * 2 * HD + hh * Dk + k] * key[hh * Dk + k];
          derived[0][(size_t)f * HDP + HD + hh] = (float)a;
        }
      }
      for (int n = 0; n < HD; ++n) derived[1][n] = bq[HD + n];
      for (int hh = 0; hh < H; ++hh) {
        double a = 0.0;
        for (int k = 0; k < Dk; ++k) a += (double)bq[hh * Dk + k] * key[hh * Dk + k];
        derived[1][HD + hh] = (float)a;
      }
      h->iAV = n_tensors;
      // columns [c0, c0 + n) of the [F, HDP] table above: value half (0, HD), head sums (HD, H)
      auto cat = [&](const float* w1, int n1, const float* b1, int c0, int n, std::vector<float>& w,
                     std::vector<float>& bb) {
        w.assign((size_t)F * (n1 + n), 0.f);
        bb.assign(n1 + n, 0.f);
        for (int f = 0; f < F; ++f) {
          for (int q = 0; q < n1; ++q) w[(size_t)f * (n1 + n) + q] = w1[(size_t)f * n1 + q];
          for (int q = 0; q < n; ++q) w[(size_t)f * (n1 + n) + n1 + q] = derived[0][(size_t)f * HDP + c0 + q];
        }
        for (int q = 0; q < n1; ++q) bb[q] = b1[q];
        for (int q = 0; q < n; ++q) bb[n1 + q] = derived[1][c0 + q];
      };
      const int o = h->iORB, ND = h->ND;
      cat(host[o + 0], ND, host[o + 1], HD, H, derived[2], derived[3]);
      cat(host[o + 2], ND, host[o + 3], HD, H, derived[4], derived[5]);
      h->iORBA = n_tensors + 2;
      int nd = 6;
      if (h->has_mlp) {
        cat(host[h->iJM + 0], h->W, host[h->iJM + 1], 0, HD, derived[6], derived[7]);
        h->iJV = n_tensors + 6;
        nd = 8;
      }
      for (int q = 0; q < nd; ++q) {
        host[n_tensors + q] = derived[q].data();
        hsize[n_tensors + q] = (int64_t)derived[q].size();
      }
    }
    for (int k = 0; k < n_all; ++k) {
      HIPCHK(hipMalloc(&h->P[k], sizeof(float) * std::max<int64_t>(hsize[k], 1)));
      HIPCHK(hipMemcpy(h->P[k], host[k], sizeof(float) * hsize[k], hipMemcpyHostToDevice));
    }
    std::vector<std::pair<int, std::pair<int, int>>> gw;
    gemm_weights(h, gw);
    for (auto& t : gw) {
      const int k = t.first, K = t.second.first, Nn = t.second.second;
      const size_t Kp = ((size_t)K + 15) / 16 * 16;  // K padded to whole 16-wide chunks (zeros)
      auto at = [Nn](int kk, int n) { return ((size_t)(kk >> 4) * Nn + n) * 16 + (kk & 15); };
      h->Pwn[k] = Nn;
      // three BF16 planes, K-blocked: x = hi + mid + lo (remainders exact in FP32)
      std::vector<uint16_t> pl[3];
      for (auto& v : pl) v.assign(Kp * Nn, 0);
      const float* src = host[k];
      auto bf2f_h = [](uint16_t b) {
        uint32_t u = (uint32_t)b << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
      };
      for (int kk = 0; kk < K; ++kk)
        for (int n = 0; n < Nn; ++n) {
          const float x = src[(size_t)kk * Nn + n];
          const uint16_t hi = host_f2bf(x);
          const float r1 = x - bf2f_h(hi);
          const uint16_t mid = host_f2bf(r1);
          pl[0][at(kk, n)] = hi;
          pl[1][at(kk, n)] = mid;
          pl[2][at(kk, n)] = host_f2bf(r1 - bf2f_h(mid));
        }
      uint16_t** dst[3] = {&h->Pbf[k], &h->Pbf_lo[k], &h->Pbf_lo2[k]};
      for (int q = 0; q < 3; ++q) {
        HIPCHK(hipMalloc(dst[q], sizeof(uint16_t) * pl[q].size()));
        HIPCHK(hipMemcpy(*dst[q], pl[q].data(), sizeof(uint16_t) * pl[q].size(), hipMemcpyHostToDevice));
      }
      // two FP16 planes of the column-scaled weight, K-blocked: s_n W = hi + lo
      std::vector<uint16_t> fh(Kp * Nn, 0), fl(Kp * Nn, 0);
      std::vector<float> winv(Nn);
      for (int n = 0; n < Nn; ++n) {
        float mx = 0.f;
        for (int kk = 0; kk < K; ++kk) mx = std::max(mx, std::fabs(src[(size_t)kk * Nn + n]));
        const float sc = host_f16_scale(mx);
        winv[n] = 1.0f / sc;  // a power of two: exact
        for (int kk = 0; kk < K; ++kk) {
          const float x = src[(size_t)kk * Nn + n] * sc;
          const uint16_t hi = host_f2h(x);
          fh[at(kk, n)] = hi;
          fl[at(kk, n)] = host_f2h(x - host_h2f(hi));
        }
      }
      HIPCHK(hipMalloc(&h->Pf16[k], sizeof(uint16_t) * fh.size()));
      HIPCHK(hipMemcpy(h->Pf16[k], fh.data(), sizeof(uint16_t) * fh.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&h->Pf16_lo[k], sizeof(uint16_t) * fl.size()));
      HIPCHK(hipMemcpy(h->Pf16_lo[k], fl.data(), sizeof(uint16_t) * fl.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&h->Pwinv[k], sizeof(float) * winv.size()));
      HIPCHK(hipMemcpy(h->Pwinv[k], winv.data(), sizeof(float) * winv.size(), hipMemcpyHostToDevice));
    }
    // FP32 transposed copies [N,K] for the FP32 MFMA kernel: every GEMM weight, and the
    // envelope weights pi per determinant ([det][ME][N] -> [det][N][ME])
    auto upload_t = [&](int k, int nb, int K, int Nn) {
      std::vector<float> tt((size_t)nb * K * Nn);
      const float* src = host[k];
      for (int b = 0; b < nb; ++b)
        for (int kk = 0; kk < K; ++kk)
          for (int n = 0; n < Nn; ++n)
            tt[(size_t)b * K * Nn + (size_t)n * K + kk] = src[(size_t)b * K * Nn + (size_t)kk * Nn + n];
      HIPCHK(hipMalloc(&h->Pt32[k], sizeof(float) * tt.size()));
      HIPCHK(hipMemcpy(h->Pt32[k], tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice));
    };
    for (auto& t : gw) upload_t(t.first, 1, t.second.first, t.second.second);
    for (int k : {h->iORB + 5, h->iORB + 7}) upload_t(k, h->D, h->M * h->Eiso, h->N);
    // a captured sweep graph holds the old parameter pointers, and the resident walker state
    // (embeddings, orbital rows, inverses, log psi) was computed with the old parameters
    if (h->st.graph) {
      HIPCHK(hipGraphExecDestroy(h->st.graph));
      h->st.graph = nullptr;
    }
    h->st.ready = h->st.pending = false;
    h->cut_learned = host[h->iSCAL][0];
    h->s1 = host[h->iSCAL][1];
    h->params_ok = true;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_profile_gemm(fire_handle* h, int32_t enable, double* total_ms, double* total_flops, double* total_bytes,
                      int64_t* n_launches) {
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
    if (total_bytes) *total_bytes = h->gemm_bytes;
    if (n_launches) *n_launches = (int64_t)h->gemm_events.size();
    h->gemm_events.clear();
    h->gemm_flops = 0.0;
    h->gemm_bytes = 0.0;
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
    h->ptr.ensure(sizeof(int64_t) * ((size_t)B * h->N + 2));
    *n_pairs = neighbours_walker_major(h, (hipStream_t)stream, r_dev, B, counts_dev, nbr_dev, cap, h->ptr.as<int64_t>());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_neighbours_noalloc(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t* counts_dev,
                            int32_t* nbr_dev, int64_t cap, int64_t* n_pairs) {
  try {
    REQUIRE(h && r_dev && counts_dev && n_pairs, "null argument");
    REQUIRE(B >= 0, "negative batch");
    REQUIRE(h->scratch != nullptr, "no static scratch address for this device");
    const int64_t need = (int64_t)B * h->N + 2;
    REQUIRE(need <= kScratchInt64, "fire_neighbours_noalloc: B * n_el = " + std::to_string((int64_t)B * h->N) +
                                       " exceeds the static scratch capacity of " + std::to_string(kScratchInt64 - 2) +
                                       " electrons");
    HIPCHK(hipSetDevice(h->device));
    if (B == 0) {
      *n_pairs = 0;
      return 0;
    }
    *n_pairs = neighbours_walker_major(h, (hipStream_t)stream, r_dev, B, counts_dev, nbr_dev, cap, h->scratch);
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
    REQUIRE(valid_precision(precision), "unknown precision mode");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, precision};
    const int N = h->N;
    const int chunk = tangent_chunk(h);
    for (int w0 = 0; w0 < B; w0 += chunk) {
      const int Bc = std::min(chunk, B - w0);
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
    REQUIRE(valid_precision(precision), "unknown precision mode");
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

int fire_state_init(fire_handle* h, void* stream, const float* r_dev, int32_t B, int32_t precision, double* logpsi_dev,
                    double* sign_dev) {
  try {
    check_ready(h);
    REQUIRE(r_dev, "null argument");
    REQUIRE(B > 0, "batch must be positive");
    REQUIRE(valid_precision(precision), "unknown precision mode");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, precision};
    state_init(c, r_dev, B);
    if (logpsi_dev)
      HIPCHK(hipMemcpyAsync(logpsi_dev, h->st.logpsi.p, sizeof(double) * B, hipMemcpyDeviceToDevice, c.s));
    if (sign_dev) HIPCHK(hipMemcpyAsync(sign_dev, h->st.sign.p, sizeof(double) * B, hipMemcpyDeviceToDevice, c.s));
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_state_propose(fire_handle* h, void* stream, int32_t k, const float* rk_dev, double* logpsi_dev,
                       double* sign_dev) {
  try {
    check_ready(h);
    REQUIRE(rk_dev, "null argument");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, h->st.prec};
    state_propose(c, k, 1, rk_dev, 0.f, 0, 0, nullptr);
    h->st.n_proposals += h->st.B;
    state_check_overflow(c);
    const int B = h->st.B;
    if (logpsi_dev)
      HIPCHK(hipMemcpyAsync(logpsi_dev, h->st.p_logpsi.p, sizeof(double) * B, hipMemcpyDeviceToDevice, c.s));
    if (sign_dev) HIPCHK(hipMemcpyAsync(sign_dev, h->st.p_sign.p, sizeof(double) * B, hipMemcpyDeviceToDevice, c.s));
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_state_commit(fire_handle* h, void* stream, const int32_t* accept_dev) {
  try {
    check_ready(h);
    REQUIRE(accept_dev, "null argument");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, h->st.prec};
    REQUIRE(h->st.ready, "no walker state (call fire_state_init first)");
    REQUIRE(h->st.pending, "no pending proposal to commit");
    HIPCHK(hipMemcpyAsync(h->st.accept.p, accept_dev, sizeof(int32_t) * h->st.B, hipMemcpyDeviceToDevice, c.s));
    state_commit(c);
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_mcmc_steps(fire_handle* h, void* stream, int32_t n_steps, float width, uint64_t seed, int64_t step0,
                    int32_t* n_accept_dev) {
  try {
    check_ready(h);
    REQUIRE(n_steps >= 0, "negative step count");
    REQUIRE(step0 >= 0, "negative step index");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, h->st.prec};
    REQUIRE(h->st.ready, "no walker state (call fire_state_init first)");
    launch_fill_i32(c.s, h->st.nacc.as<int32_t>(), h->st.B, 0);
    state_steps(c, n_steps, width, seed, step0, h->use_graphs);
    if (n_accept_dev)
      HIPCHK(hipMemcpyAsync(n_accept_dev, h->st.nacc.p, sizeof(int32_t) * h->st.B, hipMemcpyDeviceToDevice, c.s));
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_state_get(fire_handle* h, void* stream, float* r_dev, double* logpsi_dev, double* sign_dev) {
  try {
    REQUIRE(h, "null handle");
    REQUIRE(h->st.ready, "no walker state (call fire_state_init first)");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, h->st.prec};
    state_drop_pending(c);
    state_check_overflow(c);
    const int B = h->st.B;
    if (r_dev)
      HIPCHK(hipMemcpyAsync(r_dev, h->st.r.p, sizeof(float) * (size_t)B * h->N * 3, hipMemcpyDeviceToDevice, c.s));
    if (logpsi_dev)
      HIPCHK(hipMemcpyAsync(logpsi_dev, h->st.logpsi.p, sizeof(double) * B, hipMemcpyDeviceToDevice, c.s));
    if (sign_dev) HIPCHK(hipMemcpyAsync(sign_dev, h->st.sign.p, sizeof(double) * B, hipMemcpyDeviceToDevice, c.s));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_autotune(fire_handle* h, int32_t enable, int32_t* table, int32_t cap, int32_t* n_entries) {
  try {
    REQUIRE(h, "null handle");
    int n = 0;
    for (auto& kv : h->gemm_tile) {
      if (table && n < cap) {
        int bm, bn;
        gemm_bf16_tile(kv.second, &bm, &bn);
        int32_t* e = table + 7 * n;
        e[0] = std::get<0>(kv.first);
        e[1] = std::get<1>(kv.first);
        e[2] = std::get<2>(kv.first);
        e[3] = std::get<3>(kv.first);
        e[4] = kv.second;
        e[5] = bm;
        e[6] = bn;
      }
      ++n;
    }
    if (n_entries) *n_entries = n;
    REQUIRE(enable <= 3 || (enable >= 100 && enable < 100 + gemm_bf16_variants()), "autotune mode out of range");
    if (enable >= 100) {
      h->force_tile = enable - 100;
    } else if (enable >= 0) {
      h->force_tile = -1;
      h->autotune = enable != 0;
      h->autotune_near = enable == 3;
    }
    if (enable == 2) {
      h->gemm_tile.clear();
      h->gemm_tile_ms.clear();
      std::lock_guard<std::mutex> lk(tile_cache().mu);
      auto& b = tile_cache().best;
      for (auto i = b.begin(); i != b.end();) i = std::get<0>(i->first) == h->device ? b.erase(i) : std::next(i);
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_tile_forms(void) { return gemm_bf16_variants(); }

int fire_mcmc_options(fire_handle* h, int32_t use_graphs) {
  try {
    REQUIRE(h, "null handle");
    h->use_graphs = use_graphs != 0;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_precision_plan(fire_handle* h, const int32_t* planes, int32_t n, int32_t* current) {
  try {
    REQUIRE(h, "null handle");
    REQUIRE(n >= 0 && n <= 6, "at most 6 contraction classes");
    for (int k = 0; k < n; ++k)
      REQUIRE(planes[k] >= 1 && planes[k] <= 4, "planes per operand must be 1, 2 or 3 (BF16) or 4 (FP16 pair)");
    for (int k = 0; k < n; ++k) h->planes[k] = planes[k];
    if (n > 0 && h->st.graph) {  // the captured sweep holds the previous contraction choices
      HIPCHK(hipGraphExecDestroy(h->st.graph));
      h->st.graph = nullptr;
    }
    if (current)
      for (int k = 0; k < 6; ++k) current[k] = h->planes[k];
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_state_stats(fire_handle* h, int32_t* n_walkers, int64_t* n_proposals, int64_t* sum_changed,
                     int64_t* max_changed) {
  try {
    REQUIRE(h, "null handle");
    if (n_walkers) *n_walkers = h->st.ready ? h->st.B : 0;
    if (n_proposals) *n_proposals = h->st.n_proposals;
    int64_t st[2] = {0, 0};
    int32_t ov[2] = {0, 0};
    if (h->st.doverflow.p && (sum_changed || max_changed)) {
      HIPCHK(hipSetDevice(h->device));
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipMemcpy(st, h->st.dstats.p, sizeof(st), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(ov, h->st.doverflow.p, sizeof(ov), hipMemcpyDeviceToHost));
    }
    if (sum_changed) *sum_changed = st[0];
    // the largest affected set, before truncation to the capacity when a move exceeded it
    if (max_changed) *max_changed = std::max<int64_t>(st[1], ov[1]);
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_ecp_upload(fire_handle* h, int32_t n_ecp, const int32_t* atom, int32_t n_l, int32_t n_grid,
                    const double* grid, const double* v_loc, const double* v_nl, const double* cutoff,
                    const int32_t* n_quad, int32_t q_max, const double* quad_dirs, const float* z_eff) {
  try {
    REQUIRE(h, "null handle");
    REQUIRE(n_ecp >= 0 && n_ecp <= h->M, "fire_ecp_upload: n_ecp out of range");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipDeviceSynchronize());
    if (z_eff) {
      // Coulomb terms of fire_local_energy use these charges from now on
      std::vector<float> R(3 * h->M);
      HIPCHK(hipMemcpy(R.data(), h->R, sizeof(float) * 3 * h->M, hipMemcpyDeviceToHost));
      double enn = 0.0;
      for (int I = 0; I < h->M; ++I)
        for (int J = I + 1; J < h->M; ++J) {
          const double dx = (double)R[3 * I] - R[3 * J], dy = (double)R[3 * I + 1] - R[3 * J + 1],
                       dz = (double)R[3 * I + 2] - R[3 * J + 2];
          enn += (double)z_eff[I] * z_eff[J] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
      h->e_nn = enn;
      HIPCHK(hipMemcpy(h->Z, z_eff, sizeof(float) * h->M, hipMemcpyHostToDevice));
    }
    h->ecp_n = 0;
    if (n_ecp == 0) return 0;
    REQUIRE(atom && grid && v_loc && v_nl && cutoff && n_quad && quad_dirs, "null argument");
    REQUIRE(n_l >= 0 && n_l <= 4, "fire_ecp_upload: non-local channels must be l = 0 .. 3");
    REQUIRE(n_grid >= 2 && q_max >= 1, "fire_ecp_upload: empty radial grid or quadrature");
    for (int e = 0; e < n_ecp; ++e) {
      REQUIRE(atom[e] >= 0 && atom[e] < h->M, "fire_ecp_upload: atom index out of range");
      REQUIRE(n_quad[e] >= 1 && n_quad[e] <= q_max, "fire_ecp_upload: quadrature size out of range");
    }
    for (int g = 1; g < n_grid; ++g) REQUIRE(grid[g] >= grid[g - 1], "fire_ecp_upload: radial grid not ascending");
    auto put = [](DevBuf& b, const void* src, size_t bytes) {
      b.ensure(bytes);
      HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    put(h->ecp_atom, atom, sizeof(int32_t) * n_ecp);
    put(h->ecp_grid, grid, sizeof(double) * n_grid);
    put(h->ecp_vloc, v_loc, sizeof(double) * n_ecp * n_grid);
    if (n_l > 0) put(h->ecp_vnl, v_nl, sizeof(double) * n_ecp * n_l * n_grid);
    put(h->ecp_cut, cutoff, sizeof(double) * n_ecp);
    put(h->ecp_nq, n_quad, sizeof(int32_t) * n_ecp);
    put(h->ecp_qpts, quad_dirs, sizeof(double) * n_ecp * q_max * 3);
    h->ecp_n = n_ecp;
    h->ecp_nl = n_l;
    h->ecp_ng = n_grid;
    h->ecp_qmax = q_max;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_ecp_energy(fire_handle* h, void* stream, uint64_t seed, double* v_loc_dev, double* v_nl_dev,
                    int64_t* n_points) {
  try {
    check_ready(h);
    REQUIRE(h->st.ready, "no walker state (call fire_state_init first)");
    REQUIRE(v_loc_dev && v_nl_dev, "null argument");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, h->st.prec};
    auto& S = h->st;
    const int B = S.B, N = h->N;
    launch_fill_f64(c.s, v_loc_dev, B, 0.0);
    launch_fill_f64(c.s, v_nl_dev, B, 0.0);
    if (n_points) *n_points = 0;
    if (h->ecp_n == 0 || B == 0) return 0;
    state_drop_pending(c);
    const int Q = h->ecp_qmax;
    h->ecp_pts.ensure(sizeof(float) * 3 * (size_t)N * Q * B);
    h->ecp_coef.ensure(sizeof(double) * (size_t)N * Q * B);
    h->ecp_kq.ensure(sizeof(int32_t) * N);
    h->ecp_npts.ensure(sizeof(unsigned long long));
    launch_fill_i32(c.s, h->ecp_kq.as<int32_t>(), N, 0);
    launch_set_i64(c.s, h->ecp_npts.as<int64_t>(), 0);
    EcpTables t{h->ecp_n,          h->ecp_nl,          h->ecp_ng,           Q,
                h->ecp_atom.as<int32_t>(), h->ecp_grid.as<double>(), h->ecp_vloc.as<double>(),
                h->ecp_nl ? h->ecp_vnl.as<double>() : nullptr, h->ecp_cut.as<double>(), h->ecp_nq.as<int32_t>(),
                h->ecp_qpts.as<double>()};
    launch_ecp_points(c.s, S.r.as<float>(), h->R, B, N, t, seed, h->ecp_pts.as<float>(), h->ecp_coef.as<double>(),
                      h->ecp_kq.as<int32_t>(), v_loc_dev, h->ecp_npts.as<unsigned long long>());
    std::vector<int32_t> kq(N);
    int64_t npts = 0;
    HIPCHK(hipMemcpyAsync(kq.data(), h->ecp_kq.p, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipMemcpyAsync(&npts, h->ecp_npts.p, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    if (n_points) *n_points = npts;
    // psi ratios: one low-rank proposal per (electron, quadrature point) for all walkers; a
    // walker whose electron is not active there proposes its own position (coefficient 0)
    for (int k = 0; k < N; ++k)
      for (int q = 0; q < kq[k]; ++q) {
        const size_t slot = ((size_t)k * Q + q) * B;
        state_propose(c, k, 1, h->ecp_pts.as<float>() + slot * 3, 0.f, 0, 0, nullptr);
        S.n_proposals += B;
        launch_ecp_accum(c.s, B, h->ecp_coef.as<double>() + slot, S.p_logpsi.as<double>(), S.p_sign.as<double>(),
                         S.logpsi.as<double>(), S.sign.as<double>(), v_nl_dev);
      }
    state_drop_pending(c);
    state_check_overflow(c);
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_det_options(fire_handle* h, int32_t fp64_laplacian) {
  try {
    REQUIRE(h, "null handle");
    REQUIRE(fp64_laplacian == 0 || fp64_laplacian == 1, "fire_det_options: fp64_laplacian must be 0 or 1");
    h->det_lap64 = fp64_laplacian != 0;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_grad_logpsi(fire_handle* h, void* stream, const float* r_dev, int32_t B, float* grad_dev,
                     double* logpsi_dev) {
  try {
    check_ready(h);
    REQUIRE(B >= 0, "negative batch");
    if (B == 0) return 0;
    REQUIRE(r_dev && grad_dev, "null argument");
    HIPCHK(hipSetDevice(h->device));
    Ctx c{h, (hipStream_t)stream, FIRE_FP32};
    std::vector<int64_t> off(h->expect.size() + 1, 0);
    for (size_t k = 0; k < h->expect.size(); ++k) off[k + 1] = off[k] + h->expect[k];
    const int64_t n_total = off.back();
    const int N = h->N;
    // chunks bounded by the e-n edge rows (each carries about 2 F + 350 floats of workspace)
    const int64_t per_walker = (int64_t)N * h->M * (2 * h->F + 360) * 4;
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(128, (int64_t)24000000000LL / per_walker));
    for (int w0 = 0; w0 < B; w0 += chunk) {
      const int Bc = std::min(chunk, B - w0);
      run_grad(c, r_dev + (size_t)w0 * N * 3, Bc, grad_dev + (size_t)w0 * n_total, n_total, off,
               logpsi_dev ? logpsi_dev + w0 : nullptr);
    }
    HIPCHK(hipStreamSynchronize(c.s));
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_param_count(fire_handle* h, int64_t* n_floats) {
  try {
    REQUIRE(h && n_floats, "null argument");
    int64_t n = 0;
    for (int64_t v : h->expect) n += v;
    *n_floats = n;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int fire_energy_stats(fire_handle* h, void* stream, const double* e_loc_dev, int32_t B, double clip_local_energy,
                      int32_t stat, double* e_clipped_dev, double* e_diff_dev, double* stats_dev) {
  try {
    REQUIRE(e_loc_dev, "null argument");
    REQUIRE(B > 0, "fire_energy_stats: empty batch");
    REQUIRE(stat == 0 || stat == 1, "fire_energy_stats: stat must be 0 (median) or 1 (mean)");
    if (h) HIPCHK(hipSetDevice(h->device));
    fire::launch_energy_stats((hipStream_t)stream, e_loc_dev, B, clip_local_energy, stat, e_clipped_dev, e_diff_dev,
                              stats_dev);
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

}  // extern "C"
