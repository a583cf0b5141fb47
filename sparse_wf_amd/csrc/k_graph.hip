// Neighbour lists (CSR) for a batch of walkers.
//
// Follows get_pair_indices / is_neighbour (src/sparse_wf/model/new_sparse_model.py:238-263,
// src/sparse_wf/model/graph_utils.py:179-184): distance in float32 as
// sqrt((dx*dx + dy*dy) + dz*dz) with every operation rounded separately (no FMA
// contraction), diagonal excluded, strict d < cutoff, neighbours in ascending order.
//
// One 64-lane wavefront per centre electron; a ballot over 64 candidate neighbours at a
// time gives the count (pass 1) and each lane's output slot via a masked popcount (pass 2),
// so the order is deterministic.  Groups (centre electrons) are numbered up-spin first,
// walker-major (fire::group_of).
#include "fire_common.hpp"
#include "fire_kernels.hpp"

namespace fire {

__device__ __forceinline__ bool in_cutoff(const float* ri, const float* rj, float cutoff) {
  const float dx = __fsub_rn(ri[0], rj[0]);
  const float dy = __fsub_rn(ri[1], rj[1]);
  const float dz = __fsub_rn(ri[2], rj[2]);
  const float s = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
  return __fsqrt_rn(s) < cutoff;
}

// `list` (optional): centres as (walker, electron) pairs, output index = list position;
// otherwise every electron of every walker, output index in group order, or in walker-major
// order (w * N + i) with `walker_major`.
template <bool FILL>
__global__ void __launch_bounds__(256) k_nbr(const float* __restrict__ r, int B, int N, int n_up, float cutoff,
                                             int32_t* __restrict__ cnt, const int64_t* __restrict__ ptr,
                                             int32_t* __restrict__ nbr, const int2* __restrict__ list, int n_list,
                                             const int32_t* __restrict__ dn, bool walker_major) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int n_centres = list ? n_list : B * N;
  if (wave >= n_centres) return;
  if (list && dn && wave >= *dn) {  // beyond the device-side list length: empty row
    if (!FILL && lane == 0) cnt[wave] = 0;
    return;
  }
  int w, i, g;
  if (list) {
    w = list[wave].x;
    i = list[wave].y;
    g = wave;
  } else {
    w = wave / N;
    i = wave - w * N;
    g = walker_major ? wave : group_of(w, i, B, n_up, N - n_up);
  }
  const float* rw = r + (size_t)w * N * 3;
  const float ri[3] = {rw[3 * i], rw[3 * i + 1], rw[3 * i + 2]};
  int count = 0;
  int64_t base = FILL ? ptr[g] : 0;
  for (int j0 = 0; j0 < N; j0 += 64) {
    const int j = j0 + lane;
    bool in = false;
    if (j < N && j != i) in = in_cutoff(ri, rw + 3 * j, cutoff);
    const uint64_t m = __ballot(in);
    if (FILL && in) {
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      nbr[base + count + before] = j;
    }
    count += __popcll(m);
  }
  if (!FILL && lane == 0) cnt[g] = count;
}

// Exclusive prefix sum of the neighbour counts: ptr[0..G) offsets, ptr[G] total,
// ptr[G+1] largest count.  One block walks the counts in tiles of 4096 (4 consecutive per
// thread, coalesced): wave prefix sums by shuffles, one LDS pass over the 16 wave totals.
// With dn (device count) only the first *dn counts are read; later entries are zero by
// construction and get ptr = total.
__global__ void __launch_bounds__(1024) k_scan(const int32_t* __restrict__ cnt, int G, int64_t* __restrict__ ptr,
                                               const int32_t* __restrict__ dn) {
  __shared__ long s_wsum[16];
  __shared__ int32_t s_wmax[16];
  const int n = dn ? min(G, *dn) : G;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  long carry = 0;
  int32_t mx_all = 0;
  for (int base = 0; base < n; base += 4096) {
    const int k0 = base + 4 * t;
    int32_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = k0 + u < n ? cnt[k0 + u] : 0;
    const long sum = (long)c[0] + c[1] + c[2] + c[3];
    int32_t m = max(max(c[0], c[1]), max(c[2], c[3]));
    long incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const long y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if (lane == 63) s_wsum[wv] = incl;
    if (lane == 0) s_wmax[wv] = m;
    __syncthreads();
    long before = 0, tot = 0;
    int32_t mx = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const long v = s_wsum[q];
      before += q < wv ? v : 0;
      tot += v;
      mx = max(mx, s_wmax[q]);
    }
    long run = carry + before + incl - sum;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k0 + u < n) {
        ptr[k0 + u] = run;
        run += c[u];
      }
    carry += tot;
    mx_all = max(mx_all, mx);
    __syncthreads();  // s_wsum is rewritten by the next tile
  }
  for (int k = n + t; k < G; k += 1024) ptr[k] = carry;
  if (t == 0) {
    ptr[G] = carry;
    ptr[G + 1] = mx_all;
  }
}

// rev[p] for pair p = (i -> j) at slot k of i's list: the slot of i in j's list
// (binary search in j's ascending list).
__global__ void k_rev(const int64_t* __restrict__ ptr, const int32_t* __restrict__ nbr, int B, int N, int n_up,
                      int32_t* __restrict__ rev) {
  const int g = blockIdx.x;
  int w, i;
  group_inv(g, B, n_up, N - n_up, w, i);
  const int64_t p0 = ptr[g], p1 = ptr[g + 1];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int j = nbr[p];
    const int gj = group_of(w, j, B, n_up, N - n_up);
    int64_t a = ptr[gj], b = ptr[gj + 1];
    while (a < b) {
      int64_t mid = (a + b) >> 1;
      if (nbr[mid] < i) a = mid + 1; else b = mid;
    }
    rev[p] = (int32_t)(a - ptr[gj]);
  }
}

void launch_nbr_count(hipStream_t s, const float* r, int B, int N, int n_up, float cutoff, int32_t* cnt,
                      const int2* list, int n_list, const int32_t* dn, bool walker_major) {
  const int waves = list ? n_list : B * N;
  if (waves == 0) return;
  const int blocks = (waves * 64 + 255) / 256;
  k_nbr<false><<<blocks, 256, 0, s>>>(r, B, N, n_up, cutoff, cnt, nullptr, nullptr, list, n_list, dn, walker_major);
}

void launch_nbr_fill(hipStream_t s, const float* r, int B, int N, int n_up, float cutoff, const int64_t* ptr,
                     int32_t* nbr, const int2* list, int n_list, const int32_t* dn, bool walker_major) {
  const int waves = list ? n_list : B * N;
  if (waves == 0) return;
  const int blocks = (waves * 64 + 255) / 256;
  k_nbr<true><<<blocks, 256, 0, s>>>(r, B, N, n_up, cutoff, nullptr, ptr, nbr, list, n_list, dn, walker_major);
}

void launch_scan(hipStream_t s, const int32_t* cnt, int G, int64_t* ptr, const int32_t* dn) {
  k_scan<<<1, 1024, 0, s>>>(cnt, G, ptr, dn);
}

void launch_rev(hipStream_t s, const int64_t* ptr, const int32_t* nbr, int B, int N, int n_up, int32_t* rev) {
  k_rev<<<B * N, 64, 0, s>>>(ptr, nbr, B, N, n_up, rev);
}

}  // namespace fire
