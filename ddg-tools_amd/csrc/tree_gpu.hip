// BVH construction on the device (see tree_gpu.hpp).
#include <cstring>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>

#include "tree_gpu.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

namespace {

constexpr uint32_t kSplitChunk = 2048;  // elements per workgroup in the chunked split

// Total order key of a double (the order of f64::total_cmp), as an unsigned integer.
__device__ __forceinline__ uint64_t total_key_u(double x) {
  const int64_t b = __double_as_longlong(x);
  const int64_t k = b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
  return (uint64_t)k ^ 0x8000000000000000ull;
}

struct Box6 {
  double l0, l1, l2, h0, h1, h2;
};

__device__ __forceinline__ Box6 empty6() { return {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY}; }

__device__ __forceinline__ Box6 unite(const Box6& a, const Box6& b) {
  return {fmin(a.l0, b.l0), fmin(a.l1, b.l1), fmin(a.l2, b.l2), fmax(a.h0, b.h0), fmax(a.h1, b.h1), fmax(a.h2, b.h2)};
}

__device__ __forceinline__ Box6 load_box(const double* lo, const double* hi, uint32_t e) {
  return {lo[3 * e], lo[3 * e + 1], lo[3 * e + 2], hi[3 * e], hi[3 * e + 1], hi[3 * e + 2]};
}

__device__ __forceinline__ Box6 shfl_up6(const Box6& b, int o) {
  return {__shfl_up(b.l0, o), __shfl_up(b.l1, o), __shfl_up(b.l2, o),
          __shfl_up(b.h0, o), __shfl_up(b.h1, o), __shfl_up(b.h2, o)};
}

// Σ (hi − lo)² over the axes in order, 0 for an empty box (the host build's diam2).
__device__ __forceinline__ double diam2(const Box6& b) {
  if (b.l0 > b.h0 || b.l1 > b.h1 || b.l2 > b.h2) return 0.0;
  double s = 0.0, w;
  w = b.h0 - b.l0;
  s = s + w * w;
  w = b.h1 - b.l1;
  s = s + w * w;
  w = b.h2 - b.l2;
  s = s + w * w;
  return s;
}

// Inclusive union scan over a 256-thread workgroup; *total gets the union of all 256.
__device__ Box6 block_scan(Box6 v, Box6* wave_tot, Box6* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const Box6 t = shfl_up6(v, o);
    if (lane >= o) v = unite(t, v);
  }
  if (lane == 63) wave_tot[wave] = v;
  __syncthreads();
  for (int w = 0; w < wave; ++w) v = unite(wave_tot[w], v);
  Box6 all = wave_tot[0];
  for (int w = 1; w < 4; ++w) all = unite(all, wave_tot[w]);
  *total = all;
  __syncthreads();
  return v;
}

}  // namespace

__global__ __launch_bounds__(256) void bvh_keys_kernel(uint32_t n, const double* __restrict__ pts, int axis,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ iota) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  keys[e] = total_key_u(pts[3 * (size_t)e + axis]);
  iota[e] = e;
}

__global__ __launch_bounds__(256) void bvh_iota_kernel(uint32_t n, uint32_t* __restrict__ a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ __launch_bounds__(256) void bvh_fill_kernel(uint32_t n, uint32_t* __restrict__ a, uint32_t v) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// seg_of[element] = index of the splitting node holding it (one workgroup per node).
__global__ __launch_bounds__(256) void bvh_mark_kernel(const uint32_t* __restrict__ act_start,
                                                       const uint32_t* __restrict__ act_end,
                                                       const uint32_t* __restrict__ order, uint32_t* __restrict__ seg_of) {
  const uint32_t a = blockIdx.x;
  for (uint32_t p = act_start[a] + threadIdx.x; p < act_end[a]; p += 256) seg_of[order[p]] = a;
}

__global__ __launch_bounds__(256) void bvh_part_keys_kernel(uint32_t n, const uint32_t* __restrict__ axis_order,
                                                            const uint32_t* __restrict__ seg_of,
                                                            uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t e = axis_order[i];
  key[i] = seg_of[e];
  val[i] = e;
}

// The i-th element of the partitioned sequence goes to its node's range, in axis order.
__global__ __launch_bounds__(256) void bvh_scatter_kernel(uint32_t total, const uint32_t* __restrict__ key,
                                                          const uint32_t* __restrict__ val,
                                                          const uint32_t* __restrict__ act_start,
                                                          const uint32_t* __restrict__ act_prefix,
                                                          uint32_t* __restrict__ order) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const uint32_t a = key[i];
  order[act_start[a] + (i - act_prefix[a])] = val[i];
}

// Level records of the device-driven build (BvhLevel: a level's nodes are records
// [off, off + na), `total` elements, `nch` split chunks).
struct BvhLevel {
  uint32_t off, na, total, nch;
};

// The split of one node [start, start + cnt) by one workgroup: suffix boxes, then prefix boxes
// with the cost of every admissible split, reduced to the first minimum (thread 0's value).
__device__ uint32_t split_node(uint32_t start, uint32_t cnt, const uint32_t* __restrict__ order,
                               const double* __restrict__ lo, const double* __restrict__ hi, double* __restrict__ suf) {
  __shared__ Box6 wave_tot[4];
  __shared__ double red_c[256];
  __shared__ uint32_t red_k[256];
  const uint32_t t = threadIdx.x;
  const uint32_t* sl = order + start;
  double* sf = suf + 6 * (size_t)start;
  size_t low = (size_t)ceil(0.25 * (double)cnt);
  if (low < 1) low = 1;
  if (low > cnt / 2) low = cnt / 2;
  // suffix boxes: chunks from the end, lane order reversed
  Box6 carry = empty6();
  for (int64_t c1 = cnt; c1 > 0; c1 -= 256) {
    const int64_t c0 = c1 - 256 > 0 ? c1 - 256 : 0;
    const int64_t p = c1 - 1 - (int64_t)t;
    Box6 v = p >= c0 ? load_box(lo, hi, sl[p]) : empty6();
    Box6 tot;
    v = block_scan(v, wave_tot, &tot);
    v = unite(v, carry);
    if (p >= c0) {
      double* o = sf + 6 * p;
      o[0] = v.l0;
      o[1] = v.l1;
      o[2] = v.l2;
      o[3] = v.h0;
      o[4] = v.h1;
      o[5] = v.h2;
    }
    carry = unite(tot, carry);
  }
  __syncthreads();
  // prefix boxes and costs: the split k = p + 1 puts [0, k) left and [k, cnt) right
  double best = INFINITY;
  uint32_t best_k = UINT32_MAX;
  carry = empty6();
  for (uint32_t c0 = 0; c0 < cnt; c0 += 256) {
    const uint32_t p = c0 + t;
    Box6 v = p < cnt ? load_box(lo, hi, sl[p]) : empty6();
    Box6 tot;
    v = block_scan(v, wave_tot, &tot);
    v = unite(carry, v);
    const size_t k = (size_t)p + 1;
    if (p < cnt && k >= low && k <= cnt - low) {
      const double* o = sf + 6 * (size_t)k;
      const Box6 r = {o[0], o[1], o[2], o[3], o[4], o[5]};
      const double cost = diam2(v) + diam2(r);
      if (cost < best) {
        best = cost;
        best_k = (uint32_t)k;
      }
    }
    carry = unite(carry, tot);
  }
  red_c[t] = best;
  red_k[t] = best_k;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)t < w) {
      const double c2 = red_c[t + w];
      const uint32_t k2 = red_k[t + w];
      if (k2 != UINT32_MAX && (red_k[t] == UINT32_MAX || c2 < red_c[t] || (c2 == red_c[t] && k2 < red_k[t]))) {
        red_c[t] = c2;
        red_k[t] = k2;
      }
    }
    __syncthreads();
  }
  return red_k[0] == UINT32_MAX ? cnt / 2 : red_k[0];
}

// One workgroup per splitting node.
__global__ __launch_bounds__(256) void bvh_split_kernel(const uint32_t* __restrict__ act_start,
                                                        const uint32_t* __restrict__ act_end,
                                                        const uint32_t* __restrict__ order, const double* __restrict__ lo,
                                                        const double* __restrict__ hi, double* __restrict__ suf,
                                                        uint32_t* __restrict__ split) {
  const uint32_t a = blockIdx.x;
  const uint32_t k = split_node(act_start[a], act_end[a] - act_start[a], order, lo, hi, suf);
  if (threadIdx.x == 0) split[a] = k;
}

// ---- chunked split for large nodes: several workgroups per node ------------------------------
// A node's elements are cut into chunks of kSplitChunk; the chunk totals are scanned per node, then
// every chunk writes its suffix boxes (one launch) and evaluates its splits (the next launch: a
// split's right box may lie in the next chunk), and the chunks' first minima are reduced per node.
// Box unions are exact, so the boxes, costs and the chosen split equal the one-workgroup kernel's.

// The chunked kernels serve both builds: `lv` null (host-driven: act_* are the level's nodes,
// the grid is the chunk count) or the device-driven level (act_* the flat records from lv->off,
// the grid an upper bound, lv->nch the chunk count).
__global__ __launch_bounds__(256) void bvh_chunk_total_kernel(const BvhLevel* __restrict__ lv,
                                                              const uint32_t* __restrict__ ch_seg,
                                                              const uint32_t* __restrict__ ch_c0,
                                                              const uint32_t* __restrict__ act_start,
                                                              const uint32_t* __restrict__ act_end,
                                                              const uint32_t* __restrict__ order,
                                                              const double* __restrict__ lo,
                                                              const double* __restrict__ hi, Box6* __restrict__ ctot) {
  __shared__ Box6 wave_tot[4];
  const uint32_t c = blockIdx.x;
  uint32_t off = 0;
  if (lv) {
    if (c >= lv->nch) return;
    off = lv->off;
  }
  const uint32_t a = off + ch_seg[c];
  const uint32_t start = act_start[a], cnt = act_end[a] - start;
  const uint32_t c0 = ch_c0[c], c1 = min(c0 + kSplitChunk, cnt);
  Box6 v = empty6();
  for (uint32_t p = c0 + threadIdx.x; p < c1; p += 256) v = unite(v, load_box(lo, hi, order[start + p]));
  Box6 tot;
  (void)block_scan(v, wave_tot, &tot);
  if (threadIdx.x == 0) ctot[c] = tot;
}

// Per chunk: the union of the node's earlier chunks (cpre) and of its later chunks (csuf).
__global__ __launch_bounds__(256) void bvh_chunk_scan_kernel(const BvhLevel* __restrict__ lv, uint32_t nch,
                                                             const uint32_t* __restrict__ ch_seg,
                                                             const uint32_t* __restrict__ seg_chunk0,
                                                             const Box6* __restrict__ ctot, Box6* __restrict__ cpre,
                                                             Box6* __restrict__ csuf) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (lv) nch = lv->nch;
  if (c >= nch) return;
  const uint32_t a = ch_seg[c], f0 = seg_chunk0[a], f1 = seg_chunk0[a + 1];
  Box6 pre = empty6(), suf = empty6();
  for (uint32_t j = f0; j < c; ++j) pre = unite(pre, ctot[j]);
  for (uint32_t j = c + 1; j < f1; ++j) suf = unite(suf, ctot[j]);
  cpre[c] = pre;
  csuf[c] = suf;
}

__global__ __launch_bounds__(256) void bvh_chunk_suffix_kernel(const BvhLevel* __restrict__ lv,
                                                               const uint32_t* __restrict__ ch_seg,
                                                               const uint32_t* __restrict__ ch_c0,
                                                               const uint32_t* __restrict__ act_start,
                                                               const uint32_t* __restrict__ act_end,
                                                               const uint32_t* __restrict__ order,
                                                               const double* __restrict__ lo,
                                                               const double* __restrict__ hi,
                                                               const Box6* __restrict__ csuf, double* __restrict__ suf) {
  __shared__ Box6 wave_tot[4];
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  uint32_t off = 0;
  if (lv) {
    if (c >= lv->nch) return;
    off = lv->off;
  }
  const uint32_t a = off + ch_seg[c];
  const uint32_t start = act_start[a], cnt = act_end[a] - start;
  const uint32_t c0 = ch_c0[c], c1 = min(c0 + kSplitChunk, cnt);
  const uint32_t* sl = order + start;
  double* sf = suf + 6 * (size_t)start;
  Box6 carry = csuf[c];
  for (int64_t e1 = c1; e1 > (int64_t)c0; e1 -= 256) {
    const int64_t e0 = e1 - 256 > (int64_t)c0 ? e1 - 256 : (int64_t)c0;
    const int64_t p = e1 - 1 - (int64_t)t;
    Box6 v = p >= e0 ? load_box(lo, hi, sl[p]) : empty6();
    Box6 tot;
    v = block_scan(v, wave_tot, &tot);
    v = unite(v, carry);
    if (p >= e0) {
      double* o = sf + 6 * p;
      o[0] = v.l0;
      o[1] = v.l1;
      o[2] = v.l2;
      o[3] = v.h0;
      o[4] = v.h1;
      o[5] = v.h2;
    }
    carry = unite(tot, carry);
  }
}

struct SplitBest {
  double cost;
  uint32_t k;
};

__global__ __launch_bounds__(256) void bvh_chunk_cost_kernel(const BvhLevel* __restrict__ lv,
                                                             const uint32_t* __restrict__ ch_seg,
                                                             const uint32_t* __restrict__ ch_c0,
                                                             const uint32_t* __restrict__ act_start,
                                                             const uint32_t* __restrict__ act_end,
                                                             const uint32_t* __restrict__ order,
                                                             const double* __restrict__ lo,
                                                             const double* __restrict__ hi,
                                                             const Box6* __restrict__ cpre,
                                                             const double* __restrict__ suf,
                                                             SplitBest* __restrict__ cbest) {
  __shared__ Box6 wave_tot[4];
  __shared__ double red_c[256];
  __shared__ uint32_t red_k[256];
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  uint32_t off = 0;
  if (lv) {
    if (c >= lv->nch) return;
    off = lv->off;
  }
  const uint32_t a = off + ch_seg[c];
  const uint32_t start = act_start[a], cnt = act_end[a] - start;
  const uint32_t c0 = ch_c0[c], c1 = min(c0 + kSplitChunk, cnt);
  const uint32_t* sl = order + start;
  const double* sf = suf + 6 * (size_t)start;
  size_t low = (size_t)ceil(0.25 * (double)cnt);
  if (low < 1) low = 1;
  if (low > cnt / 2) low = cnt / 2;
  double best = INFINITY;
  uint32_t best_k = UINT32_MAX;
  Box6 carry = cpre[c];
  for (uint32_t e0 = c0; e0 < c1; e0 += 256) {
    const uint32_t p = e0 + t;
    Box6 v = p < c1 ? load_box(lo, hi, sl[p]) : empty6();
    Box6 tot;
    v = block_scan(v, wave_tot, &tot);
    v = unite(carry, v);
    const size_t k = (size_t)p + 1;
    if (p < c1 && k >= low && k <= cnt - low) {
      const double* o = sf + 6 * k;
      const Box6 r = {o[0], o[1], o[2], o[3], o[4], o[5]};
      const double cost = diam2(v) + diam2(r);
      if (cost < best) {
        best = cost;
        best_k = (uint32_t)k;
      }
    }
    carry = unite(carry, tot);
  }
  red_c[t] = best;
  red_k[t] = best_k;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)t < w) {
      const double c2 = red_c[t + w];
      const uint32_t k2 = red_k[t + w];
      if (k2 != UINT32_MAX && (red_k[t] == UINT32_MAX || c2 < red_c[t] || (c2 == red_c[t] && k2 < red_k[t]))) {
        red_c[t] = c2;
        red_k[t] = k2;
      }
    }
    __syncthreads();
  }
  if (t == 0) cbest[c] = {red_c[0], red_k[0]};
}

// Per node: the first minimum over its chunks (chunks in increasing k).
__global__ __launch_bounds__(256) void bvh_chunk_pick_kernel(const BvhLevel* __restrict__ lv, uint32_t na,
                                                             const uint32_t* __restrict__ seg_chunk0,
                                                             const uint32_t* __restrict__ act_start,
                                                             const uint32_t* __restrict__ act_end,
                                                             const SplitBest* __restrict__ cbest,
                                                             uint32_t* __restrict__ split) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t off = 0;
  if (lv) {
    na = lv->na;
    off = lv->off;
  }
  if (i >= na) return;
  const uint32_t a = i;
  double best = INFINITY;
  uint32_t best_k = UINT32_MAX;
  for (uint32_t j = seg_chunk0[a]; j < seg_chunk0[a + 1]; ++j) {
    const SplitBest b = cbest[j];
    if (b.k != UINT32_MAX && (best_k == UINT32_MAX || b.cost < best)) {
      best = b.cost;
      best_k = b.k;
    }
  }
  const uint32_t cnt = act_end[off + a] - act_start[off + a];
  split[off + a] = best_k == UINT32_MAX ? cnt / 2 : best_k;
}

// ---- device-driven levels ----------------------------------------------------------------------
// The host enqueues levels without reading anything back: a level's nodes are records
// [off, off + na) of flat arrays (first element, end, element prefix, split), written by the
// previous level's bvh_next kernel; every kernel reads its level's counts on the device and is
// launched with an upper bound.  The host reads the splits once at the end and creates and
// numbers the nodes exactly as the host-driven build does.

struct BvhRecords {
  uint32_t *start, *end, *prefix, *split;
};

// seg_of[order[p]] = the level's node holding position p, or the sentinel (one thread per
// position: the nodes are disjoint ranges in increasing order).
__global__ __launch_bounds__(256) void bvh_mark_all_kernel(uint32_t n, const BvhLevel* __restrict__ lv, BvhRecords R,
                                                           const uint32_t* __restrict__ order,
                                                           uint32_t* __restrict__ seg_of) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const uint32_t off = lv->off, na = lv->na;
  const uint32_t* st = R.start + off;
  uint32_t l = 0, h = na;  // the number of nodes starting at or before p
  while (l < h) {
    const uint32_t m = (l + h) >> 1;
    if (st[m] <= p)
      l = m + 1;
    else
      h = m;
  }
  seg_of[order[p]] = (l > 0 && p < R.end[off + l - 1]) ? l - 1 : UINT32_MAX;
}

__global__ __launch_bounds__(256) void bvh_scatter_lv_kernel(const BvhLevel* __restrict__ lv, BvhRecords R,
                                                             const uint32_t* __restrict__ key,
                                                             const uint32_t* __restrict__ val,
                                                             uint32_t* __restrict__ order) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= lv->total) return;
  const uint32_t a = lv->off + key[i];
  order[R.start[a] + (i - R.prefix[a])] = val[i];
}

__global__ __launch_bounds__(256) void bvh_split_lv_kernel(const BvhLevel* __restrict__ lv, BvhRecords R,
                                                           const uint32_t* __restrict__ order,
                                                           const double* __restrict__ lo, const double* __restrict__ hi,
                                                           double* __restrict__ suf) {
  if (blockIdx.x >= lv->na) return;  // uniform per workgroup
  const uint32_t a = lv->off + blockIdx.x;
  const uint32_t k = split_node(R.start[a], R.end[a] - R.start[a], order, lo, hi, suf);
  if (threadIdx.x == 0) R.split[a] = k;
}

// Inclusive sums over a 256-thread workgroup of two counters; `tot` gets the workgroup's totals.
__device__ void block_scan2(uint32_t& x, uint32_t& y, uint32_t* wx, uint32_t* wy, uint32_t& tx, uint32_t& ty) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(x, o), v = __shfl_up(y, o);
    if (lane >= o) {
      x += u;
      y += v;
    }
  }
  if (lane == 63) {
    wx[wave] = x;
    wy[wave] = y;
  }
  __syncthreads();
  uint32_t bx = 0, by = 0;
  for (int w = 0; w < wave; ++w) {
    bx += wx[w];
    by += wy[w];
  }
  x += bx;
  y += by;
  tx = wx[0] + wx[1] + wx[2] + wx[3];
  ty = wy[0] + wy[1] + wy[2] + wy[3];
  __syncthreads();
}

// Inclusive sum over a workgroup of W waves; tot gets the workgroup's total.
template <int W>
__device__ void block_scan_n(uint32_t& x, uint32_t* wx, uint32_t& tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(x, o);
    if (lane >= o) x += u;
  }
  if (lane == 63) wx[wave] = x;
  __syncthreads();
  uint32_t b = 0, all = 0;
  for (int w = 0; w < W; ++w) {
    if (w < wave) b += wx[w];
    all += wx[w];
  }
  x += b;
  tot = all;
  __syncthreads();
}

// The level's split chunks (⌈cnt / kSplitChunk⌉ per node, in node order): per chunk its node and
// first element, per node its first chunk (one workgroup).
__global__ __launch_bounds__(256) void bvh_chunks_lv_kernel(BvhLevel* __restrict__ lv, BvhRecords R,
                                                            uint32_t* __restrict__ ch_seg, uint32_t* __restrict__ ch_c0,
                                                            uint32_t* __restrict__ seg_chunk0) {
  __shared__ uint32_t wx[4], wy[4];
  const uint32_t off = lv->off, na = lv->na;
  uint32_t base = 0;
  for (uint32_t a0 = 0; a0 < na; a0 += 256) {
    const uint32_t a = a0 + threadIdx.x;
    const uint32_t cnt = a < na ? R.end[off + a] - R.start[off + a] : 0;
    const uint32_t c = (cnt + kSplitChunk - 1) / kSplitChunk;
    uint32_t x = c, y = 0, tx, ty;
    block_scan2(x, y, wx, wy, tx, ty);
    if (a < na) {
      const uint32_t first = base + x - c;
      seg_chunk0[a] = first;
      for (uint32_t j = 0; j < c; ++j) {
        ch_seg[first + j] = a;
        ch_c0[first + j] = j * kSplitChunk;
      }
    }
    base += tx;
  }
  if (threadIdx.x == 0) {
    seg_chunk0[na] = base;
    lv->nch = base;
  }
}

// The next level: the children of this level's nodes that split (more than leaf_size elements),
// in order (left before right), with their element prefix (one workgroup).
__global__ __launch_bounds__(256) void bvh_next_kernel(BvhLevel* __restrict__ lv, BvhRecords R, uint32_t leaf_size) {
  __shared__ uint32_t wx[4], wy[4];
  const uint32_t off = lv->off, na = lv->na, off2 = off + na;
  uint32_t base_n = 0, base_e = 0;
  for (uint32_t a0 = 0; a0 < na; a0 += 256) {
    const uint32_t a = a0 + threadIdx.x;
    uint32_t s0 = 0, m = 0, e0 = 0;
    bool la = false, ra = false;
    if (a < na) {
      s0 = R.start[off + a];
      e0 = R.end[off + a];
      m = s0 + R.split[off + a];
      la = m - s0 > leaf_size;
      ra = e0 - m > leaf_size;
    }
    const uint32_t cn = (uint32_t)la + (uint32_t)ra, ce = (la ? m - s0 : 0) + (ra ? e0 - m : 0);
    uint32_t x = cn, y = ce, tx, ty;
    block_scan2(x, y, wx, wy, tx, ty);
    uint32_t q = off2 + base_n + x - cn, pe = base_e + y - ce;
    if (la) {
      R.start[q] = s0;
      R.end[q] = m;
      R.prefix[q] = pe;
      ++q;
      pe += m - s0;
    }
    if (ra) {
      R.start[q] = m;
      R.end[q] = e0;
      R.prefix[q] = pe;
    }
    base_n += tx;
    base_e += ty;
  }
  if (threadIdx.x == 0) lv[1] = BvhLevel{off2, base_n, base_e, 0};
}

__global__ void bvh_root_kernel(uint32_t n, BvhLevel* __restrict__ lv, BvhRecords R) {
  lv[0] = BvhLevel{0, 1, n, 0};
  R.start[0] = 0;
  R.end[0] = n;
  R.prefix[0] = 0;
}

// ---- subtrees finished in one workgroup ---------------------------------------------------------
// Below a cut depth every node has at most kSubMax elements; one workgroup builds each such node's
// whole subtree in LDS.  The elements get local ids (their rank in the node's axis-0 order) and the
// workgroup keeps, per axis, the node-ordered element sequence in that axis's (key, index) order.
// Level by level, each wave splits one node at a time: the node's elements in its axis order are
// its range of that axis's sequence, the split is the first minimum of the same costs as
// split_node's (chunk unions, then per-chunk prefix and suffix scans; unions are exact), and every
// axis sequence is then stably partitioned into the children's ranges (ballot ranks), which keeps
// each child's elements in every axis order.  A child of at most leaf_size elements writes its
// elements to the final order in its parent's axis order, as the level-by-level build leaves them.
// The splits go out in level order per subtree; the host numbers the nodes as before.
constexpr uint32_t kSubMax = 2048;
constexpr uint32_t kSubNodes = kSubMax / 2;  // nodes of one local level (each holds 2 or more elements)

struct SubtreeArgs {
  const BvhLevel* lv;     // the cut level's records
  BvhRecords R;
  const uint32_t* lists;  // 3 n: per axis, each cut node's elements in that axis's order
  uint32_t* inv;          // n: element → local id (scratch)
  const double *lo, *hi;
  uint32_t* order;        // n: the leaves' elements are written here
  uint32_t* sub_split;    // n: per cut node, from its first position, its subtree's splits in level order
  uint32_t* sub_count;    // per cut node: splits written, UINT32_MAX when it has more than kSubMax elements
  uint32_t n, leaf, axis0;
};

namespace {
__device__ __forceinline__ Box6 shfl_down6(const Box6& b, int o) {
  return {__shfl_down(b.l0, o), __shfl_down(b.l1, o), __shfl_down(b.l2, o),
          __shfl_down(b.h0, o), __shfl_down(b.h1, o), __shfl_down(b.h2, o)};
}

// lane `src` (wave-uniform) of every component, by scalar reads of the lane
__device__ __forceinline__ double readlane_d(double x, int src) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, src);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ Box6 shfl6(const Box6& b, int src) {
  return {readlane_d(b.l0, src), readlane_d(b.l1, src), readlane_d(b.l2, src),
          readlane_d(b.h0, src), readlane_d(b.h1, src), readlane_d(b.h2, src)};
}

// inclusive unions over lanes 0..lane (fwd) or lane..63 (bwd)
__device__ __forceinline__ Box6 wave_union_fwd(Box6 v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const Box6 t = shfl_up6(v, o);
    if (lane >= o) v = unite(t, v);
  }
  return v;
}

__device__ __forceinline__ Box6 wave_union_bwd(Box6 v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const Box6 t = shfl_down6(v, o);
    if (lane + o < 64) v = unite(v, t);
  }
  return v;
}

// The split of the node at local positions [s, e) along the axis whose sequence is L (one wave).
__device__ uint32_t wave_split(uint32_t s, uint32_t e, const uint16_t* __restrict__ L, const uint32_t* __restrict__ gid,
                               const double* __restrict__ lo, const double* __restrict__ hi, int lane) {
  const uint32_t cnt = e - s, nch = (cnt + 63) / 64;
  size_t low = (size_t)ceil(0.25 * (double)cnt);
  if (low < 1) low = 1;
  if (low > cnt / 2) low = cnt / 2;
  // lane c: the unions of the chunks before c and after c (none for a node of one chunk)
  Box6 pre = empty6(), post = empty6();
  if (nch > 1) {
    Box6 mine = empty6();
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t p = s + c * 64 + lane;
      Box6 v = p < e ? load_box(lo, hi, gid[L[p]]) : empty6();
      v = shfl6(wave_union_fwd(v, lane), 63);
      if ((uint32_t)lane == c) mine = v;
    }
    pre = shfl_up6(wave_union_fwd(mine, lane), 1);
    if (lane == 0) pre = empty6();
    post = shfl_down6(wave_union_bwd(mine, lane), 1);
    if (lane == 63) post = empty6();
  }
  double best = INFINITY;
  uint32_t best_k = UINT32_MAX;
  for (uint32_t c = 0; c < nch; ++c) {
    const Box6 pre_c = shfl6(pre, c), post_c = shfl6(post, c);
    const uint32_t p = s + c * 64 + lane;
    const Box6 v = p < e ? load_box(lo, hi, gid[L[p]]) : empty6();
    const Box6 left = unite(pre_c, wave_union_fwd(v, lane));  // positions s..p
    Box6 rest = shfl_down6(wave_union_bwd(v, lane), 1);        // positions p+1.. to the chunk's end
    if (lane == 63) rest = empty6();
    const Box6 right = unite(rest, post_c);
    const size_t k = (size_t)(p - s) + 1;
    if (p < e && k >= low && k <= cnt - low) {
      const double cost = diam2(left) + diam2(right);
      if (cost < best) {
        best = cost;
        best_k = (uint32_t)k;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double c2 = __shfl_xor(best, o);
    const uint32_t k2 = __shfl_xor(best_k, o);
    if (k2 != UINT32_MAX && (best_k == UINT32_MAX || c2 < best || (c2 == best && k2 < best_k))) {
      best = c2;
      best_k = k2;
    }
  }
  return best_k == UINT32_MAX ? cnt / 2 : best_k;
}
}  // namespace

constexpr uint32_t kSubThreads = 512;

__global__ __launch_bounds__(512) void bvh_subtree_kernel(SubtreeArgs A) {
  __shared__ uint32_t gid[kSubMax];
  __shared__ uint16_t seq[2][3][kSubMax];  // [level parity][axis][position] → local id
  __shared__ uint8_t flag[kSubMax];        // local id → in the left child
  __shared__ uint32_t nodes[2][kSubNodes]; // per local level: s | e << 16
  __shared__ uint16_t ksplit[kSubNodes];
  __shared__ uint32_t count[2];
  __shared__ uint32_t wx[kSubThreads / 64], wy[kSubThreads / 64];
  const uint32_t g = blockIdx.x, t = threadIdx.x;
  const int lane = (int)(t & 63), wave = (int)(t >> 6);
  if (g >= A.lv->na) return;  // uniform per workgroup
  const uint32_t a = A.lv->off + g, s0 = A.R.start[a], m = A.R.end[a] - s0;
  if (m > kSubMax) {
    if (t == 0) A.sub_count[g] = UINT32_MAX;
    return;
  }
  for (uint32_t i = t; i < m; i += kSubThreads) {
    const uint32_t x = A.lists[s0 + i];
    gid[i] = x;
    A.inv[x] = i;
    seq[0][0][i] = (uint16_t)i;
  }
  __threadfence_block();
  __syncthreads();
  for (int b = 1; b < 3; ++b)
    for (uint32_t i = t; i < m; i += kSubThreads) seq[0][b][i] = (uint16_t)A.inv[A.lists[(size_t)b * A.n + s0 + i]];
  if (t == 0) {
    nodes[0][0] = m << 16;
    count[0] = 1;
  }
  __syncthreads();
  uint32_t nsplit = 0;
  for (uint32_t d = 0;; ++d) {
    const uint32_t par = d & 1, nc = count[par];
    if (nc == 0) break;
    const uint32_t ax = (A.axis0 + d) % 3;
    for (uint32_t i = (uint32_t)wave; i < nc; i += kSubThreads / 64) {
      const uint32_t se = nodes[par][i], s = se & 0xffffu, e = se >> 16;
      const uint16_t* src = seq[par][ax];
      const uint32_t k = wave_split(s, e, src, gid, A.lo, A.hi, lane), mid = s + k;
      if (lane == 0) ksplit[i] = (uint16_t)k;
      for (uint32_t p = s + lane; p < e; p += 64) flag[src[p]] = p < mid ? 1 : 0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const unsigned long long below = (1ull << lane) - 1ull;
      for (int b = 0; b < 3; ++b) {
        const uint16_t* S = seq[par][b];
        uint16_t* D = seq[par ^ 1][b];
        uint32_t nl = 0, nr = 0;
        for (uint32_t c0 = s; c0 < e; c0 += 64) {
          const uint32_t p = c0 + lane;
          const bool valid = p < e;
          const uint16_t x = valid ? S[p] : 0;
          const bool f = valid && flag[x];
          const bool r = valid && !flag[x];
          const unsigned long long bl = __ballot(f), br = __ballot(r);
          if (f) D[s + nl + __popcll(bl & below)] = x;
          if (r) D[mid + nr + __popcll(br & below)] = x;
          nl += __popcll(bl);
          nr += __popcll(br);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // children of at most leaf_size elements are leaves: their elements in this axis's order
      const uint16_t* dst = seq[par ^ 1][ax];
      if (mid - s <= A.leaf)
        for (uint32_t p = s + lane; p < mid; p += 64) A.order[s0 + p] = gid[dst[p]];
      if (e - mid <= A.leaf)
        for (uint32_t p = mid + lane; p < e; p += 64) A.order[s0 + p] = gid[dst[p]];
    }
    __syncthreads();
    // the splits in level order, and the next level: the children that split, left before right
    for (uint32_t i = t; i < nc; i += kSubThreads) A.sub_split[s0 + nsplit + i] = ksplit[i];
    nsplit += nc;
    uint32_t base = 0;
    for (uint32_t i0 = 0; i0 < nc; i0 += kSubThreads) {
      const uint32_t i = i0 + t;
      uint32_t s = 0, e = 0, mid = 0;
      bool la = false, ra = false;
      if (i < nc) {
        const uint32_t se = nodes[par][i];
        s = se & 0xffffu;
        e = se >> 16;
        mid = s + ksplit[i];
        la = mid - s > A.leaf;
        ra = e - mid > A.leaf;
      }
      const uint32_t cn = (uint32_t)la + (uint32_t)ra;
      uint32_t x = cn, y = 0, tx, ty;
      block_scan_n<kSubThreads / 64>(x, wx, tx);
      (void)y;
      (void)wy;
      (void)ty;
      uint32_t q = base + x - cn;
      if (la) nodes[par ^ 1][q++] = s | (mid << 16);
      if (ra) nodes[par ^ 1][q] = mid | (e << 16);
      base += tx;
    }
    if (t == 0) count[par ^ 1] = base;
    __syncthreads();
  }
  if (t == 0) A.sub_count[g] = nsplit;
}

Topology GpuBvhBuilder::build_levels_on_device(ddgrs_ctx* ctx, size_t n, const double* lo, const double* hi,
                                               size_t leaf_size, bool allow_cut) {
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  // DDGRS_BVH_SUBTREES=0: every level through the level loop.  DDGRS_BVH_CUT_SIZE: the cut goes to
  // the first level whose largest node had at most this many elements in the previous build.
  const char* sub_env = std::getenv("DDGRS_BVH_SUBTREES");
  const char* cut_env = std::getenv("DDGRS_BVH_CUT_SIZE");
  const bool subtrees = !(sub_env && sub_env[0] == '0');
  const double cut_size = cut_env ? std::atof(cut_env) : 0.5 * kSubMax;
  const auto t_build = std::chrono::steady_clock::now();
  const uint32_t un = (uint32_t)n, leaf = (uint32_t)leaf_size;
  rec.reserve(4 * (n + 1));
  BvhRecords R{rec.ptr, rec.ptr + (n + 1), rec.ptr + 2 * (n + 1), rec.ptr + 3 * (n + 1)};
  lv_dev.reserve(kMaxBvhLevels + 1);
  const uint32_t max_nodes = un / (leaf + 1) + 1;  // nodes of a level have more than leaf_size elements
  ch_seg.reserve(un / kSplitChunk + max_nodes + 1);
  ch_c0.reserve(un / kSplitChunk + max_nodes + 1);
  seg_chunk0.reserve(max_nodes + 1);
  ch_box.reserve(18 * (size_t)(un / kSplitChunk + max_nodes + 1));
  ch_best.reserve(2 * (size_t)(un / kSplitChunk + max_nodes + 1));
  std::vector<BvhLevel> lv_host;
  std::vector<uint32_t> splits, sub_count_h, sub_split_h;
  int levels = 0;
  // the cut: the first level whose largest node the previous build found small enough for one
  // workgroup's subtree build
  int cut = -1;
  if (allow_cut && subtrees && n > leaf_size) {  // (a root of at most leaf_size elements is the tree)
    for (size_t L = 0; L < prev_levels.size(); ++L)
      if ((double)prev_levels[L] <= cut_size) {
        cut = (int)L;
        break;
      }
    // a cut build records the levels down to its cut only: below them the largest node shrinks by
    // at least a quarter a level (the split rule keeps a quarter on each side); a first build
    // starts that bound from the root
    if (cut < 0) {
      double est = prev_levels.empty() ? (double)n : prev_levels.back();
      int L = prev_levels.empty() ? 0 : (int)prev_levels.size() - 1;
      while (est > cut_size && L < (int)kMaxBvhLevels - 1) {
        est *= 0.75;
        ++L;
      }
      if (est <= cut_size) cut = L;
    }
  }
  if (n > leaf_size) {
    launch(ctx, "bvh_root", bvh_root_kernel, dim3(1), dim3(1), 0, un, lv_dev.ptr, R);
    uint32_t* pin = (uint32_t*)ctx->staging(sizeof(BvhLevel));
    // one level: its nodes' elements in the level's axis order, the splits, the next level's nodes
    auto run_level = [&](int L) {
      const BvhLevel* lv = lv_dev.ptr + L;
      // upper bounds of this level's nodes and of its largest node
      const uint64_t na_bound = std::min<uint64_t>(L < 32 ? (1ull << L) : max_nodes, max_nodes);
      const double size_bound = L < (int)prev_levels.size() ? 2.0 * prev_levels[L] : (double)n * std::pow(0.75, L);
      launch(ctx, "bvh_mark", bvh_mark_all_kernel, dim3(blocks_for(un, 256)), dim3(256), 0, un, lv, R,
             (const uint32_t*)order.ptr, seg_of.ptr);
      launch(ctx, "bvh_part_keys", bvh_part_keys_kernel, dim3(blocks_for(un, 256)), dim3(256), 0, un,
             (const uint32_t*)(axis_order.ptr + (size_t)(L % 3) * n), (const uint32_t*)seg_of.ptr, part_key_in.ptr,
             part_val_in.ptr);
      int bits = 1;
      while ((1ull << bits) - 1 < na_bound) ++bits;  // the sentinel truncates to 2^bits − 1 ≥ na
      size_t tb = temp.n;
      HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, part_key_in.ptr, part_key_out.ptr, part_val_in.ptr,
                                         part_val_out.ptr, n, 0, bits, ctx->stream));
      launch(ctx, "bvh_scatter", bvh_scatter_lv_kernel, dim3(blocks_for(un, 256)), dim3(256), 0, lv, R,
             (const uint32_t*)part_key_out.ptr, (const uint32_t*)part_val_out.ptr, order.ptr);
      if (size_bound <= (double)kSplitChunk) {
        launch(ctx, "bvh_split", bvh_split_lv_kernel, dim3((unsigned)na_bound), dim3(256), 0, lv, R,
               (const uint32_t*)order.ptr, lo, hi, suf.ptr);
      } else {
        const unsigned nch_bound = (unsigned)std::min<uint64_t>(un / kSplitChunk + na_bound, un / kSplitChunk + max_nodes);
        Box6* ctot = (Box6*)ch_box.ptr;
        Box6* cpre = ctot + (un / kSplitChunk + max_nodes + 1);
        Box6* csuf = cpre + (un / kSplitChunk + max_nodes + 1);
        launch(ctx, "bvh_chunks", bvh_chunks_lv_kernel, dim3(1), dim3(256), 0, lv_dev.ptr + L, R, ch_seg.ptr,
               ch_c0.ptr, seg_chunk0.ptr);
        launch(ctx, "bvh_chunk_total", bvh_chunk_total_kernel, dim3(nch_bound), dim3(256), 0, lv,
               (const uint32_t*)ch_seg.ptr, (const uint32_t*)ch_c0.ptr, (const uint32_t*)R.start,
               (const uint32_t*)R.end, (const uint32_t*)order.ptr, lo, hi, ctot);
        launch(ctx, "bvh_chunk_scan", bvh_chunk_scan_kernel, dim3(blocks_for(nch_bound, 256)), dim3(256), 0, lv, 0u,
               (const uint32_t*)ch_seg.ptr, (const uint32_t*)seg_chunk0.ptr, (const Box6*)ctot, cpre, csuf);
        launch(ctx, "bvh_chunk_suffix", bvh_chunk_suffix_kernel, dim3(nch_bound), dim3(256), 0, lv,
               (const uint32_t*)ch_seg.ptr, (const uint32_t*)ch_c0.ptr, (const uint32_t*)R.start,
               (const uint32_t*)R.end, (const uint32_t*)order.ptr, lo, hi, (const Box6*)csuf, suf.ptr);
        launch(ctx, "bvh_chunk_cost", bvh_chunk_cost_kernel, dim3(nch_bound), dim3(256), 0, lv,
               (const uint32_t*)ch_seg.ptr, (const uint32_t*)ch_c0.ptr, (const uint32_t*)R.start,
               (const uint32_t*)R.end, (const uint32_t*)order.ptr, lo, hi, (const Box6*)cpre, (const double*)suf.ptr,
               (SplitBest*)ch_best.ptr);
        launch(ctx, "bvh_chunk_pick", bvh_chunk_pick_kernel, dim3(blocks_for(na_bound, 256)), dim3(256), 0, lv, 0u,
               (const uint32_t*)seg_chunk0.ptr, (const uint32_t*)R.start, (const uint32_t*)R.end,
               (const SplitBest*)ch_best.ptr, R.split);
      }
      launch(ctx, "bvh_next", bvh_next_kernel, dim3(1), dim3(256), 0, lv_dev.ptr + L, R, leaf);
    };
    if (cut >= 0) {
      for (; levels < cut; ++levels) run_level(levels);
      // the cut level: its nodes' elements in each axis's order, then one workgroup per node
      const BvhLevel* lv = lv_dev.ptr + cut;
      const uint64_t na_bound = std::min<uint64_t>(cut < 32 ? (1ull << cut) : max_nodes, max_nodes);
      int bits = 1;
      while ((1ull << bits) - 1 < na_bound) ++bits;
      cut_lists.reserve(3 * n);
      sub_split.reserve(n);
      sub_count.reserve(na_bound);
      launch(ctx, "bvh_mark", bvh_mark_all_kernel, dim3(blocks_for(un, 256)), dim3(256), 0, un, lv, R,
             (const uint32_t*)order.ptr, seg_of.ptr);
      for (int b = 0; b < 3; ++b) {
        launch(ctx, "bvh_part_keys", bvh_part_keys_kernel, dim3(blocks_for(un, 256)), dim3(256), 0, un,
               (const uint32_t*)(axis_order.ptr + (size_t)b * n), (const uint32_t*)seg_of.ptr, part_key_in.ptr,
               part_val_in.ptr);
        size_t tb = temp.n;
        HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, part_key_in.ptr, part_key_out.ptr, part_val_in.ptr,
                                           part_val_out.ptr, n, 0, bits, ctx->stream));
        launch(ctx, "bvh_scatter", bvh_scatter_lv_kernel, dim3(blocks_for(un, 256)), dim3(256), 0, lv, R,
               (const uint32_t*)part_key_out.ptr, (const uint32_t*)part_val_out.ptr, cut_lists.ptr + (size_t)b * n);
      }
      // seg_of is free again: the subtrees' element → local id map
      const SubtreeArgs sa{lv,       R,      (const uint32_t*)cut_lists.ptr, seg_of.ptr, lo,
                           hi,       order.ptr, sub_split.ptr, sub_count.ptr, un,
                           leaf,     (uint32_t)(cut % 3)};
      launch(ctx, "bvh_subtree", bvh_subtree_kernel, dim3((unsigned)na_bound), dim3(kSubThreads), 0, sa);
      levels = cut;
      lv_host.resize(cut + 1);
      download(ctx, lv_host.data(), (const BvhLevel*)lv_dev.ptr, lv_host.size());
      sub_count_h.resize(na_bound);
      download(ctx, sub_count_h.data(), (const uint32_t*)sub_count.ptr, sub_count_h.size());
      ctx->sync();
      uint32_t too_large = UINT32_MAX;
      for (uint32_t g = 0; g < lv_host[cut].na && too_large == UINT32_MAX; ++g)
        if (sub_count_h[g] == UINT32_MAX) too_large = g;
      if (too_large == UINT32_MAX) {
        const size_t rec_bound = std::min<uint64_t>(cut < 32 ? (1ull << cut) : n + 1, n + 1);  // records above the cut
        splits.resize(rec_bound);
        download(ctx, splits.data(), (const uint32_t*)R.split, splits.size());
        sub_split_h.resize(n);
        download(ctx, sub_split_h.data(), (const uint32_t*)sub_split.ptr, n);
      } else {
        // a node too large for one workgroup: the level loop goes on from the cut level (the cut
        // level's records stand, and every cut node's range of `order` still holds its elements)
        if (trace)
          std::fprintf(stderr, "bvh: %zu elements, cut at level %d: node %u of %u too large\n", n, cut, too_large,
                       lv_host[cut].na);
        cut = -1;
      }
    }
    int target = prev_levels.empty() ? 16 : (int)prev_levels.size();
    for (;;) {
      if (cut >= 0) break;
      for (; levels < target; ++levels) run_level(levels);
      download(ctx, (BvhLevel*)pin, (const BvhLevel*)lv_dev.ptr + levels, 1);
      ctx->sync();
      if (((BvhLevel*)pin)->na == 0) break;
      if (levels >= (int)kMaxBvhLevels) throw Error(DDGRS_GPU, "BVH build", "more levels than the device build allows");
      target = std::min<int>(levels + 4, (int)kMaxBvhLevels);
    }
    if (cut < 0) {
      lv_host.resize(levels + 1);
      download(ctx, lv_host.data(), (const BvhLevel*)lv_dev.ptr, lv_host.size());
      ctx->sync();
      splits.resize(lv_host[levels].off);
      download(ctx, splits.data(), (const uint32_t*)R.split, splits.size());
    }
  }
  std::vector<uint32_t> ord(n);
  download(ctx, ord.data(), (const uint32_t*)order.ptr, n);
  ctx->sync();
  // the nodes in creation order, as the host-driven build makes them
  std::vector<TreeSegment> segs{{0, un, 0}};
  std::vector<size_t> level{0};
  std::vector<uint32_t> level_max;
  for (int L = 0; !level.empty(); ++L) {
    std::vector<size_t> next;
    uint32_t i = 0, biggest = 0;
    if (L == cut) {
      if ((size_t)L >= lv_host.size()) throw Error(DDGRS_GPU, "BVH build", "the cut level was not built");
      // each cut node's subtree from its splits in level order, left child before right
      for (size_t si : level) {
        if (segs[si].end - segs[si].start <= leaf_size) continue;
        if (i >= lv_host[L].na) throw Error(DDGRS_GPU, "BVH build", "device levels disagree with the nodes");
        biggest = std::max(biggest, segs[si].end - segs[si].start);
        const uint32_t s0 = segs[si].start, cnt = sub_count_h[i++];
        uint32_t j = 0;
        std::vector<size_t> q{si}, nq;
        while (!q.empty()) {
          nq.clear();
          for (size_t u : q) {
            if (j >= cnt || s0 + j >= sub_split_h.size())
              throw Error(DDGRS_GPU, "BVH build", "a subtree's splits disagree with its nodes");
            const uint32_t mid = segs[u].start + sub_split_h[s0 + j++];
            TreeSegment l{segs[u].start, mid, segs[u].depth + 1};
            TreeSegment r{mid, segs[u].end, segs[u].depth + 1};
            segs[u].left = (int64_t)segs.size();
            segs.push_back(l);
            segs[u].right = (int64_t)segs.size();
            segs.push_back(r);
            if (l.end - l.start > leaf_size) nq.push_back((size_t)segs[u].left);
            if (r.end - r.start > leaf_size) nq.push_back((size_t)segs[u].right);
          }
          std::swap(q, nq);
        }
        if (j != cnt) throw Error(DDGRS_GPU, "BVH build", "a subtree's splits disagree with its nodes");
      }
      if (i != lv_host[L].na) throw Error(DDGRS_GPU, "BVH build", "device levels disagree with the nodes");
      if (i > 0) level_max.push_back(biggest);
      break;
    }
    for (size_t si : level) {
      if (segs[si].end - segs[si].start <= leaf_size) continue;
      if (L >= levels || i >= lv_host[L].na) throw Error(DDGRS_GPU, "BVH build", "device levels disagree with the nodes");
      biggest = std::max(biggest, segs[si].end - segs[si].start);
      const uint32_t mid = segs[si].start + splits[lv_host[L].off + i++];
      TreeSegment l{segs[si].start, mid, segs[si].depth + 1};
      TreeSegment r{mid, segs[si].end, segs[si].depth + 1};
      segs[si].left = (int64_t)segs.size();
      segs.push_back(l);
      segs[si].right = (int64_t)segs.size();
      segs.push_back(r);
      next.push_back((size_t)segs[si].left);
      next.push_back((size_t)segs[si].right);
    }
    if (i == 0) break;
    if (i != lv_host[L].na) throw Error(DDGRS_GPU, "BVH build", "device levels disagree with the nodes");
    level_max.push_back(biggest);
    level = std::move(next);
  }
  prev_levels = std::move(level_max);
  if (trace)
    std::fprintf(stderr, "bvh: %zu elements, %d levels on the device%s, %.2f ms total\n", n, levels,
                 cut >= 0 ? " then the subtrees" : "",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t_build).count() * 1e3);
  return finish_topology(segs, std::move(ord));
}

Topology GpuBvhBuilder::build(ddgrs_ctx* ctx, size_t n, const double* pts, const double* lo, const double* hi,
                              size_t leaf_size) {
  const uint32_t un = (uint32_t)n;
  // DDGRS_TRACE_STEP=1: levels and the time spent waiting for each level's split positions
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  const auto t_build = std::chrono::steady_clock::now();
  double wait_ms = 0.0;
  int n_levels = 0;
  key_in.reserve(n);
  key_out.reserve(n);
  iota.reserve(n);
  axis_order.reserve(3 * n);
  order.reserve(n);
  seg_of.reserve(n);
  part_key_in.reserve(n);
  part_key_out.reserve(n);
  part_val_in.reserve(n);
  part_val_out.reserve(n);
  suf.reserve(6 * n);
  size_t t1 = 0, t2 = 0;
  HIPCHECK(rocprim::radix_sort_pairs(nullptr, t1, key_in.ptr, key_out.ptr, iota.ptr, iota.ptr, n, 0, 64, ctx->stream));
  HIPCHECK(rocprim::radix_sort_pairs(nullptr, t2, part_key_in.ptr, part_key_out.ptr, part_val_in.ptr, part_val_out.ptr,
                                     n, 0, 32, ctx->stream));
  temp.reserve(std::max(t1, t2));
  // each axis once, by (key, index): the input is in index order and the sort is stable
  for (int ax = 0; ax < 3; ++ax) {
    launch(ctx, "bvh_keys", bvh_keys_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, un, pts, ax, key_in.ptr,
           iota.ptr);
    size_t tb = temp.n;
    HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, key_in.ptr, key_out.ptr, iota.ptr, axis_order.ptr + (size_t)ax * n, n,
                                       0, 64, ctx->stream));
  }
  launch(ctx, "bvh_iota", bvh_iota_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, un, order.ptr);
  const char* host_levels = std::getenv("DDGRS_BVH_LEVELS_ON_HOST");
  if (!(host_levels && host_levels[0] == '1')) return build_levels_on_device(ctx, n, lo, hi, leaf_size);
  std::vector<TreeSegment> segs{{0, un, 0}};
  std::vector<size_t> level{0};
  std::vector<uint32_t> a_start, a_end, a_prefix, a_split;
  std::vector<size_t> a_seg;
  while (!level.empty()) {
    a_start.clear();
    a_end.clear();
    a_prefix.clear();
    a_seg.clear();
    uint32_t total = 0;
    for (size_t si : level) {
      const TreeSegment& s = segs[si];
      if (s.end - s.start <= leaf_size) continue;
      a_seg.push_back(si);
      a_start.push_back(s.start);
      a_end.push_back(s.end);
      a_prefix.push_back(total);
      total += s.end - s.start;
    }
    if (a_seg.empty()) break;
    const uint32_t na = (uint32_t)a_seg.size();
    const int axis = (int)(segs[a_seg[0]].depth % 3);
    upload(ctx, act_start, a_start);
    upload(ctx, act_end, a_end);
    upload(ctx, act_prefix, a_prefix);
    act_split.reserve(na);
    launch(ctx, "bvh_fill", bvh_fill_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, un, seg_of.ptr, UINT32_MAX);
    launch(ctx, "bvh_mark", bvh_mark_kernel, dim3(na), dim3(256), 0, (const uint32_t*)act_start.ptr,
           (const uint32_t*)act_end.ptr, (const uint32_t*)order.ptr, seg_of.ptr);
    launch(ctx, "bvh_part_keys", bvh_part_keys_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, un,
           (const uint32_t*)(axis_order.ptr + (size_t)axis * n), (const uint32_t*)seg_of.ptr, part_key_in.ptr,
           part_val_in.ptr);
    int bits = 1;
    while ((1u << bits) - 1u < na) ++bits;  // the sentinel UINT32_MAX truncates to 2^bits − 1 ≥ na
    size_t tb = temp.n;
    HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, part_key_in.ptr, part_key_out.ptr, part_val_in.ptr,
                                       part_val_out.ptr, n, 0, bits, ctx->stream));
    launch(ctx, "bvh_scatter", bvh_scatter_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, total,
           (const uint32_t*)part_key_out.ptr, (const uint32_t*)part_val_out.ptr, (const uint32_t*)act_start.ptr,
           (const uint32_t*)act_prefix.ptr, order.ptr);
    uint32_t longest = 0;
    for (uint32_t i = 0; i < na; ++i) longest = std::max(longest, a_end[i] - a_start[i]);
    if (longest <= kSplitChunk) {
      launch(ctx, "bvh_split", bvh_split_kernel, dim3(na), dim3(256), 0, (const uint32_t*)act_start.ptr,
             (const uint32_t*)act_end.ptr, (const uint32_t*)order.ptr, lo, hi, suf.ptr, act_split.ptr);
    } else {
      c_seg.clear();
      c_c0.clear();
      c_first.assign(1, 0);
      for (uint32_t i = 0; i < na; ++i) {
        for (uint32_t c0 = 0; c0 < a_end[i] - a_start[i]; c0 += kSplitChunk) {
          c_seg.push_back(i);
          c_c0.push_back(c0);
        }
        c_first.push_back((uint32_t)c_seg.size());
      }
      const uint32_t nch = (uint32_t)c_seg.size();
      upload(ctx, ch_seg, c_seg);
      upload(ctx, ch_c0, c_c0);
      upload(ctx, seg_chunk0, c_first);
      static_assert(sizeof(Box6) == 6 * sizeof(double) && sizeof(SplitBest) == 2 * sizeof(uint64_t), "scratch layout");
      ch_box.reserve(18 * (size_t)nch);
      ch_best.reserve(2 * (size_t)nch);
      Box6* ctot = (Box6*)ch_box.ptr;
      Box6* cpre = ctot + nch;
      Box6* csuf = cpre + nch;
      launch(ctx, "bvh_chunk_total", bvh_chunk_total_kernel, dim3(nch), dim3(256), 0, (const BvhLevel*)nullptr,
             (const uint32_t*)ch_seg.ptr,
             (const uint32_t*)ch_c0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const uint32_t*)order.ptr, lo, hi, ctot);
      launch(ctx, "bvh_chunk_scan", bvh_chunk_scan_kernel, dim3(blocks_for(nch, 256)), dim3(256), 0,
             (const BvhLevel*)nullptr, nch,
             (const uint32_t*)ch_seg.ptr, (const uint32_t*)seg_chunk0.ptr, (const Box6*)ctot, cpre, csuf);
      launch(ctx, "bvh_chunk_suffix", bvh_chunk_suffix_kernel, dim3(nch), dim3(256), 0, (const BvhLevel*)nullptr,
             (const uint32_t*)ch_seg.ptr,
             (const uint32_t*)ch_c0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const uint32_t*)order.ptr, lo, hi, (const Box6*)csuf, suf.ptr);
      launch(ctx, "bvh_chunk_cost", bvh_chunk_cost_kernel, dim3(nch), dim3(256), 0, (const BvhLevel*)nullptr,
             (const uint32_t*)ch_seg.ptr,
             (const uint32_t*)ch_c0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const uint32_t*)order.ptr, lo, hi, (const Box6*)cpre, (const double*)suf.ptr, (SplitBest*)ch_best.ptr);
      launch(ctx, "bvh_chunk_pick", bvh_chunk_pick_kernel, dim3(blocks_for(na, 256)), dim3(256), 0,
             (const BvhLevel*)nullptr, na,
             (const uint32_t*)seg_chunk0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const SplitBest*)ch_best.ptr, act_split.ptr);
    }
    a_split.resize(na);
    download(ctx, a_split.data(), (const uint32_t*)act_split.ptr, na);
    const auto t_wait = std::chrono::steady_clock::now();
    ctx->sync();
    if (trace) {
      wait_ms += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_wait).count() * 1e3;
      ++n_levels;
    }
    std::vector<size_t> next;
    for (uint32_t i = 0; i < na; ++i) {
      const size_t si = a_seg[i];
      const uint32_t mid = segs[si].start + a_split[i];
      TreeSegment l{segs[si].start, mid, segs[si].depth + 1};
      TreeSegment r{mid, segs[si].end, segs[si].depth + 1};
      segs[si].left = (int64_t)segs.size();
      segs.push_back(l);
      segs[si].right = (int64_t)segs.size();
      segs.push_back(r);
      next.push_back((size_t)segs[si].left);
      next.push_back((size_t)segs[si].right);
    }
    level = std::move(next);
  }
  std::vector<uint32_t> ord(n);
  download(ctx, ord.data(), (const uint32_t*)order.ptr, n);
  ctx->sync();
  if (trace)
    std::fprintf(stderr, "bvh: %zu elements, %d levels, %.2f ms waiting for the device, %.2f ms total\n", n,
                 n_levels, wait_ms, std::chrono::duration<double>(std::chrono::steady_clock::now() - t_build).count() * 1e3);
  return finish_topology(segs, std::move(ord));
}

}  // namespace ddgrs
