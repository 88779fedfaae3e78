// BVH construction on the device (see tree_gpu.hpp).
#include <cstring>
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

// One workgroup per splitting node: suffix boxes, then prefix boxes with the cost of every
// admissible split, reduced to the first minimum.
__global__ __launch_bounds__(256) void bvh_split_kernel(const uint32_t* __restrict__ act_start,
                                                        const uint32_t* __restrict__ act_end,
                                                        const uint32_t* __restrict__ order, const double* __restrict__ lo,
                                                        const double* __restrict__ hi, double* __restrict__ suf,
                                                        uint32_t* __restrict__ split) {
  __shared__ Box6 wave_tot[4];
  __shared__ double red_c[256];
  __shared__ uint32_t red_k[256];
  const uint32_t a = blockIdx.x, t = threadIdx.x;
  const uint32_t start = act_start[a], cnt = act_end[a] - start;
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
  if (t == 0) split[a] = red_k[0] == UINT32_MAX ? cnt / 2 : red_k[0];
}

// ---- chunked split for large nodes: several workgroups per node ------------------------------
// A node's elements are cut into chunks of kSplitChunk; the chunk totals are scanned per node, then
// every chunk writes its suffix boxes (one launch) and evaluates its splits (the next launch: a
// split's right box may lie in the next chunk), and the chunks' first minima are reduced per node.
// Box unions are exact, so the boxes, costs and the chosen split equal the one-workgroup kernel's.

__global__ __launch_bounds__(256) void bvh_chunk_total_kernel(const uint32_t* __restrict__ ch_seg,
                                                              const uint32_t* __restrict__ ch_c0,
                                                              const uint32_t* __restrict__ act_start,
                                                              const uint32_t* __restrict__ act_end,
                                                              const uint32_t* __restrict__ order,
                                                              const double* __restrict__ lo,
                                                              const double* __restrict__ hi, Box6* __restrict__ ctot) {
  __shared__ Box6 wave_tot[4];
  const uint32_t c = blockIdx.x, a = ch_seg[c];
  const uint32_t start = act_start[a], cnt = act_end[a] - start;
  const uint32_t c0 = ch_c0[c], c1 = min(c0 + kSplitChunk, cnt);
  Box6 v = empty6();
  for (uint32_t p = c0 + threadIdx.x; p < c1; p += 256) v = unite(v, load_box(lo, hi, order[start + p]));
  Box6 tot;
  (void)block_scan(v, wave_tot, &tot);
  if (threadIdx.x == 0) ctot[c] = tot;
}

// Per chunk: the union of the node's earlier chunks (cpre) and of its later chunks (csuf).
__global__ __launch_bounds__(256) void bvh_chunk_scan_kernel(uint32_t nch, const uint32_t* __restrict__ ch_seg,
                                                             const uint32_t* __restrict__ seg_chunk0,
                                                             const Box6* __restrict__ ctot, Box6* __restrict__ cpre,
                                                             Box6* __restrict__ csuf) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nch) return;
  const uint32_t a = ch_seg[c], f0 = seg_chunk0[a], f1 = seg_chunk0[a + 1];
  Box6 pre = empty6(), suf = empty6();
  for (uint32_t j = f0; j < c; ++j) pre = unite(pre, ctot[j]);
  for (uint32_t j = c + 1; j < f1; ++j) suf = unite(suf, ctot[j]);
  cpre[c] = pre;
  csuf[c] = suf;
}

__global__ __launch_bounds__(256) void bvh_chunk_suffix_kernel(const uint32_t* __restrict__ ch_seg,
                                                               const uint32_t* __restrict__ ch_c0,
                                                               const uint32_t* __restrict__ act_start,
                                                               const uint32_t* __restrict__ act_end,
                                                               const uint32_t* __restrict__ order,
                                                               const double* __restrict__ lo,
                                                               const double* __restrict__ hi,
                                                               const Box6* __restrict__ csuf, double* __restrict__ suf) {
  __shared__ Box6 wave_tot[4];
  const uint32_t c = blockIdx.x, a = ch_seg[c], t = threadIdx.x;
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

__global__ __launch_bounds__(256) void bvh_chunk_cost_kernel(const uint32_t* __restrict__ ch_seg,
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
  const uint32_t c = blockIdx.x, a = ch_seg[c], t = threadIdx.x;
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
__global__ __launch_bounds__(256) void bvh_chunk_pick_kernel(uint32_t na, const uint32_t* __restrict__ seg_chunk0,
                                                             const uint32_t* __restrict__ act_start,
                                                             const uint32_t* __restrict__ act_end,
                                                             const SplitBest* __restrict__ cbest,
                                                             uint32_t* __restrict__ split) {
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a >= na) return;
  double best = INFINITY;
  uint32_t best_k = UINT32_MAX;
  for (uint32_t j = seg_chunk0[a]; j < seg_chunk0[a + 1]; ++j) {
    const SplitBest b = cbest[j];
    if (b.k != UINT32_MAX && (best_k == UINT32_MAX || b.cost < best)) {
      best = b.cost;
      best_k = b.k;
    }
  }
  const uint32_t cnt = act_end[a] - act_start[a];
  split[a] = best_k == UINT32_MAX ? cnt / 2 : best_k;
}

Topology GpuBvhBuilder::build(ddgrs_ctx* ctx, size_t n, const double* pts, const double* lo, const double* hi,
                              size_t leaf_size) {
  const uint32_t un = (uint32_t)n;
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
      launch(ctx, "bvh_chunk_total", bvh_chunk_total_kernel, dim3(nch), dim3(256), 0, (const uint32_t*)ch_seg.ptr,
             (const uint32_t*)ch_c0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const uint32_t*)order.ptr, lo, hi, ctot);
      launch(ctx, "bvh_chunk_scan", bvh_chunk_scan_kernel, dim3(blocks_for(nch, 256)), dim3(256), 0, nch,
             (const uint32_t*)ch_seg.ptr, (const uint32_t*)seg_chunk0.ptr, (const Box6*)ctot, cpre, csuf);
      launch(ctx, "bvh_chunk_suffix", bvh_chunk_suffix_kernel, dim3(nch), dim3(256), 0, (const uint32_t*)ch_seg.ptr,
             (const uint32_t*)ch_c0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const uint32_t*)order.ptr, lo, hi, (const Box6*)csuf, suf.ptr);
      launch(ctx, "bvh_chunk_cost", bvh_chunk_cost_kernel, dim3(nch), dim3(256), 0, (const uint32_t*)ch_seg.ptr,
             (const uint32_t*)ch_c0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const uint32_t*)order.ptr, lo, hi, (const Box6*)cpre, (const double*)suf.ptr, (SplitBest*)ch_best.ptr);
      launch(ctx, "bvh_chunk_pick", bvh_chunk_pick_kernel, dim3(blocks_for(na, 256)), dim3(256), 0, na,
             (const uint32_t*)seg_chunk0.ptr, (const uint32_t*)act_start.ptr, (const uint32_t*)act_end.ptr,
             (const SplitBest*)ch_best.ptr, act_split.ptr);
    }
    a_split.resize(na);
    download(ctx, a_split.data(), (const uint32_t*)act_split.ptr, na);
    ctx->sync();
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
  return finish_topology(segs, std::move(ord));
}

}  // namespace ddgrs
