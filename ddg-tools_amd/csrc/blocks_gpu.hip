// Block cluster tree structure and near-field panels built on the device.
//
// Follows RS/src/spatial/block_tree.rs:129 build_impl: starting from (root, root), every pair of
// the current level is admissible (a ≠ b and the boxes χ-separated: d² > 0 and
// max(r_a, r_b)² ≤ χ² d², RS/src/metric/hierarchical.rs:116), near (both leaves, or
// |a|·|b| ≤ leaf_pairs), or split into the pairs of children (the row node's children outer, left
// child first).  Admissible and near pairs are kept in creation order and then grouped by row node
// stably, which is the order of build_block_structure.  The separation test uses the same FP64
// operations (no contraction), so the lists are bitwise the host build's.
//
// Per level: a classification kernel, a scan of the (admissible, near, children) counts and a
// scatter kernel; the host reads the three totals.  Grouping: a stable radix sort by row node and a
// binary search for the row pointers.  Panels: per leaf the column ranges of the near blocks of the
// leaf and its ancestors (each near block adds its column range to the leaves under its row node, a
// consecutive range of leaf ordinals), sorted by first column (the ranges of one leaf are disjoint:
// the blocks of a chain partition the columns), then the column lists and the panel table.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "blocktree.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

BlockStats block_stats(const BlockStructure& s) {
  BlockStats b;
  b.far_blocks = s.far_col.size();
  b.near_blocks = s.near_col.size();
  b.admissible_pairs = s.admissible_pairs;
  b.near_entries = s.near_entries;
  b.near_total = s.near_total;
  b.depth = s.depth;
  return b;
}

namespace {

struct AddCounts {
  __host__ __device__ BlockCounts operator()(const BlockCounts& x, const BlockCounts& y) const {
    return {x.adm + y.adm, x.nearb + y.nearb, x.kids + y.kids};
  }
};

struct BlockGeo {
  const uint32_t *start, *end, *right;
  const double *lo, *hi, *radius;
  double chi;
  uint64_t leaf_pairs;
};

__device__ __forceinline__ bool separated(const BlockGeo& g, uint32_t a, uint32_t b) {
  if (!(g.chi > 0.0)) return false;
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double d =
        fmax(fmax(g.lo[3 * a + k], g.lo[3 * b + k]) - fmin(g.hi[3 * a + k], g.hi[3 * b + k]), 0.0);
    d2 += d * d;
  }
  const double h = fmax(g.radius[a], g.radius[b]);
  return d2 > 0.0 && h * h <= g.chi * g.chi * d2;
}

}  // namespace

// The class of pair (a, b) (0 admissible, 1 near, 2 split) and what it adds to each list.
__device__ __forceinline__ BlockCounts classify_pair(const BlockGeo& g, uint32_t a, uint32_t b, uint8_t& c) {
  const bool la = g.right[a] == 0, lb = g.right[b] == 0;
  const uint64_t na = g.end[a] - g.start[a], nb = g.end[b] - g.start[b];
  if (a != b && separated(g, a, b))
    c = 0;
  else if ((la && lb) || na * nb <= g.leaf_pairs)
    c = 1;
  else
    c = 2;
  return {c == 0 ? 1u : 0u, c == 1 ? 1u : 0u, c == 2 ? (la ? 1u : 2u) * (lb ? 1u : 2u) : 0u};
}

// cls: 0 admissible, 1 near, 2 split; cnt: what the pair adds to each list.
// position[order[p]] = p (p < n) and inv_mass[a] = 1 / mass[a] (a < nn).
__global__ __launch_bounds__(256) void blk_tree_maps_kernel(uint32_t n, uint32_t nn, const uint32_t* __restrict__ order,
                                                            const double* __restrict__ mass,
                                                            uint32_t* __restrict__ position,
                                                            double* __restrict__ inv_mass) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) position[order[i]] = i;
  if (i < nn) inv_mass[i] = 1.0 / mass[i];
}

__global__ __launch_bounds__(256) void blk_classify_kernel(uint32_t m, const uint32_t* __restrict__ pa,
                                                           const uint32_t* __restrict__ pb, BlockGeo g,
                                                           uint8_t* __restrict__ cls, BlockCounts* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  uint8_t c;
  cnt[i] = classify_pair(g, pa[i], pb[i], c);
  cls[i] = c;
}

// ---- levels driven by the device ----
// The same classification without a host round trip per level: the level's pair count stays on the
// device, and each kernel covers it with a grid of at most kLevelBlocks blocks (sized from the
// previous build's level sizes), block b taking a
// contiguous range of the pairs.  A scan over the blocks' counts gives each block its first output
// slot, and the block scans its range in order, so the lists come out in the same order as with
// the per-level global scan.  List capacities are given; a level that would exceed one sets
// `overflow` and the level and all later ones do nothing.
constexpr uint32_t kLevelBlocks = 1024;
constexpr uint32_t kMaxLevels = 64;  // levels whose sizes are recorded for the next build's grids
// A first build's list capacities per tree node: pairs of a level, admissible blocks, near blocks
// (surface meshes of 10⁴–10⁵ faces have about 23, 70 and 30 per node).
constexpr uint64_t kFirstPairsPerNode = 48, kFirstAdmPerNode = 128, kFirstNearPerNode = 64;

struct LevelState {
  uint32_t m[2];                // pairs of the level, by level parity
  uint32_t adm_base, near_base; // list lengths before the current level
  uint32_t n_adm, n_near;       // list lengths after it
  uint32_t depth, max_m, overflow, pad;
};
struct LevelCaps {
  uint32_t pairs, adm, near;
};

__device__ __forceinline__ void level_range(uint32_t m, uint32_t& lo, uint32_t& hi) {
  const uint32_t per = (m + gridDim.x - 1) / gridDim.x;
  lo = min(m, blockIdx.x * per);
  hi = min(m, lo + per);
}

// Sum of 3 counts over the block's 256 threads (exclusive prefix per thread in `pre`, total returned).
__device__ __forceinline__ BlockCounts block_scan_counts(BlockCounts v, BlockCounts& pre) {
  __shared__ uint32_t sa[256], sn[256], sk[256];
  const int t = threadIdx.x;
  sa[t] = v.adm;
  sn[t] = v.nearb;
  sk[t] = v.kids;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    uint32_t xa = 0, xn = 0, xk = 0;
    if (t >= d) {
      xa = sa[t - d];
      xn = sn[t - d];
      xk = sk[t - d];
    }
    __syncthreads();
    sa[t] += xa;
    sn[t] += xn;
    sk[t] += xk;
    __syncthreads();
  }
  pre = {sa[t] - v.adm, sn[t] - v.nearb, sk[t] - v.kids};
  const BlockCounts tot{sa[255], sn[255], sk[255]};
  __syncthreads();
  return tot;
}

__global__ __launch_bounds__(256) void blk_level_count_kernel(const LevelState* __restrict__ st, int par,
                                                              const uint32_t* __restrict__ pa,
                                                              const uint32_t* __restrict__ pb, BlockGeo g,
                                                              uint8_t* __restrict__ cls,
                                                              BlockCounts* __restrict__ bsum) {
  const uint32_t m = st->overflow ? 0u : st->m[par];
  uint32_t lo, hi;
  level_range(m, lo, hi);
  BlockCounts acc{0, 0, 0};
  for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
    uint8_t c;
    const BlockCounts x = classify_pair(g, pa[i], pb[i], c);
    cls[i] = c;
    acc.adm += x.adm;
    acc.nearb += x.nearb;
    acc.kids += x.kids;
  }
  BlockCounts pre;
  const BlockCounts tot = block_scan_counts(acc, pre);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// One block of kLevelBlocks threads: the blocks' first output slots, the level's totals, the next
// level's count.
__global__ __launch_bounds__(1024) void blk_level_scan_kernel(LevelState* __restrict__ st, int par, uint32_t nblk,
                                                              uint32_t lev, const BlockCounts* __restrict__ bsum,
                                                              BlockCounts* __restrict__ boff, uint32_t* __restrict__ level_m,
                                                              LevelCaps caps) {
  __shared__ uint32_t sa[kLevelBlocks], sn[kLevelBlocks], sk[kLevelBlocks];
  const int t = threadIdx.x;
  const BlockCounts v = (uint32_t)t < nblk ? bsum[t] : BlockCounts{0, 0, 0};
  sa[t] = v.adm;
  sn[t] = v.nearb;
  sk[t] = v.kids;
  __syncthreads();
  for (int d = 1; d < (int)kLevelBlocks; d <<= 1) {
    uint32_t xa = 0, xn = 0, xk = 0;
    if (t >= d) {
      xa = sa[t - d];
      xn = sn[t - d];
      xk = sk[t - d];
    }
    __syncthreads();
    sa[t] += xa;
    sn[t] += xn;
    sk[t] += xk;
    __syncthreads();
  }
  boff[t] = {sa[t] - v.adm, sn[t] - v.nearb, sk[t] - v.kids};
  if (t == (int)kLevelBlocks - 1) {
    const uint32_t ta = sa[t], tn = sn[t], tk = sk[t];
    if (st->overflow) {
      st->m[par ^ 1] = 0;
      return;
    }
    const uint32_t m = st->m[par];
    if (lev < kMaxLevels) level_m[lev] = m;
    if (m > 0) ++st->depth;
    st->max_m = max(st->max_m, m);
    st->adm_base = st->n_adm;
    st->near_base = st->n_near;
    if ((uint64_t)st->n_adm + ta > caps.adm || (uint64_t)st->n_near + tn > caps.near || tk > caps.pairs) {
      st->overflow = 1;
      st->m[par ^ 1] = 0;
      return;
    }
    st->n_adm += ta;
    st->n_near += tn;
    st->m[par ^ 1] = tk;
  }
}

__global__ __launch_bounds__(256) void blk_level_scatter_kernel(const LevelState* __restrict__ st, int par,
                                                                const uint32_t* __restrict__ pa,
                                                                const uint32_t* __restrict__ pb,
                                                                const uint32_t* __restrict__ right,
                                                                const uint8_t* __restrict__ cls,
                                                                const BlockCounts* __restrict__ boff,
                                                                uint32_t* __restrict__ adm_a,
                                                                uint32_t* __restrict__ adm_b,
                                                                uint32_t* __restrict__ near_a,
                                                                uint32_t* __restrict__ near_b,
                                                                uint32_t* __restrict__ qa, uint32_t* __restrict__ qb) {
  if (st->overflow) return;
  const uint32_t m = st->m[par];
  uint32_t lo, hi;
  level_range(m, lo, hi);
  BlockCounts base = boff[blockIdx.x];
  base.adm += st->adm_base;
  base.nearb += st->near_base;
  for (uint32_t tile = lo; tile < hi; tile += 256) {
    const uint32_t i = tile + threadIdx.x;
    uint32_t a = 0, b = 0, ra = 0, rb = 0;
    uint8_t c = 3;
    BlockCounts x{0, 0, 0};
    if (i < hi) {
      a = pa[i];
      b = pb[i];
      c = cls[i];
      ra = right[a];
      rb = right[b];
      x = {c == 0 ? 1u : 0u, c == 1 ? 1u : 0u, c == 2 ? (ra ? 2u : 1u) * (rb ? 2u : 1u) : 0u};
    }
    BlockCounts pre;
    const BlockCounts tot = block_scan_counts(x, pre);
    if (c == 0) {
      adm_a[base.adm + pre.adm] = a;
      adm_b[base.adm + pre.adm] = b;
    } else if (c == 1) {
      near_a[base.nearb + pre.nearb] = a;
      near_b[base.nearb + pre.nearb] = b;
    } else if (c == 2) {
      const uint32_t ka[2] = {ra ? a + 1 : a, ra}, kb[2] = {rb ? b + 1 : b, rb};
      const int xa = ra ? 2 : 1, xb = rb ? 2 : 1;
      uint32_t q = base.kids + pre.kids;
      for (int u = 0; u < xa; ++u)
        for (int w = 0; w < xb; ++w) {
          qa[q] = ka[u];
          qb[q] = kb[w];
          ++q;
        }
    }
    base.adm += tot.adm;
    base.nearb += tot.nearb;
    base.kids += tot.kids;
  }
}

__global__ __launch_bounds__(256) void blk_scatter_kernel(uint32_t m, const uint32_t* __restrict__ pa,
                                                          const uint32_t* __restrict__ pb,
                                                          const uint32_t* __restrict__ right,
                                                          const uint8_t* __restrict__ cls,
                                                          const BlockCounts* __restrict__ off, uint32_t adm0,
                                                          uint32_t near0, uint32_t* __restrict__ adm_a,
                                                          uint32_t* __restrict__ adm_b, uint32_t* __restrict__ near_a,
                                                          uint32_t* __restrict__ near_b, uint32_t* __restrict__ qa,
                                                          uint32_t* __restrict__ qb) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const uint32_t a = pa[i], b = pb[i];
  const BlockCounts o = off[i];
  if (cls[i] == 0) {
    adm_a[adm0 + o.adm] = a;
    adm_b[adm0 + o.adm] = b;
  } else if (cls[i] == 1) {
    near_a[near0 + o.nearb] = a;
    near_b[near0 + o.nearb] = b;
  } else {
    const uint32_t ra = right[a], rb = right[b];
    const uint32_t ka[2] = {ra ? a + 1 : a, ra}, kb[2] = {rb ? b + 1 : b, rb};
    const int xa = ra ? 2 : 1, xb = rb ? 2 : 1;
    uint32_t q = o.kids;
    for (int x = 0; x < xa; ++x)
      for (int y = 0; y < xb; ++y) {
        qa[q] = ka[x];
        qb[q] = kb[y];
        ++q;
      }
  }
}

// ptr[a] = first position of row a in the sorted rows (a in [0, nn]).
__global__ __launch_bounds__(256) void blk_row_ptr_kernel(uint32_t nn, uint32_t m, const uint32_t* __restrict__ rows,
                                                          uint32_t* __restrict__ ptr) {
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a > nn) return;
  uint32_t lo = 0, hi = m;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (rows[mid] < a)
      lo = mid + 1;
    else
      hi = mid;
  }
  ptr[a] = lo;
}

// out[0] += Σ |a||b|;  out[1] += Σ (a = b ? |a|(|a| − 1) : |a||b|)  (integer sums).
__global__ __launch_bounds__(256) void blk_stats_kernel(uint32_t m, const uint32_t* __restrict__ pa,
                                                        const uint32_t* __restrict__ pb,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ end,
                                                        unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s0[256], s1[256];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  unsigned long long x = 0, y = 0;
  if (i < m) {
    const uint32_t a = pa[i], b = pb[i];
    const uint64_t na = end[a] - start[a], nb = end[b] - start[b];
    x = na * nb;
    y = a == b ? na * (na - 1) : na * nb;
  }
  s0[threadIdx.x] = x;
  s1[threadIdx.x] = y;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], s0[0]);
    atomicAdd(&out[1], s1[0]);
  }
}

// Each near block (A, B) gives every leaf under A (leaf ordinals [lf_A, le_A)) the column range of B:
// per leaf the number of ranges and of columns (integer atomic sums, independent of order).
__global__ __launch_bounds__(256) void blk_near_count_kernel(uint32_t n_near, const uint32_t* __restrict__ near_row,
                                                             const uint32_t* __restrict__ near_col,
                                                             const uint32_t* __restrict__ start,
                                                             const uint32_t* __restrict__ end,
                                                             const uint32_t* __restrict__ leaf_first,
                                                             const uint32_t* __restrict__ leaf_end,
                                                             uint32_t* __restrict__ rcount,
                                                             uint32_t* __restrict__ kcount) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n_near) return;
  const uint32_t a = near_row[q], b = near_col[q], w = end[b] - start[b];
  for (uint32_t k = leaf_first[a]; k < leaf_end[a]; ++k) {
    atomicAdd(&rcount[k], 1u);
    atomicAdd(&kcount[k], w);
  }
}

// Per leaf ordinal k < nl: panels, their segments and stored entries; zeros at k = nl (the scans
// then end with the totals).
__global__ __launch_bounds__(256) void blk_panel_count_kernel(uint32_t nl, const uint32_t* __restrict__ leaves,
                                                              const uint32_t* __restrict__ start,
                                                              const uint32_t* __restrict__ end,
                                                              const uint32_t* __restrict__ kcount, uint32_t seg_cols,
                                                              uint32_t* __restrict__ pcount,
                                                              uint32_t* __restrict__ scount,
                                                              uint64_t* __restrict__ entries) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k > nl) return;
  if (k == nl) {
    pcount[k] = 0;
    scount[k] = 0;
    entries[k] = 0;
    return;
  }
  const uint32_t leaf = leaves[k], ns = end[leaf] - start[leaf], K = kcount[k];
  pcount[k] = (ns + kPanelRows - 1) / kPanelRows;
  scount[k] = pcount[k] * max(1u, (K + seg_cols - 1) / seg_cols);
  entries[k] = (uint64_t)ns * K;
}

// The ranges of each leaf, in any order (a per-leaf cursor); they are sorted by first column next.
__global__ __launch_bounds__(256) void blk_range_fill_kernel(uint32_t n_near, const uint32_t* __restrict__ near_row,
                                                             const uint32_t* __restrict__ near_col,
                                                             const uint32_t* __restrict__ start,
                                                             const uint32_t* __restrict__ end,
                                                             const uint32_t* __restrict__ leaf_first,
                                                             const uint32_t* __restrict__ leaf_end,
                                                             const uint32_t* __restrict__ range_ptr,
                                                             uint32_t* __restrict__ cursor, uint64_t* __restrict__ key,
                                                             uint32_t* __restrict__ rend) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n_near) return;
  const uint32_t a = near_row[q], b = near_col[q];
  for (uint32_t k = leaf_first[a]; k < leaf_end[a]; ++k) {
    const uint32_t o = range_ptr[k] + atomicAdd(&cursor[k], 1u);
    key[o] = ((uint64_t)k << 32) | start[b];
    rend[o] = end[b];
  }
}

// Columns of each sorted range (the ranges sorted by (leaf, first column) are the leaves' column
// lists back to back, so their offsets are the exclusive sums of their lengths).
__global__ __launch_bounds__(256) void blk_range_len_kernel(uint32_t nr, const uint64_t* __restrict__ key,
                                                            const uint32_t* __restrict__ rend,
                                                            uint32_t* __restrict__ len) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r < nr) len[r] = rend[r] - (uint32_t)key[r];
}

// One thread per range: its columns at its offset.  A range is one near block's column node, which
// is a leaf or has at most leaf_pairs elements, so the loop is short.
__global__ __launch_bounds__(256) void blk_cols_kernel(uint32_t nr, const uint64_t* __restrict__ key,
                                                       const uint32_t* __restrict__ rend,
                                                       const uint32_t* __restrict__ off,
                                                       uint32_t* __restrict__ cols) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nr) return;
  const uint32_t b0 = (uint32_t)key[r], b1 = rend[r], o = off[r];
  for (uint32_t x = b0; x < b1; ++x) cols[o + (x - b0)] = x;
}

struct PanelOut {
  uint32_t *leaf, *row0, *rows;
  uint64_t* val;
  uint32_t *seg_panel, *seg_col0, *seg_ptr;
};

// The panels of each leaf and their segments (panel after panel, columns ascending).
__global__ __launch_bounds__(256) void blk_panel_fill_kernel(uint32_t nl, const uint32_t* __restrict__ leaves,
                                                             const uint32_t* __restrict__ start,
                                                             const uint32_t* __restrict__ end,
                                                             const uint32_t* __restrict__ col_ptr,
                                                             const uint32_t* __restrict__ pan_ptr,
                                                             const uint32_t* __restrict__ seg_base,
                                                             const uint64_t* __restrict__ val_base, uint32_t seg_cols,
                                                             PanelOut o) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nl) return;
  const uint32_t leaf = leaves[k], K = col_ptr[k + 1] - col_ptr[k];
  const uint32_t nseg = max(1u, (K + seg_cols - 1) / seg_cols);
  uint32_t p = pan_ptr[k], sg = seg_base[k];
  for (uint32_t r0 = start[leaf]; r0 < end[leaf]; r0 += kPanelRows, ++p) {
    o.leaf[p] = k;
    o.row0[p] = r0;
    o.rows[p] = min(kPanelRows, end[leaf] - r0);
    o.val[p] = val_base[k] + (uint64_t)(r0 - start[leaf]) * K;
    o.seg_ptr[p] = sg;
    for (uint32_t j = 0; j < nseg; ++j, ++sg) {
      o.seg_panel[sg] = p;
      o.seg_col0[sg] = j * seg_cols;
    }
  }
  if (k == nl - 1) o.seg_ptr[pan_ptr[nl]] = seg_base[nl];
}

__global__ __launch_bounds__(256) void blk_weight_kernel(uint32_t n, const uint32_t* __restrict__ order,
                                                         const double* __restrict__ w, double* __restrict__ out) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p < n) out[p] = w[order[p]];
}

namespace {

template <class T>
void scan_exclusive(ddgrs_ctx* ctx, DevBuf<uint8_t>& temp, const T* in, T* out, size_t n) {
  size_t tb = 0;
  HIPCHECK(rocprim::exclusive_scan(nullptr, tb, in, out, T(0), n, rocprim::plus<T>(), ctx->stream));
  temp.reserve(tb);
  tb = temp.n;
  HIPCHECK(rocprim::exclusive_scan(temp.ptr, tb, in, out, T(0), n, rocprim::plus<T>(), ctx->stream));
}

template <class K>
void sort_pairs(ddgrs_ctx* ctx, DevBuf<uint8_t>& temp, const K* ki, K* ko, const uint32_t* vi, uint32_t* vo, size_t n,
                int bits) {
  size_t tb = 0;
  HIPCHECK(rocprim::radix_sort_pairs(nullptr, tb, ki, ko, vi, vo, n, 0, bits, ctx->stream));
  temp.reserve(tb);
  tb = temp.n;
  HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, ki, ko, vi, vo, n, 0, bits, ctx->stream));
}

int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (1ull << b) <= maxval) ++b;
  return b;
}

// Grows `b` to hold `need` elements, keeping the first `used`.
void grow_keep(ddgrs_ctx* ctx, DevBuf<uint32_t>& b, size_t used, size_t need) {
  if (need <= b.n && b.ptr) return;
  DevBuf<uint32_t> nb(std::max<size_t>(need, 2 * b.n));
  if (used) HIPCHECK(hipMemcpyAsync(nb.ptr, b.ptr, used * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
  ctx->sync();
  b = std::move(nb);
}

}  // namespace

// Per-level totals read back to the host: the level's lists are sized exactly.
static void host_levels(ddgrs_ctx* ctx, BlockBuildScratch& S, const BlockGeo& g, const uint32_t* right, uint32_t& n_adm,
                        uint32_t& n_near, uint32_t& depth, uint32_t& max_m) {
  S.pa.reserve(1);
  S.pb.reserve(1);
  HIPCHECK(hipMemsetAsync(S.pa.ptr, 0, sizeof(uint32_t), ctx->stream));  // the root pair
  HIPCHECK(hipMemsetAsync(S.pb.ptr, 0, sizeof(uint32_t), ctx->stream));
  uint32_t m = 1;
  n_adm = n_near = depth = 0;
  max_m = 1;
  BlockCounts* pin = (BlockCounts*)ctx->staging(2 * sizeof(BlockCounts));
  S.prev_m.clear();
  while (m > 0) {
    ++depth;
    max_m = std::max(max_m, m);
    if (S.prev_m.size() < kMaxLevels) S.prev_m.push_back(m);
    S.cls.reserve(m);
    S.cnt.reserve(m);
    S.off.reserve(m);
    launch(ctx, "blk_classify", blk_classify_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, m,
           (const uint32_t*)S.pa.ptr, (const uint32_t*)S.pb.ptr, g, S.cls.ptr, S.cnt.ptr);
    size_t tb = 0;
    const BlockCounts zero{0, 0, 0};
    HIPCHECK(rocprim::exclusive_scan(nullptr, tb, S.cnt.ptr, S.off.ptr, zero, m, AddCounts(), ctx->stream));
    S.temp.reserve(tb);
    tb = S.temp.n;
    HIPCHECK(rocprim::exclusive_scan(S.temp.ptr, tb, S.cnt.ptr, S.off.ptr, zero, m, AddCounts(), ctx->stream));
    download(ctx, pin, (const BlockCounts*)S.off.ptr + (m - 1), 1);
    download(ctx, pin + 1, (const BlockCounts*)S.cnt.ptr + (m - 1), 1);
    ctx->sync();
    const BlockCounts tot = AddCounts()(pin[0], pin[1]);
    grow_keep(ctx, S.adm_a, n_adm, (size_t)n_adm + tot.adm);
    grow_keep(ctx, S.adm_b, n_adm, (size_t)n_adm + tot.adm);
    grow_keep(ctx, S.near_a, n_near, (size_t)n_near + tot.nearb);
    grow_keep(ctx, S.near_b, n_near, (size_t)n_near + tot.nearb);
    S.qa.reserve(std::max<uint32_t>(tot.kids, 1));
    S.qb.reserve(std::max<uint32_t>(tot.kids, 1));
    launch(ctx, "blk_scatter", blk_scatter_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, m,
           (const uint32_t*)S.pa.ptr, (const uint32_t*)S.pb.ptr, right, (const uint8_t*)S.cls.ptr,
           (const BlockCounts*)S.off.ptr, n_adm, n_near, S.adm_a.ptr, S.adm_b.ptr, S.near_a.ptr, S.near_b.ptr,
           S.qa.ptr, S.qb.ptr);
    n_adm += tot.adm;
    n_near += tot.nearb;
    m = tot.kids;
    std::swap(S.pa, S.qa);
    std::swap(S.pb, S.qb);
  }
}

// Levels driven by the device within the scratch's capacities: batches of levels launched without
// reading anything back, then one read of the level state.  False when a list would have exceeded
// its capacity (nothing usable was produced).
static bool device_levels(ddgrs_ctx* ctx, BlockBuildScratch& S, const BlockGeo& g, const uint32_t* right, uint32_t& n_adm,
                          uint32_t& n_near, uint32_t& depth, uint32_t& max_m) {
  const LevelCaps caps{(uint32_t)std::min<uint64_t>(S.cap_pairs, UINT32_MAX),
                       (uint32_t)std::min<uint64_t>(S.cap_adm, UINT32_MAX),
                       (uint32_t)std::min<uint64_t>(S.cap_near, UINT32_MAX)};
  S.pa.reserve(caps.pairs);
  S.pb.reserve(caps.pairs);
  S.qa.reserve(caps.pairs);
  S.qb.reserve(caps.pairs);
  S.cls.reserve(caps.pairs);
  S.adm_a.reserve(caps.adm);
  S.adm_b.reserve(caps.adm);
  S.near_a.reserve(caps.near);
  S.near_b.reserve(caps.near);
  S.bsum.reserve(kLevelBlocks);
  S.boff.reserve(kLevelBlocks);
  S.level_m.reserve(kMaxLevels);
  // a level's grid: about 512 pairs per block for the size the previous build had at that level
  auto grid_for = [&](uint32_t lev) -> uint32_t {
    if (S.prev_m.empty()) return kLevelBlocks / 4;  // a first build: no sizes yet
    const uint64_t m = lev < S.prev_m.size() ? S.prev_m[lev] : 0;
    return (uint32_t)std::min<uint64_t>(kLevelBlocks, std::max<uint64_t>(1, (m + m / 4 + 511) / 512));
  };
  S.lstate.reserve(sizeof(LevelState) / sizeof(uint32_t));
  LevelState* st = (LevelState*)S.lstate.ptr;
  HIPCHECK(hipMemsetAsync(st, 0, sizeof(LevelState), ctx->stream));
  HIPCHECK(hipMemsetD32Async((hipDeviceptr_t)&st->m[0], 1u, 1, ctx->stream));  // level 0: the root pair
  HIPCHECK(hipMemsetAsync(S.pa.ptr, 0, sizeof(uint32_t), ctx->stream));
  HIPCHECK(hipMemsetAsync(S.pb.ptr, 0, sizeof(uint32_t), ctx->stream));
  uint32_t *xa = S.pa.ptr, *xb = S.pb.ptr, *ya = S.qa.ptr, *yb = S.qb.ptr;
  LevelState* hs = (LevelState*)ctx->staging(sizeof(LevelState));
  uint32_t lev = 0, batch = S.last_depth + 2;
  for (;;) {
    for (uint32_t k = 0; k < batch; ++k, ++lev) {
      const int par = (int)(lev & 1);
      const uint32_t nblk = grid_for(lev);
      launch(ctx, "blk_level_count", blk_level_count_kernel, dim3(nblk), dim3(256), 0, (const LevelState*)st, par,
             (const uint32_t*)xa, (const uint32_t*)xb, g, S.cls.ptr, S.bsum.ptr);
      launch(ctx, "blk_level_scan", blk_level_scan_kernel, dim3(1), dim3(kLevelBlocks), 0, st, par, nblk, lev,
             (const BlockCounts*)S.bsum.ptr, S.boff.ptr, S.level_m.ptr, caps);
      launch(ctx, "blk_level_scatter", blk_level_scatter_kernel, dim3(nblk), dim3(256), 0,
             (const LevelState*)st, par, (const uint32_t*)xa, (const uint32_t*)xb, right, (const uint8_t*)S.cls.ptr,
             (const BlockCounts*)S.boff.ptr, S.adm_a.ptr, S.adm_b.ptr, S.near_a.ptr, S.near_b.ptr, ya, yb);
      std::swap(xa, ya);
      std::swap(xb, yb);
    }
    download(ctx, hs, (const LevelState*)st, 1);
    ctx->sync();
    if (hs->overflow) return false;
    if (hs->m[lev & 1] == 0) break;
    batch = 4;
  }
  n_adm = hs->n_adm;
  n_near = hs->n_near;
  depth = hs->depth;
  max_m = hs->max_m;
  S.prev_m.resize(std::min<uint32_t>(lev, kMaxLevels));
  download(ctx, S.prev_m.data(), (const uint32_t*)S.level_m.ptr, S.prev_m.size());
  ctx->sync();
  return true;
}

BlockStats BlockTreeDev::build_structure(ddgrs_ctx* ctx, const Topology& t, const Clusters& cl, double chi,
                                         size_t leaf_pairs, int nk_, const double* elem_weight, const DeviceTree* dt) {
  BlockBuildScratch& S = scratch;
  // DDGRS_TRACE_STEP=1: wall times of the build's phases on stderr
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  auto t_part = std::chrono::steady_clock::now();
  auto mark = [&](const char* what) {
    if (!trace) return;
    ctx->sync();
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "blocks(%d): %-14s %8.2f ms\n", n, what,
                 std::chrono::duration<double>(now - t_part).count() * 1e3);
    t_part = now;
  };
  nk = nk_;
  n = (int)t.order.size();
  nn = (int)t.n_nodes();
  nl = (int)t.leaves.size();
  if (dt) {
    auto copy = [&](auto& d, const auto* src, size_t cnt) {
      d.reserve(cnt);
      HIPCHECK(hipMemcpyAsync(d.ptr, src, cnt * sizeof(*src), hipMemcpyDeviceToDevice, ctx->stream));
    };
    copy(order, dt->order, (size_t)n);
    copy(start, dt->start, (size_t)nn);
    copy(end, dt->end, (size_t)nn);
    copy(right, dt->right, (size_t)nn);
    copy(parent, dt->parent, (size_t)nn);
    copy(leaves, dt->leaves, (size_t)nl);
    copy(S.lo, dt->lo, 3 * (size_t)nn);
    copy(S.hi, dt->hi, 3 * (size_t)nn);
    copy(S.radius, dt->radius, (size_t)nn);
    position.reserve(n);
    inv_mass.reserve(nn);
    launch(ctx, "blk_tree_maps", blk_tree_maps_kernel, dim3(blocks_for(std::max(n, nn), 256)), dim3(256), 0,
           (uint32_t)n, (uint32_t)nn, dt->order, dt->mass, position.ptr, inv_mass.ptr);
  } else {
    ddgrs::upload(ctx, order, t.order);
    std::vector<uint32_t> pos(n);
    for (int p = 0; p < n; ++p) pos[t.order[p]] = (uint32_t)p;
    ddgrs::upload(ctx, position, pos);
    ddgrs::upload(ctx, start, t.start);
    ddgrs::upload(ctx, end, t.end);
    ddgrs::upload(ctx, right, t.right);
    ddgrs::upload(ctx, parent, t.parent);
    ddgrs::upload(ctx, leaves, t.leaves);
    std::vector<double> im(nn);
    for (int a = 0; a < nn; ++a) im[a] = 1.0 / cl.mass[a];
    ddgrs::upload(ctx, inv_mass, im);
    ddgrs::upload(ctx, S.lo, cl.lo);
    ddgrs::upload(ctx, S.hi, cl.hi);
    ddgrs::upload(ctx, S.radius, cl.radius);
  }
  weight.reserve(n);
  launch(ctx, "blk_weight", blk_weight_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (uint32_t)n,
         (const uint32_t*)order.ptr, elem_weight, weight.ptr);
  const BlockGeo g{start.ptr, end.ptr, right.ptr, S.lo.ptr, S.hi.ptr, S.radius.ptr, chi, (uint64_t)leaf_pairs};

  mark("uploads");
  // level-synchronous classification: driven by the device with the capacities of an earlier build,
  // else (first build, or a list outgrew its capacity) with the totals read back per level
  // DDGRS_BLOCK_LEVELS_ON_HOST=1 always reads the totals back per level (for comparisons)
  const bool levels_on_host = std::getenv("DDGRS_BLOCK_LEVELS_ON_HOST") != nullptr;
  uint32_t n_adm = 0, n_near = 0, depth = 0, max_m = 1;
  // a first build sizes its lists from the node count; a list that outgrows its capacity makes the
  // device pass start over with four times the capacities (the host pass after the third)
  if (!S.cap_pairs) {
    S.cap_pairs = kFirstPairsPerNode * (uint64_t)nn + 4096;
    S.cap_adm = kFirstAdmPerNode * (uint64_t)nn + 4096;
    S.cap_near = kFirstNearPerNode * (uint64_t)nn + 4096;
  }
  bool on_device = false;
  for (int attempt = 0; attempt < 3 && !levels_on_host && !on_device; ++attempt) {
    on_device = device_levels(ctx, S, g, right.ptr, n_adm, n_near, depth, max_m);
    if (!on_device) {
      if (trace) std::fprintf(stderr, "blocks(%d): device levels overflowed their capacities\n", n);
      S.cap_pairs *= 4;
      S.cap_adm *= 4;
      S.cap_near *= 4;
    }
  }
  if (!on_device) host_levels(ctx, S, g, right.ptr, n_adm, n_near, depth, max_m);
  if (trace)
    std::fprintf(stderr, "blocks(%d): levels %s, depth %u, largest level %u pairs, %u admissible, %u near\n", n,
                 on_device ? "on the device" : "on the host", depth, max_m, n_adm, n_near);
  S.cap_pairs = std::max<uint64_t>(S.cap_pairs, (uint64_t)max_m + max_m / 4 + 1024);
  S.cap_adm = std::max<uint64_t>(S.cap_adm, (uint64_t)n_adm + n_adm / 4 + 1024);
  S.cap_near = std::max<uint64_t>(S.cap_near, (uint64_t)n_near + n_near / 4 + 1024);
  S.last_depth = depth;

  mark("levels");
  // grouped by row node, creation order kept
  const int row_bits = bits_for((uint64_t)nn);
  n_far = n_adm;
  far_row.reserve(std::max<uint32_t>(n_far, 1));
  far_col.reserve(std::max<uint32_t>(n_far, 1));
  far_ptr.reserve(nn + 1);
  if (n_far) sort_pairs<uint32_t>(ctx, S.temp, S.adm_a.ptr, far_row.ptr, S.adm_b.ptr, far_col.ptr, n_far, row_bits);
  launch(ctx, "blk_row_ptr", blk_row_ptr_kernel, dim3(blocks_for(nn + 1, 256)), dim3(256), 0, (uint32_t)nn, n_far,
         (const uint32_t*)far_row.ptr, far_ptr.ptr);
  S.near_row.reserve(std::max<uint32_t>(n_near, 1));
  S.near_col.reserve(std::max<uint32_t>(n_near, 1));
  S.near_ptr.reserve(nn + 1);
  this->n_near = n_near;
  if (n_near)
    sort_pairs<uint32_t>(ctx, S.temp, S.near_a.ptr, S.near_row.ptr, S.near_b.ptr, S.near_col.ptr, n_near, row_bits);
  launch(ctx, "blk_row_ptr", blk_row_ptr_kernel, dim3(blocks_for(nn + 1, 256)), dim3(256), 0, (uint32_t)nn, n_near,
         (const uint32_t*)S.near_row.ptr, S.near_ptr.ptr);
  S.sums.reserve(4);
  HIPCHECK(hipMemsetAsync(S.sums.ptr, 0, 4 * sizeof(unsigned long long), ctx->stream));
  launch(ctx, "blk_stats", blk_stats_kernel, dim3(blocks_for(n_far, 256)), dim3(256), 0, n_far,
         (const uint32_t*)far_row.ptr, (const uint32_t*)far_col.ptr, (const uint32_t*)start.ptr,
         (const uint32_t*)end.ptr, S.sums.ptr);
  launch(ctx, "blk_stats", blk_stats_kernel, dim3(blocks_for(n_near, 256)), dim3(256), 0, n_near,
         (const uint32_t*)S.near_row.ptr, (const uint32_t*)S.near_col.ptr, (const uint32_t*)start.ptr,
         (const uint32_t*)end.ptr, S.sums.ptr + 2);

  mark("grouping");
  // panels
  const size_t nl1 = (size_t)nl + 1;
  S.rcount.reserve(nl1);
  S.kcount.reserve(nl1);
  S.pcount.reserve(nl1);
  S.entries.reserve(nl1);
  S.range_ptr.reserve(nl1);
  S.pan_ptr.reserve(nl1);
  S.val_base.reserve(nl1);
  col_ptr.reserve(nl1);
  // leaf ordinal ranges of the nodes (pre-order: a node's leaves are consecutive)
  {
    std::vector<uint32_t> lf(nn), le(nn);
    for (int x = nn; x-- > 0;) {
      if (t.right[x] == 0) {
        lf[x] = t.leaf_ordinal[x];
        le[x] = t.leaf_ordinal[x] + 1;
      } else {
        lf[x] = lf[x + 1];
        le[x] = le[t.right[x]];
      }
    }
    ddgrs::upload(ctx, S.leaf_first, lf);
    ddgrs::upload(ctx, S.leaf_end, le);
    S.cursor.reserve(nl1);
    HIPCHECK(hipMemsetAsync(S.rcount.ptr, 0, nl1 * sizeof(uint32_t), ctx->stream));
    HIPCHECK(hipMemsetAsync(S.kcount.ptr, 0, nl1 * sizeof(uint32_t), ctx->stream));
    HIPCHECK(hipMemsetAsync(S.cursor.ptr, 0, nl1 * sizeof(uint32_t), ctx->stream));
    launch(ctx, "blk_near_count", blk_near_count_kernel, dim3(blocks_for(n_near, 256)), dim3(256), 0, n_near,
           (const uint32_t*)S.near_row.ptr, (const uint32_t*)S.near_col.ptr, (const uint32_t*)start.ptr,
           (const uint32_t*)end.ptr, (const uint32_t*)S.leaf_first.ptr, (const uint32_t*)S.leaf_end.ptr, S.rcount.ptr,
           S.kcount.ptr);
    ctx->sync();  // lf, le are released on return
  }
  seg_cols = panel_segment_cols();
  S.scount.reserve(nl1);
  S.seg_base.reserve(nl1);
  launch(ctx, "blk_panel_count", blk_panel_count_kernel, dim3(blocks_for(nl1, 256)), dim3(256), 0, (uint32_t)nl,
         (const uint32_t*)leaves.ptr, (const uint32_t*)start.ptr, (const uint32_t*)end.ptr, (const uint32_t*)S.kcount.ptr,
         seg_cols, S.pcount.ptr, S.scount.ptr, S.entries.ptr);
  scan_exclusive<uint32_t>(ctx, S.temp, S.rcount.ptr, S.range_ptr.ptr, nl1);
  scan_exclusive<uint32_t>(ctx, S.temp, S.kcount.ptr, col_ptr.ptr, nl1);
  scan_exclusive<uint32_t>(ctx, S.temp, S.pcount.ptr, S.pan_ptr.ptr, nl1);
  scan_exclusive<uint32_t>(ctx, S.temp, S.scount.ptr, S.seg_base.ptr, nl1);
  scan_exclusive<uint64_t>(ctx, S.temp, S.entries.ptr, S.val_base.ptr, nl1);
  struct Tails {
    uint32_t ranges, cols, panels, segs;
    uint64_t entries;
    unsigned long long sums[4];
  };
  Tails* tl = (Tails*)ctx->staging(sizeof(Tails));
  download(ctx, &tl->ranges, (const uint32_t*)S.range_ptr.ptr + nl, 1);
  download(ctx, &tl->cols, (const uint32_t*)col_ptr.ptr + nl, 1);
  download(ctx, &tl->panels, (const uint32_t*)S.pan_ptr.ptr + nl, 1);
  download(ctx, &tl->segs, (const uint32_t*)S.seg_base.ptr + nl, 1);
  download(ctx, &tl->entries, (const uint64_t*)S.val_base.ptr + nl, 1);
  download(ctx, tl->sums, (const unsigned long long*)S.sums.ptr, 4);
  ctx->sync();
  const Tails T = *tl;
  np = (int)T.panels;
  nseg = (int)T.segs;
  near_total = T.entries;
  const size_t nr = std::max<uint32_t>(T.ranges, 1);
  S.range_key.reserve(nr);
  S.range_key_s.reserve(nr);
  S.range_end.reserve(nr);
  S.range_end_s.reserve(nr);
  launch(ctx, "blk_range_fill", blk_range_fill_kernel, dim3(blocks_for(n_near, 256)), dim3(256), 0, n_near,
         (const uint32_t*)S.near_row.ptr, (const uint32_t*)S.near_col.ptr, (const uint32_t*)start.ptr,
         (const uint32_t*)end.ptr, (const uint32_t*)S.leaf_first.ptr, (const uint32_t*)S.leaf_end.ptr,
         (const uint32_t*)S.range_ptr.ptr, S.cursor.ptr, S.range_key.ptr, S.range_end.ptr);
  if (T.ranges)
    sort_pairs<uint64_t>(ctx, S.temp, S.range_key.ptr, S.range_key_s.ptr, S.range_end.ptr, S.range_end_s.ptr, T.ranges,
                         32 + bits_for((uint64_t)nl));
  cols.reserve(std::max<uint32_t>(T.cols, 1));
  if (T.ranges) {
    S.range_len.reserve(T.ranges);
    S.range_off.reserve(T.ranges);
    launch(ctx, "blk_range_len", blk_range_len_kernel, dim3(blocks_for(T.ranges, 256)), dim3(256), 0, T.ranges,
           (const uint64_t*)S.range_key_s.ptr, (const uint32_t*)S.range_end_s.ptr, S.range_len.ptr);
    scan_exclusive<uint32_t>(ctx, S.temp, S.range_len.ptr, S.range_off.ptr, T.ranges);
    launch(ctx, "blk_cols", blk_cols_kernel, dim3(blocks_for(T.ranges, 256)), dim3(256), 0, T.ranges,
           (const uint64_t*)S.range_key_s.ptr, (const uint32_t*)S.range_end_s.ptr, (const uint32_t*)S.range_off.ptr,
           cols.ptr);
  }
  panel_leaf.reserve(std::max(np, 1));
  panel_row0.reserve(std::max(np, 1));
  panel_rows.reserve(std::max(np, 1));
  panel_val.reserve(std::max(np, 1));
  seg_panel.reserve(std::max(nseg, 1));
  seg_col0.reserve(std::max(nseg, 1));
  seg_ptr.reserve((size_t)np + 1);
  launch(ctx, "blk_panel_fill", blk_panel_fill_kernel, dim3(blocks_for(nl, 256)), dim3(256), 0, (uint32_t)nl,
         (const uint32_t*)leaves.ptr, (const uint32_t*)start.ptr, (const uint32_t*)end.ptr, (const uint32_t*)col_ptr.ptr,
         (const uint32_t*)S.pan_ptr.ptr, (const uint32_t*)S.seg_base.ptr, (const uint64_t*)S.val_base.ptr, seg_cols,
         PanelOut{panel_leaf.ptr, panel_row0.ptr, panel_rows.ptr, panel_val.ptr, seg_panel.ptr, seg_col0.ptr,
                  seg_ptr.ptr});
  seg_arrive.reserve(std::max(np, 1));
  HIPCHECK(hipMemsetAsync(seg_arrive.ptr, 0, (size_t)std::max(np, 1) * sizeof(uint32_t), ctx->stream));
  reserve_values(ctx);
  far_coef.reserve((size_t)std::max<uint32_t>(n_far, 1) * nk);
  counters.reserve(nn);
  HIPCHECK(hipMemsetAsync(counters.ptr, 0, (size_t)nn * sizeof(uint32_t), ctx->stream));
  mark("panels");
  BlockStats st;
  st.far_blocks = n_far;
  st.near_blocks = n_near;
  st.admissible_pairs = T.sums[0];
  st.near_total = T.sums[2];
  st.near_entries = T.sums[3];
  st.depth = depth;
  return st;
}

void BlockTreeDev::set_near_lists(ddgrs_ctx* ctx, const BlockStructure& s) {
  n_near = (uint32_t)s.near_col.size();
  ddgrs::upload(ctx, scratch.near_ptr, s.near_ptr);
  ddgrs::upload(ctx, scratch.near_col, s.near_col);
}

void BlockTreeDev::lists(ddgrs_ctx* ctx, std::vector<uint32_t>* order_h, std::vector<uint32_t>* start_h,
                         std::vector<uint32_t>* end_h, std::vector<uint32_t>* right_h,
                         std::vector<uint32_t>* far_ptr_h, std::vector<uint32_t>* far_col_h,
                         std::vector<uint32_t>* near_ptr_h, std::vector<uint32_t>* near_col_h) {
  auto get = [&](std::vector<uint32_t>* h, const DevBuf<uint32_t>& d, size_t count) {
    h->assign(count, 0u);
    download(ctx, h->data(), (const uint32_t*)d.ptr, count);
  };
  get(order_h, order, (size_t)n);
  get(start_h, start, (size_t)nn);
  get(end_h, end, (size_t)nn);
  get(right_h, right, (size_t)nn);
  get(far_ptr_h, far_ptr, (size_t)nn + 1);
  get(far_col_h, far_col, n_far);
  get(near_ptr_h, scratch.near_ptr, (size_t)nn + 1);
  get(near_col_h, scratch.near_col, n_near);
  ctx->sync();
}

}  // namespace ddgrs
