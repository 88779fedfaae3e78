// Barnes–Hut interaction lists built on the device (RS/src/energy/barnes_hut.rs:743 build_lists).
//
// For each leaf L the reference walks the tree from the root depth first, left child first: a node
// A is lumped when max(r_L, r_A)² < θ² dist²(box_L, box_A), an exact leaf when it is a leaf, and
// opened otherwise.  With the nodes numbered in pre-order the walk visits nodes in increasing
// order, so it is done here without a stack: after a lumped node or a leaf the walk continues at
// the node after A's subtree (its escape index), after an opened node at A + 1.
//
// A leaf's walk is split into tasks, one per root of a cut of the tree (the nodes at depth D and
// the shallower leaves, in pre-order; a deeper cut for the few leaves much wider than the median):
// a task first tests the ancestors of its root from the top;
// if one of them is lumped or a leaf, the walk never enters the root's subtree and the task emits
// that ancestor only if its root is the first cut node below it (reached from it by left children
// alone); otherwise the task walks its root's subtree.  The tasks of a leaf, in pre-order of their
// roots, emit exactly the walk's sequence, so counts scanned in (leaf, task) order place every
// entry.  One thread per (leaf, task) walks twice (counts, then the entries); the transposes are
// stable radix sorts of the entries by target.  The admissibility test uses the host's FP64
// operations without contraction, so the lists are the host build's.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "bh.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

namespace {
// A node as the walks read it: box, radius, right child and escape index in one 64-byte record
// (one cache line per visited node).
struct alignas(64) BhNode {
  double lo[3], hi[3], radius;
  uint32_t right, esc;
};
static_assert(sizeof(BhNode) == 64, "one cache line per node");

struct BhTree {
  const BhNode* node;
  const uint32_t *start, *end, *leaves;
  uint32_t nn;
  double th2;
};

// The walking leaf's box and radius.
struct LeafBox {
  double lo[3], hi[3], radius;
};

__device__ __forceinline__ LeafBox leaf_box(const BhTree& T, uint32_t l) {
  const BhNode& n = T.node[l];
  return {{n.lo[0], n.lo[1], n.lo[2]}, {n.hi[0], n.hi[1], n.hi[2]}, n.radius};
}

__device__ __forceinline__ bool bh_lumped(const BhTree& T, const LeafBox& L, const BhNode& A) {
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double d = fmax(fmax(L.lo[k], A.lo[k]) - fmin(L.hi[k], A.hi[k]), 0.0);
    d2 += d * d;
  }
  const double h = fmax(L.radius, A.radius);
  return h * h < T.th2 * d2;
}
}  // namespace

// Two cuts: at depth kCutDepth for most leaves, and at kLongCutDepth for leaves whose radius exceeds
// kLongLeafRadius times the median leaf radius (the long thin leaves along mesh seams, whose walks
// open most of the tree): their walks are split into up to 2^kLongCutDepth tasks.
constexpr int kCutDepth = 5;
constexpr int kLongCutDepth = 9;
constexpr double kLongLeafRadius = 3.0;

struct BhTasks {
  const uint32_t* root;      // per cut node (the short cut's, then the long cut's): the node
  const uint32_t* depth;     // per cut node: its depth
  const uint32_t* parent;
  const uint32_t* task_ptr;  // per leaf ordinal: its first task (nl + 1)
  const uint32_t* cut_base;  // per leaf ordinal: its cut's first entry in root/depth
  uint32_t nl;
};

// The leaf ordinal k of task q (task_ptr[k] ≤ q < task_ptr[k + 1]) and the task's cut node index.
__device__ __forceinline__ void bh_task_of(const BhTasks& W, uint32_t q, uint32_t& k, uint32_t& j) {
  uint32_t lo = 0, hi = W.nl;  // the last k in [0, nl) with task_ptr[k] ≤ q
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (W.task_ptr[mid] <= q)
      lo = mid;
    else
      hi = mid;
  }
  k = lo;
  j = W.cut_base[k] + (q - W.task_ptr[k]);
}

// Walks task j (a cut node index) of leaf k, calling emit(node, lumped) for each entry in order.
template <class Emit>
__device__ __forceinline__ void bh_task_walk(const BhTree& T, const BhTasks& W, uint32_t k, uint32_t j, Emit&& emit) {
  const uint32_t l = T.leaves[k], c = W.root[j], m = W.depth[j];
  const LeafBox L = leaf_box(T, l);
  uint32_t path[kLongCutDepth];
  {
    uint32_t x = c;
    for (int i = (int)m - 1; i >= 0; --i) {
      x = W.parent[x];
      path[i] = x;
    }
  }
  for (uint32_t i = 0; i < m; ++i) {
    const uint32_t a = path[i];
    const BhNode A = T.node[a];
    const bool lum = bh_lumped(T, L, A);
    if (lum || A.right == 0) {
      if (c == a + (m - i)) emit(a, lum);
      return;
    }
  }
  const uint32_t stop = T.node[c].esc;
  for (uint32_t a = c; a < stop;) {
    const BhNode A = T.node[a];
    if (bh_lumped(T, L, A)) {
      emit(a, true);
      a = A.esc;
    } else if (A.right == 0) {
      emit(a, false);
      a = A.esc;
    } else {
      a = a + 1;
    }
  }
}

// Per (leaf, task) q = k·tasks + j < nl·tasks: lumped and exact counts, and Σ over exact nodes of
// the ordered pairs; zeros at q = nl·tasks (the scans end with the totals).
__global__ __launch_bounds__(256) void bhl_count_kernel(BhTree T, BhTasks W, uint32_t nl, uint32_t* __restrict__ lcount,
                                                        uint32_t* __restrict__ ecount,
                                                        unsigned long long* __restrict__ pairs) {
  __shared__ unsigned long long red[256];
  const uint32_t q = blockIdx.x * 256 + threadIdx.x, nq = W.task_ptr[nl];
  unsigned long long pk = 0;
  if (q < nq) {
    uint32_t k, j;
    bh_task_of(W, q, k, j);
    const uint32_t l = T.leaves[k];
    const uint64_t ns = T.end[l] - T.start[l];
    uint32_t nlum = 0, nex = 0;
    bh_task_walk(T, W, k, j, [&](uint32_t a, bool lum) {
      if (lum) {
        ++nlum;
      } else {
        ++nex;
        pk += a == l ? ns * (ns - 1) : ns * (uint64_t)(T.end[a] - T.start[a]);
      }
    });
    lcount[q] = nlum;
    ecount[q] = nex;
  } else if (q == nq) {
    lcount[q] = ecount[q] = 0;
  }
  red[threadIdx.x] = pk;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(pairs, red[0]);  // an integer sum
}

// The entries of each (leaf, task), with each exact entry's target leaf ordinal and source leaf
// ordinal and the lumped entries' indices; per leaf the offsets of its first task.
__global__ __launch_bounds__(256) void bhl_fill_kernel(BhTree T, BhTasks W, uint32_t nl,
                                                       const uint32_t* __restrict__ leaf_ordinal,
                                                       const uint32_t* __restrict__ ltask_off,
                                                       const uint32_t* __restrict__ etask_off,
                                                       uint32_t* __restrict__ lumped_off, uint32_t* __restrict__ exact_off,
                                                       uint32_t* __restrict__ lumped, uint32_t* __restrict__ lumped_e,
                                                       uint32_t* __restrict__ exact, uint32_t* __restrict__ exact_tgt,
                                                       uint32_t* __restrict__ exact_src) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x, nq = W.task_ptr[nl];
  if (q > nq) return;
  if (q == nq) {  // the totals
    lumped_off[nl] = ltask_off[q];
    exact_off[nl] = etask_off[q];
    return;
  }
  uint32_t k, j;
  bh_task_of(W, q, k, j);
  if (q == W.task_ptr[k]) {  // the leaf's first task
    lumped_off[k] = ltask_off[q];
    exact_off[k] = etask_off[q];
  }
  uint32_t ql = ltask_off[q], qe = etask_off[q];
  bh_task_walk(T, W, k, j, [&](uint32_t a, bool lum) {
    if (lum) {
      lumped[ql] = a;
      lumped_e[ql] = ql;
      ++ql;
    } else {
      exact[qe] = a;
      exact_tgt[qe] = leaf_ordinal[a];
      exact_src[qe] = k;
      ++qe;
    }
  });
}

// ptr[x] = first position of key x in the sorted keys (x in [0, m]).
__global__ __launch_bounds__(256) void bhl_ptr_kernel(uint32_t m, uint32_t count, const uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ ptr) {
  const uint32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x > m) return;
  uint32_t lo = 0, hi = count;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  ptr[x] = lo;
}

namespace {
int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 32 && (1ull << b) <= maxval) ++b;
  return b;
}
}  // namespace

void BhEval::build_lists_device(ddgrs_ctx* ctx) {
  const Topology& t = topo;
  // escape indices: the node after each subtree in pre-order
  std::vector<uint32_t> esc(nn);
  esc[0] = (uint32_t)nn;
  for (int a = 0; a < nn; ++a)
    if (t.right[a] != 0) {
      esc[a + 1] = t.right[a];
      esc[t.right[a]] = esc[a];
    }
  // the cuts: nodes at depth D and shallower leaves, in pre-order (the short cut, then the long one)
  std::vector<uint32_t> troot, tdepth;
  uint32_t ncut[2];
  for (int c = 0; c < 2; ++c) {
    const uint32_t D = std::min<uint32_t>(c ? kLongCutDepth : kCutDepth, t.height > 0 ? t.height - 1 : 0);
    const size_t n0 = troot.size();
    for (int a = 0; a < nn; ++a)
      if (t.depth[a] == D || (t.depth[a] < D && t.right[a] == 0)) {
        troot.push_back((uint32_t)a);
        tdepth.push_back(t.depth[a]);
      }
    ncut[c] = (uint32_t)(troot.size() - n0);
  }
  // per leaf its tasks: the long cut for leaves much wider than the median leaf
  std::vector<double> lr(nl);
  for (int k = 0; k < nl; ++k) lr[k] = clusters.radius[t.leaves[k]];
  std::vector<double> srt = lr;
  std::nth_element(srt.begin(), srt.begin() + nl / 2, srt.end());
  const double long_r = kLongLeafRadius * (nl ? srt[nl / 2] : 0.0);
  std::vector<uint32_t> task_ptr(nl + 1, 0), cut_base(nl);
  for (int k = 0; k < nl; ++k) {
    const bool wide = lr[k] > long_r;
    cut_base[k] = wide ? ncut[0] : 0;
    const uint64_t next = (uint64_t)task_ptr[k] + (wide ? ncut[1] : ncut[0]);
    if (next >= UINT32_MAX) throw Error(DDGRS_GPU, "bh_lists", "too many (leaf, task) pairs");
    task_ptr[k + 1] = (uint32_t)next;
  }
  // pinned staging: a copy from pageable memory would wait for the stream
  BhNode* nodes = (BhNode*)ctx->arena_take((size_t)nn * sizeof(BhNode));
  for (int a = 0; a < nn; ++a) {
    BhNode& r = nodes[a];
    for (int k = 0; k < 3; ++k) {
      r.lo[k] = clusters.lo[3 * (size_t)a + k];
      r.hi[k] = clusters.hi[3 * (size_t)a + k];
    }
    r.radius = clusters.radius[a];
    r.right = t.right[a];
    r.esc = esc[a];
  }
  d_bhnode.reserve((size_t)nn * sizeof(BhNode));
  HIPCHECK(hipMemcpyAsync(d_bhnode.ptr, nodes, (size_t)nn * sizeof(BhNode), hipMemcpyHostToDevice, ctx->stream));
  ddgrs::upload(ctx, d_leaf_ordinal, t.leaf_ordinal);
  ddgrs::upload(ctx, d_parent, t.parent);
  ddgrs::upload(ctx, d_task_root, troot);
  ddgrs::upload(ctx, d_task_depth, tdepth);
  ddgrs::upload(ctx, d_task_ptr, task_ptr);
  ddgrs::upload(ctx, d_cut_base, cut_base);
  const BhTree T{(const BhNode*)d_bhnode.ptr, d_start.ptr, d_end.ptr, d_leaves.ptr, (uint32_t)nn, theta * theta};
  const BhTasks W{d_task_root.ptr, d_task_depth.ptr, d_parent.ptr, d_task_ptr.ptr, d_cut_base.ptr, (uint32_t)nl};
  const size_t nq1 = (size_t)task_ptr[nl] + 1, nl1 = (size_t)nl + 1;
  l_count.reserve(nq1);
  e_count.reserve(nq1);
  l_task_off.reserve(nq1);
  e_task_off.reserve(nq1);
  d_lumped_off.reserve(nl1);
  d_exact_off.reserve(nl1);
  d_pairs.reserve(1);
  HIPCHECK(hipMemsetAsync(d_pairs.ptr, 0, sizeof(unsigned long long), ctx->stream));
  launch(ctx, "bhl_count", bhl_count_kernel, dim3(blocks_for((int64_t)nq1, 256)), dim3(256), 0, T, W, (uint32_t)nl,
         l_count.ptr, e_count.ptr, d_pairs.ptr);
  for (int which = 0; which < 2; ++which) {
    const uint32_t* in = which ? e_count.ptr : l_count.ptr;
    uint32_t* out = which ? e_task_off.ptr : l_task_off.ptr;
    size_t tb = 0;
    HIPCHECK(rocprim::exclusive_scan(nullptr, tb, in, out, 0u, nq1, rocprim::plus<uint32_t>(), ctx->stream));
    temp.reserve(tb);
    tb = temp.n;
    HIPCHECK(rocprim::exclusive_scan(temp.ptr, tb, in, out, 0u, nq1, rocprim::plus<uint32_t>(), ctx->stream));
  }
  struct Totals {
    uint32_t lumped, exact;
    unsigned long long pairs;
  };
  Totals* tot = (Totals*)ctx->staging(sizeof(Totals));
  download(ctx, &tot->lumped, (const uint32_t*)l_task_off.ptr + (nq1 - 1), 1);
  download(ctx, &tot->exact, (const uint32_t*)e_task_off.ptr + (nq1 - 1), 1);
  download(ctx, &tot->pairs, (const unsigned long long*)d_pairs.ptr, 1);
  ctx->sync();
  n_lumped = tot->lumped;
  n_exact = tot->exact;
  pairs_exact = tot->pairs;
  const size_t ml = std::max<size_t>(n_lumped, 1), me = std::max<size_t>(n_exact, 1);
  d_lumped.reserve(ml);
  l_idx.reserve(ml);
  l_key_s.reserve(ml);
  d_lumped_t.reserve(ml);
  d_exact.reserve(me);
  e_tgt.reserve(me);
  e_src.reserve(me);
  e_tgt_s.reserve(me);
  d_exact_t.reserve(me);
  d_lumped_t_off.reserve((size_t)nn + 1);
  d_exact_t_off.reserve(nl1);
  launch(ctx, "bhl_fill", bhl_fill_kernel, dim3(blocks_for((int64_t)nq1, 256)), dim3(256), 0, T, W, (uint32_t)nl,
         (const uint32_t*)d_leaf_ordinal.ptr, (const uint32_t*)l_task_off.ptr, (const uint32_t*)e_task_off.ptr,
         d_lumped_off.ptr, d_exact_off.ptr, d_lumped.ptr, l_idx.ptr, d_exact.ptr, e_tgt.ptr, e_src.ptr);
  // transposes: entries grouped by target, in entry order within a target
  auto sort = [&](const uint32_t* ki, uint32_t* ko, const uint32_t* vi, uint32_t* vo, size_t n, int bits) {
    size_t tb = 0;
    HIPCHECK(rocprim::radix_sort_pairs(nullptr, tb, ki, ko, vi, vo, n, 0, bits, ctx->stream));
    temp.reserve(tb);
    tb = temp.n;
    HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, ki, ko, vi, vo, n, 0, bits, ctx->stream));
  };
  if (n_lumped) sort(d_lumped.ptr, l_key_s.ptr, l_idx.ptr, d_lumped_t.ptr, n_lumped, bits_for((uint64_t)nn));
  launch(ctx, "bhl_ptr", bhl_ptr_kernel, dim3(blocks_for((int64_t)nn + 1, 256)), dim3(256), 0, (uint32_t)nn,
         (uint32_t)n_lumped, (const uint32_t*)l_key_s.ptr, d_lumped_t_off.ptr);
  if (n_exact) sort(e_tgt.ptr, e_tgt_s.ptr, e_src.ptr, d_exact_t.ptr, n_exact, bits_for((uint64_t)nl));
  launch(ctx, "bhl_ptr", bhl_ptr_kernel, dim3(blocks_for((int64_t)nl + 1, 256)), dim3(256), 0, (uint32_t)nl,
         (uint32_t)n_exact, (const uint32_t*)e_tgt_s.ptr, d_exact_t_off.ptr);
}

// ---- flattened exact lists ---------------------------------------------------------------------

// size[q] = faces of the node of entry q (own side: node exact[q]; partner side: leaf
// leaves[exact_t[q]]), 0 at q = count.
__global__ __launch_bounds__(256) void bhl_flat_size_kernel(uint32_t count, const uint32_t* __restrict__ node,
                                                            const uint32_t* __restrict__ leaves,
                                                            const uint32_t* __restrict__ start,
                                                            const uint32_t* __restrict__ end,
                                                            uint32_t* __restrict__ size) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q > count) return;
  if (q == count) {
    size[q] = 0;
    return;
  }
  const uint32_t a = leaves ? leaves[node[q]] : node[q];
  size[q] = end[a] - start[a];
}

__global__ __launch_bounds__(256) void bhl_flat_fill_kernel(uint32_t count, const uint32_t* __restrict__ node,
                                                            const uint32_t* __restrict__ leaves,
                                                            const uint32_t* __restrict__ start,
                                                            const uint32_t* __restrict__ end,
                                                            const uint32_t* __restrict__ fo, uint32_t* __restrict__ pos) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= count) return;
  const uint32_t a = leaves ? leaves[node[q]] : node[q];
  uint32_t o = fo[q];
  for (uint32_t p = start[a]; p < end[a]; ++p) pos[o++] = p;
}

// per leaf: the flat offset of its first entry (k = nl: the total)
__global__ __launch_bounds__(256) void bhl_flat_off_kernel(uint32_t nl, const uint32_t* __restrict__ entry_off,
                                                           const uint32_t* __restrict__ fo, uint32_t* __restrict__ off) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k <= nl) off[k] = fo[entry_off[k]];
}

// Work items per leaf ordinal k < nl (zero at k = nl, so the scan ends with the total): the leaf's
// steps (lumped, exact and partner entries, 8 a step) in pieces of kBhItemSteps, at least one.
__global__ __launch_bounds__(256) void bh_item_count_kernel(uint32_t nl, const uint32_t* __restrict__ lumped_off,
                                                            const uint32_t* __restrict__ ex_off,
                                                            const uint32_t* __restrict__ pt_off,
                                                            uint32_t* __restrict__ cnt) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k > nl) return;
  if (k == nl) {
    cnt[k] = 0;
    return;
  }
  const uint32_t steps = (lumped_off[k + 1] - lumped_off[k] + 7) / 8 + (ex_off[k + 1] - ex_off[k] + 7) / 8 +
                         (pt_off[k + 1] - pt_off[k] + 7) / 8;
  cnt[k] = max(1u, (steps + kBhItemSteps - 1) / kBhItemSteps);
}

__global__ __launch_bounds__(256) void bh_item_fill_kernel(uint32_t nl, const uint32_t* __restrict__ item_ptr,
                                                           uint32_t* __restrict__ item_leaf,
                                                           uint32_t* __restrict__ item_g0) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nl) return;
  for (uint32_t q = item_ptr[k], g = 0; q < item_ptr[k + 1]; ++q, g += kBhItemSteps) {
    item_leaf[q] = k;
    item_g0[q] = g;
  }
}

void BhEval::build_items(ddgrs_ctx* ctx) {
  const size_t nl1 = (size_t)nl + 1;
  item_cnt.reserve(nl1);
  item_ptr.reserve(nl1);
  launch(ctx, "bh_item_count", bh_item_count_kernel, dim3(blocks_for((int64_t)nl1, 256)), dim3(256), 0, (uint32_t)nl,
         (const uint32_t*)d_lumped_off.ptr, (const uint32_t*)d_ex_off.ptr, (const uint32_t*)d_pt_off.ptr,
         item_cnt.ptr);
  size_t tb = 0;
  HIPCHECK(rocprim::exclusive_scan(nullptr, tb, item_cnt.ptr, item_ptr.ptr, 0u, nl1, rocprim::plus<uint32_t>(),
                                   ctx->stream));
  temp.reserve(tb);
  tb = temp.n;
  HIPCHECK(rocprim::exclusive_scan(temp.ptr, tb, item_cnt.ptr, item_ptr.ptr, 0u, nl1, rocprim::plus<uint32_t>(),
                                   ctx->stream));
  uint32_t* pin = (uint32_t*)ctx->staging(sizeof(uint32_t));
  download(ctx, pin, (const uint32_t*)item_ptr.ptr + nl, 1);
  ctx->sync();
  n_items = (int)*pin;
  item_leaf.reserve(std::max(n_items, 1));
  item_g0.reserve(std::max(n_items, 1));
  launch(ctx, "bh_item_fill", bh_item_fill_kernel, dim3(blocks_for(nl, 256)), dim3(256), 0, (uint32_t)nl,
         (const uint32_t*)item_ptr.ptr, item_leaf.ptr, item_g0.ptr);
  part.reserve(7 * 8 * (size_t)std::max(n_items, 1));
  item_energy.reserve(std::max(n_items, 1));
}

void BhEval::flatten_lists(ddgrs_ctx* ctx) {
  const size_t nl1 = (size_t)nl + 1;
  d_ex_off.reserve(nl1);
  d_pt_off.reserve(nl1);
  for (int side = 0; side < 2; ++side) {
    const uint32_t count = (uint32_t)n_exact;
    const uint32_t* node = side ? d_exact_t.ptr : d_exact.ptr;
    const uint32_t* lv = side ? d_leaves.ptr : nullptr;
    const uint32_t* entry_off = side ? d_exact_t_off.ptr : d_exact_off.ptr;
    DevBuf<uint32_t>& size = side ? pt_size : ex_size;
    DevBuf<uint32_t>& fo = side ? pt_fo : ex_fo;
    DevBuf<uint32_t>& pos = side ? d_pt_pos : d_ex_pos;
    DevBuf<uint32_t>& off = side ? d_pt_off : d_ex_off;
    size.reserve((size_t)count + 1);
    fo.reserve((size_t)count + 1);
    launch(ctx, "bhl_flat_size", bhl_flat_size_kernel, dim3(blocks_for((int64_t)count + 1, 256)), dim3(256), 0, count,
           node, lv, (const uint32_t*)d_start.ptr, (const uint32_t*)d_end.ptr, size.ptr);
    size_t tb = 0;
    HIPCHECK(rocprim::exclusive_scan(nullptr, tb, size.ptr, fo.ptr, 0u, (size_t)count + 1, rocprim::plus<uint32_t>(),
                                     ctx->stream));
    temp.reserve(tb);
    tb = temp.n;
    HIPCHECK(rocprim::exclusive_scan(temp.ptr, tb, size.ptr, fo.ptr, 0u, (size_t)count + 1, rocprim::plus<uint32_t>(),
                                     ctx->stream));
    // the total, to size the positions
    uint32_t* pin = (uint32_t*)ctx->staging(sizeof(uint32_t));
    download(ctx, pin, (const uint32_t*)fo.ptr + count, 1);
    ctx->sync();
    const uint32_t total = *pin;
    pos.reserve(std::max<uint32_t>(total, 1));
    launch(ctx, "bhl_flat_fill", bhl_flat_fill_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, count, node, lv,
           (const uint32_t*)d_start.ptr, (const uint32_t*)d_end.ptr, (const uint32_t*)fo.ptr, pos.ptr);
    launch(ctx, "bhl_flat_off", bhl_flat_off_kernel, dim3(blocks_for((int64_t)nl1, 256)), dim3(256), 0, (uint32_t)nl,
           entry_off, (const uint32_t*)fo.ptr, off.ptr);
  }
}

}  // namespace ddgrs
