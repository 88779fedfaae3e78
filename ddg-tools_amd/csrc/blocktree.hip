// Hierarchical products on the device in difference form.
//
// Follows RS/src/spatial/block_tree.rs:536 multiply_core (far field per admissible block, near
// blocks of the chain root → leaf, diagonal excluded) and RS/src/metric/hierarchical.rs
// (FaceKernels :192-236, product :696: m = M x, ψ = a m, z = w (H̃a·m − H̃ψ),
// y = 2 Σ_{S∋v} a_S (Mᵀ row)·z_S; VertexFractionalTree::product :1313).  The products are evaluated
// as Σ_T v_ST (u_S − u_T) (see blocktree.hpp), which is the same operator.
//
// Precision: inputs u are FP64 split into FP32 hi/lo pairs, so u_S − u_T = (hi_S − hi_T) +
// (lo_S − lo_T) keeps its relative accuracy however close u_S and u_T are; near values and far
// coefficients are evaluated in FP64 and stored FP32 (a fixed perturbation of the operator, the
// same for every product); near sums accumulate in FP32 per lane and reduce pairwise across the
// wave; node means, the far field and the epilogue are FP64; outputs are FP32 hi/lo pairs.
//
// Kernels of one product: upward node sums (one thread per leaf, climbing with a per-node arrival
// counter: the second child to arrive sums left + right, so the values do not depend on timing),
// far field (one wave per node), leaf far terms (one thread per leaf and channel), panels (one wave
// per panel: lanes are columns, registers hold the panel's rows × C channels).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "blocktree.hpp"
#include "host_parallel.hpp"

namespace ddgrs {

__device__ __forceinline__ float wsum64(float x) {
  x += __shfl_xor(x, 32);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 1);
  return x;
}

// Sums N per-lane values over the 64 lanes of a wave, the work halved at each step: across lane
// bit MASK a lane keeps one half of its values and adds the partner's copy of that half (the
// partner keeps the other), until N is odd; the rest is summed by a butterfly.  Afterwards lane L
// holds the sums of values [base, base + n) in v[0, n), with n and base from rs_slots.
template <int N, int MASK, class T>
__device__ __forceinline__ void reduce_scatter(T* v, int lane) {
  if constexpr (MASK >= 1) {
    if constexpr (N % 2 == 0) {
      const bool up = (lane & MASK) != 0;
#pragma unroll
      for (int i = 0; i < N / 2; ++i) {
        const T keep = up ? v[N / 2 + i] : v[i];
        const T send = up ? v[i] : v[N / 2 + i];
        v[i] = keep + __shfl_xor(send, MASK);
      }
      reduce_scatter<N / 2, MASK / 2, T>(v, lane);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int m = MASK; m >= 1; m >>= 1) v[i] += __shfl_xor(v[i], m);
    }
  }
}
// The values that lane L holds after reduce_scatter<N, 32>: count n and first index base.
template <int N, int MASK>
__device__ __forceinline__ void rs_slots(int lane, int& n, int& base) {
  if constexpr (MASK >= 1 && N % 2 == 0) {
    rs_slots<N / 2, MASK / 2>(lane, n, base);
    if (lane & MASK) base += N / 2;
  } else {
    n = N;
    base = 0;
  }
}
__device__ __forceinline__ double wsum64d(double x) {
  x += __shfl_xor(x, 32);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 1);
  return x;
}

__device__ __forceinline__ uint32_t wave_id() {
  return __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
}

// 2C floats (C hi then C lo) of one element.
template <int C>
__device__ __forceinline__ void load_hl(const float* __restrict__ p, float* hi, float* lo) {
  if constexpr (C % 2 == 0) {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int i = 0; i < C / 2; ++i) {
      float4 t = q[i];
      float* dst = (4 * i < C) ? hi : lo;
      const int o = (4 * i < C) ? 4 * i : 4 * i - C;
      dst[o] = t.x;
      dst[o + 1] = t.y;
      dst[o + 2] = t.z;
      dst[o + 3] = t.w;
    }
  } else {
    const float2* q = reinterpret_cast<const float2*>(p);
    float t[2 * C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
      float2 u = q[i];
      t[2 * i] = u.x;
      t[2 * i + 1] = u.y;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      hi[c] = t[c];
      lo[c] = t[C + c];
    }
  }
}

// ---- one product ---------------------------------------------------------------------------

struct TreeArgs {
  const uint32_t *start, *end, *right, *parent, *leaves;
  const double *weight, *inv_mass;
  uint32_t* counters;
  int nl;
};

// Node sums σ_A = Σ_{T∈A} a_T u_T and means ū_A = σ_A / A_A (FP64), all nodes.
template <int C>
__global__ __launch_bounds__(256) void bt_upward_kernel(TreeArgs t, const float* __restrict__ u_hl,
                                                        double* __restrict__ sigma, double* __restrict__ mean) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= t.nl) return;
  uint32_t node = t.leaves[k];
  double s[C];
#pragma unroll
  for (int c = 0; c < C; ++c) s[c] = 0.0;
  for (uint32_t p = t.start[node]; p < t.end[node]; ++p) {
    const double a = t.weight[p];
    const float* q = u_hl + (size_t)p * 2 * C;
#pragma unroll
    for (int c = 0; c < C; ++c) s[c] += a * ((double)q[c] + (double)q[C + c]);
  }
  for (;;) {
    const double im = t.inv_mass[node];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      publish(&sigma[(size_t)node * C + c], s[c]);
      mean[(size_t)node * C + c] = s[c] * im;
    }
    if (node == 0) return;
    const uint32_t p = t.parent[node];
    published();  // σ_node is visible before the arrival at the parent
    if (atomicAdd(&t.counters[p], 1u) == 0u) return;  // the sibling is not done: it will climb
    t.counters[p] = 0u;
    const uint32_t l = p + 1, r = t.right[p];
#pragma unroll
    for (int c = 0; c < C; ++c) s[c] = collect(&sigma[(size_t)l * C + c]) + collect(&sigma[(size_t)r * C + c]);
    node = p;
  }
}

// γ_A = Σ_{B far of A} c_AB (ū_A − ū_B) per channel (FP64), one wave per node.
template <int C, int H, int NK>
__global__ __launch_bounds__(256) void bt_far_kernel(int nn, const uint32_t* __restrict__ far_ptr,
                                                     const uint32_t* __restrict__ far_col, const float* __restrict__ coef,
                                                     const double* __restrict__ mean, double* __restrict__ gamma) {
  const uint32_t a = wave_id();
  const int lane = threadIdx.x & 63;
  if (a >= (uint32_t)nn) return;
  const uint32_t q0 = far_ptr[a], q1 = far_ptr[a + 1];
  double ma[C], g[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    ma[c] = mean[(size_t)a * C + c];
    g[c] = 0.0;
  }
  for (uint32_t q = q0 + lane; q < q1; q += 64) {
    const uint32_t b = far_col[q];
    const double k0 = (double)coef[(size_t)q * NK], k1 = (double)coef[(size_t)q * NK + NK - 1];
#pragma unroll
    for (int c = 0; c < C; ++c) g[c] += (c < H ? k0 : k1) * (ma[c] - mean[(size_t)b * C + c]);
  }
  if (q1 - q0 > 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) g[c] = wsum64d(g[c]);
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) gamma[(size_t)a * C + c] = g[c];
  }
}

// δ_L = Σ_{A ⊇ L} [α_A (ū_L − ū_A) + γ_A], from the leaf to the root (FP64).
template <int C, int H, int NK>
__global__ __launch_bounds__(256) void bt_leaf_far_kernel(TreeArgs t, const double* __restrict__ alpha_node,
                                                          const double* __restrict__ mean,
                                                          const double* __restrict__ gamma, double* __restrict__ delta) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.nl * C) return;
  const int k = i / C, c = i % C;
  const int kern = c < H ? 0 : NK - 1;
  const uint32_t leaf = t.leaves[k];
  const double ml = mean[(size_t)leaf * C + c];
  double d = 0.0;
  for (uint32_t a = leaf;; a = t.parent[a]) {
    d += alpha_node[(size_t)a * NK + kern] * (ml - mean[(size_t)a * C + c]) + gamma[(size_t)a * C + c];
    if (a == 0) break;
  }
  delta[(size_t)k * C + c] = d;
}

struct PanelArgs {
  const uint32_t *panel_leaf, *panel_row0, *panel_rows, *col_ptr, *cols, *leaves;
  const uint64_t* panel_val;
  const float* val;
  const double *alpha_leaf, *mean, *delta;
  int np;
  const float* vzero;  // NK zeros after the values
};

// Panel segments: per segment its panel and first column, per panel its first segment.
struct SegArgs {
  const uint32_t *panel, *col0, *ptr;
  uint32_t* arrive;
  float* partial;     // FP32 products
  double* partial64;  // extended products
  uint32_t cols;
  int n;
};

// z of row r, channel c of panel p from its near sum: z = w_k (near + ᾱ_L (u_S − ū_L) + δ_L).
template <int C, int H, int NK>
__device__ __forceinline__ void panel_finish(const PanelArgs& P, uint32_t k, uint32_t row0, int r, int c, double near,
                                             const float* __restrict__ u_hl, double w0, double w1,
                                             float* __restrict__ z_hl) {
  const int kern = c < H ? 0 : NK - 1;
  const uint32_t srow = row0 + r;
  const float* qv = u_hl + (size_t)srow * 2 * C;
  const double us = (double)qv[c] + (double)qv[C + c];
  const double far =
      P.alpha_leaf[(size_t)k * NK + kern] * (us - P.mean[(size_t)P.leaves[k] * C + c]) + P.delta[(size_t)k * C + c];
  const double z = (c < H ? w0 : w1) * (near + far);
  const float zh = (float)z;
  z_hl[(size_t)srow * 2 * C + c] = zh;
  z_hl[(size_t)srow * 2 * C + C + c] = (float)(z - (double)zh);
}

// One 64-column chunk of a panel segment in registers: the columns' u (hi, lo) and the rows' values.
template <int C, int NK>
struct PanelChunk {
  float th[C], tl[C], v0[kPanelRows], v1[kPanelRows];
};

// Lanes past the segment's end and rows past the panel's read the zero entry after the values
// (`vzero`), so no load result is masked and the loads stay in flight until their chunk is summed.
template <int C, int NK>
__device__ __forceinline__ void panel_load(PanelChunk<C, NK>& x, const float* __restrict__ u_hl, uint32_t tcol,
                                           const float* __restrict__ vbase, const float* __restrict__ vzero,
                                           uint32_t K, uint32_t kk, bool act, uint32_t R) {
  load_hl<C>(u_hl + (size_t)tcol * 2 * C, x.th, x.tl);
#pragma unroll
  for (int r = 0; r < (int)kPanelRows; ++r) {
    const bool ok = act && (uint32_t)r < R;
    const float* v = ok ? vbase + ((size_t)r * K + kk) * NK : vzero;
    x.v0[r] = v[0];
    x.v1[r] = v[NK - 1];
  }
}

// Acc float: the difference and the sum in FP32.  Acc double (the extended product): the
// difference of the hi/lo pairs exactly and the sum in FP64, so that the product's rounding is
// that of the FP32 values alone.
template <int C, int H, int NK, class Acc>
__device__ __forceinline__ void panel_compute(Acc (&acc)[kPanelRows][C], const PanelChunk<C, NK>& x,
                                              const float* __restrict__ u_hl, uint32_t row0, uint32_t R) {
#pragma unroll
  for (int r = 0; r < (int)kPanelRows; ++r) {
    if ((uint32_t)r < R) {
      const float* srow = u_hl + (size_t)(row0 + r) * 2 * C;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if constexpr (sizeof(Acc) == sizeof(float)) {
          const float d = (srow[c] - x.th[c]) + (srow[C + c] - x.tl[c]);
          acc[r][c] += (c < H ? x.v0[r] : x.v1[r]) * d;
        } else {
          const double d = ((double)srow[c] - (double)x.th[c]) + ((double)srow[C + c] - (double)x.tl[c]);
          acc[r][c] += (double)(c < H ? x.v0[r] : x.v1[r]) * d;
        }
      }
    }
  }
}

// Near field of one panel segment per wave.  A panel of one segment is finished by its wave; the
// segments of a longer panel store their sums, and the last to arrive finishes the panel from the
// stored sums in segment order (the result does not depend on the arrival order).  The chunks go
// through two register sets: the loads of the next chunk are in flight while one is summed.
template <int C, int H, int NK, class Acc>
__global__ __launch_bounds__(256) void bt_panel_kernel(PanelArgs P, SegArgs G, const float* __restrict__ u_hl,
                                                       double w0, double w1, float* __restrict__ z_hl) {
  const uint32_t sg = xcd_wave_id();
  const int lane = threadIdx.x & 63;
  if (sg >= (uint32_t)G.n) return;
  const uint32_t p = G.panel[sg], j0 = G.col0[sg];
  const uint32_t k = P.panel_leaf[p], row0 = P.panel_row0[p], R = P.panel_rows[p];
  const uint32_t c0 = P.col_ptr[k], K = P.col_ptr[k + 1] - c0;
  const uint32_t kend = min(K, j0 + G.cols);
  const uint32_t nch = (kend - j0 + 63) / 64;  // chunks of the segment (at least one)
  const float* vbase = P.val + P.panel_val[p] * NK;
  Acc acc[kPanelRows][C];
#pragma unroll
  for (int r = 0; r < (int)kPanelRows; ++r)
#pragma unroll
    for (int c = 0; c < C; ++c) acc[r][c] = 0.f;
  // Chunk m's batch of loads: the column index of chunk m + 1, then chunk m's u and values (column
  // indices clamped to the segment's last column, so every load is in bounds; every panel has a
  // column).  A batch is issued before the previous chunk is summed; the loop ends after one batch
  // past the segment's last chunk (its results unused).
  const uint32_t last = c0 + kend - 1;
  uint32_t cia = P.cols[min(c0 + j0 + lane, last)], cib;
  PanelChunk<C, NK> a, b;
  for (uint32_t m = 0;; m += 2) {
    const uint32_t km = j0 + 64 * m + lane;
    cib = P.cols[min(c0 + km + 64, last)];
    panel_load<C, NK>(a, u_hl, cia, vbase, P.vzero, K, km, km < kend, R);
    if (m > 0) panel_compute<C, H, NK, Acc>(acc, b, u_hl, row0, R);
    if (m >= nch) break;
    cia = P.cols[min(c0 + km + 128, last)];
    panel_load<C, NK>(b, u_hl, cib, vbase, P.vzero, K, km + 64, km + 64 < kend, R);
    panel_compute<C, H, NK, Acc>(acc, a, u_hl, row0, R);
    if (m + 1 >= nch) break;
  }
  // the sums over the wave: reduce-scatter, then each lane whose low bits are 0 finishes its values
  constexpr int NV = (int)kPanelRows * C;
  Acc* flat = &acc[0][0];
  reduce_scatter<NV, 32, Acc>(flat, lane);
  int nmine, base;
  rs_slots<NV, 32>(lane, nmine, base);
  // lanes that differ only in the bits the butterfly summed over hold the same sums: the lowest writes
  int used = 0;
  for (int m = 32, nn = NV; m >= 1 && nn % 2 == 0; m >>= 1, nn /= 2) used |= m;
  const bool owner = (lane & ~used & 63) == 0;
  const uint32_t s0 = G.ptr[p], ns = G.ptr[p + 1] - s0;
  if (ns == 1) {
    if (owner) {
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        if (q >= nmine) break;
        const int idx = base + q;
        if ((uint32_t)(idx / C) < R) panel_finish<C, H, NK>(P, k, row0, idx / C, idx % C, (double)flat[q], u_hl, w0, w1, z_hl);
      }
    }
    return;
  }
  if (owner) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      if (q >= nmine) break;
      if constexpr (sizeof(Acc) == sizeof(float))
        publish(&G.partial[(size_t)sg * NV + base + q], flat[q]);
      else
        publish(&G.partial64[(size_t)sg * NV + base + q], flat[q]);
    }
  }
  published();  // the partial sums are visible before the arrival
  uint32_t arrived = 0;
  if (lane == 0) arrived = atomicAdd(&G.arrive[p], 1u);
  arrived = __shfl(arrived, 0);
  if (arrived != ns - 1) return;
  if (lane == 0) G.arrive[p] = 0u;
  for (int idx = lane; idx < NV; idx += 64) {
    if ((uint32_t)(idx / C) >= R) continue;
    double near = 0.0;
    for (uint32_t t = s0; t < s0 + ns; ++t) {
      if constexpr (sizeof(Acc) == sizeof(float))
        near += (double)collect(&G.partial[(size_t)t * NV + idx]);
      else
        near += collect(&G.partial64[(size_t)t * NV + idx]);
    }
    panel_finish<C, H, NK>(P, k, row0, idx / C, idx % C, near, u_hl, w0, w1, z_hl);
  }
}

// ---- structure -----------------------------------------------------------------------------

// α_A = Σ_B c_AB per kernel (FP64), one wave per node.
__global__ __launch_bounds__(256) void bt_alpha_node_kernel(int nn, int nk, const uint32_t* __restrict__ far_ptr,
                                                            const float* __restrict__ coef, double* __restrict__ alpha) {
  const uint32_t a = wave_id();
  const int lane = threadIdx.x & 63;
  if (a >= (uint32_t)nn) return;
  for (int k = 0; k < nk; ++k) {
    double s = 0.0;
    for (uint32_t q = far_ptr[a] + lane; q < far_ptr[a + 1]; q += 64) s += (double)coef[(size_t)q * nk + k];
    s = wsum64d(s);
    if (lane == 0) alpha[(size_t)a * nk + k] = s;
  }
}

// ᾱ_L = Σ_{A ⊇ L} α_A (leaf to root), one thread per leaf.
__global__ __launch_bounds__(256) void bt_alpha_leaf_kernel(int nl, int nk, const uint32_t* __restrict__ leaves,
                                                            const uint32_t* __restrict__ parent,
                                                            const double* __restrict__ alpha,
                                                            double* __restrict__ alpha_leaf) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nl) return;
  for (int j = 0; j < nk; ++j) {
    double s = 0.0;
    for (uint32_t a = leaves[k];; a = parent[a]) {
      s += alpha[(size_t)a * nk + j];
      if (a == 0) break;
    }
    alpha_leaf[(size_t)k * nk + j] = s;
  }
}

uint32_t panel_segment_cols() {
  const char* e = std::getenv("DDGRS_SEGMENT_COLUMNS");
  if (!e || !*e) return kSegCols;
  const long v = std::strtol(e, nullptr, 10);
  if (v <= 0 || v % 64 != 0 || v > (long)kSegCols)
    invalid("DDGRS_SEGMENT_COLUMNS", e, "a positive multiple of 64, at most " + std::to_string(kSegCols));
  return (uint32_t)v;
}

BlockTreePlan plan_block_tree(const Topology& t, const BlockStructure& s, int nk, std::vector<double> weight_bvh,
                              const std::vector<double>& node_mass) {
  constexpr uint32_t kRows = kPanelRows;
  BlockTreePlan P;
  P.nk = nk;
  P.n = (int)t.order.size();
  P.nn = (int)t.n_nodes();
  P.nl = (int)t.leaves.size();
  P.order = t.order;
  P.position.resize(P.n);
  for (int p = 0; p < P.n; ++p) P.position[t.order[p]] = (uint32_t)p;
  P.start = t.start;
  P.end = t.end;
  P.right = t.right;
  P.parent = t.parent;
  P.leaves = t.leaves;
  // panels: per leaf, the column ranges of the near blocks of its ancestors, ascending
  const int nl = P.nl;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> ranges(nl);
  parallel((size_t)nl, [&](size_t k) {
    auto& rg = ranges[k];
    for (uint32_t ci = s.chain_ptr[k]; ci < s.chain_ptr[k + 1]; ++ci) {
      const uint32_t a = s.chain[ci];
      for (uint32_t q = s.near_ptr[a]; q < s.near_ptr[a + 1]; ++q) {
        const uint32_t b = s.near_col[q];
        rg.push_back({t.start[b], t.end[b]});
      }
    }
    std::sort(rg.begin(), rg.end());
  });
  P.col_ptr.assign(nl + 1, 0);
  for (int k = 0; k < nl; ++k) {
    uint32_t K = 0;
    for (auto [b0, b1] : ranges[k]) K += b1 - b0;
    P.col_ptr[k + 1] = P.col_ptr[k] + K;
  }
  P.cols.resize(P.col_ptr[nl]);
  parallel((size_t)nl, [&](size_t k) {
    uint32_t o = P.col_ptr[k];
    for (auto [b0, b1] : ranges[k])
      for (uint32_t x = b0; x < b1; ++x) P.cols[o++] = x;
  });
  uint64_t total = 0;
  for (int k = 0; k < nl; ++k) {
    const uint32_t K = P.col_ptr[k + 1] - P.col_ptr[k], leaf = t.leaves[k];
    for (uint32_t r0 = t.start[leaf]; r0 < t.end[leaf]; r0 += kRows) {
      const uint32_t rows = std::min(kRows, t.end[leaf] - r0);
      P.panel_leaf.push_back((uint32_t)k);
      P.panel_row0.push_back(r0);
      P.panel_rows.push_back(rows);
      P.panel_val.push_back(total);
      total += (uint64_t)rows * K;
    }
  }
  P.np = (int)P.panel_leaf.size();
  P.near_total = total;
  P.seg_cols = panel_segment_cols();
  P.seg_ptr.assign(P.np + 1, 0);
  for (int p = 0; p < P.np; ++p) {
    const uint32_t K = P.col_ptr[P.panel_leaf[p] + 1] - P.col_ptr[P.panel_leaf[p]];
    const uint32_t ns = std::max<uint32_t>(1, (K + P.seg_cols - 1) / P.seg_cols);
    P.seg_ptr[p + 1] = P.seg_ptr[p] + ns;
    for (uint32_t j = 0; j < ns; ++j) {
      P.seg_panel.push_back((uint32_t)p);
      P.seg_col0.push_back(j * P.seg_cols);
    }
  }
  P.far_ptr = s.far_ptr;
  P.far_col = s.far_col;
  P.far_row.resize(s.far_col.size());
  for (int a = 0; a < P.nn; ++a)
    for (uint32_t q = s.far_ptr[a]; q < s.far_ptr[a + 1]; ++q) P.far_row[q] = (uint32_t)a;
  P.weight = std::move(weight_bvh);
  P.inv_mass.resize(P.nn);
  for (int a = 0; a < P.nn; ++a) P.inv_mass[a] = 1.0 / node_mass[a];
  return P;
}

void BlockTreeDev::upload(ddgrs_ctx* ctx, const BlockTreePlan& P) {
  nk = P.nk;
  n = P.n;
  nn = P.nn;
  nl = P.nl;
  np = P.np;
  near_total = P.near_total;
  n_far = (uint32_t)P.far_col.size();
  ddgrs::upload(ctx, order, P.order);
  ddgrs::upload(ctx, position, P.position);
  ddgrs::upload(ctx, start, P.start);
  ddgrs::upload(ctx, end, P.end);
  ddgrs::upload(ctx, right, P.right);
  ddgrs::upload(ctx, parent, P.parent);
  ddgrs::upload(ctx, leaves, P.leaves);
  ddgrs::upload(ctx, col_ptr, P.col_ptr);
  ddgrs::upload(ctx, cols, P.cols);
  ddgrs::upload(ctx, panel_leaf, P.panel_leaf);
  ddgrs::upload(ctx, panel_row0, P.panel_row0);
  ddgrs::upload(ctx, panel_rows, P.panel_rows);
  ddgrs::upload(ctx, panel_val, P.panel_val);
  seg_cols = P.seg_cols;
  nseg = (int)P.seg_panel.size();
  ddgrs::upload(ctx, seg_panel, P.seg_panel);
  ddgrs::upload(ctx, seg_col0, P.seg_col0);
  ddgrs::upload(ctx, seg_ptr, P.seg_ptr);
  seg_arrive.reserve(std::max(np, 1));
  HIPCHECK(hipMemsetAsync(seg_arrive.ptr, 0, (size_t)std::max(np, 1) * sizeof(uint32_t), ctx->stream));
  reserve_values(ctx);
  ddgrs::upload(ctx, far_ptr, P.far_ptr);
  ddgrs::upload(ctx, far_col, P.far_col);
  ddgrs::upload(ctx, far_row, P.far_row);
  far_coef.reserve((size_t)std::max<uint32_t>(n_far, 1) * nk);
  ddgrs::upload(ctx, weight, P.weight);
  ddgrs::upload(ctx, inv_mass, P.inv_mass);
  counters.reserve(nn);
  HIPCHECK(hipMemsetAsync(counters.ptr, 0, (size_t)nn * sizeof(uint32_t), ctx->stream));
}

void BlockTreeDev::reserve_values(ddgrs_ctx* ctx) {
  val.reserve((size_t)(near_total + 1) * nk);
  HIPCHECK(hipMemsetAsync(val.ptr + (size_t)near_total * nk, 0, (size_t)nk * sizeof(float), ctx->stream));
}

void BlockTreeDev::finish_far(ddgrs_ctx* ctx) {
  alpha_node.reserve((size_t)nn * nk);
  alpha_leaf.reserve((size_t)nl * nk);
  launch(ctx, "bt_alpha_node", bt_alpha_node_kernel, dim3(blocks_for(nn, 4)), dim3(256), 0, nn, nk,
         (const uint32_t*)far_ptr.ptr, (const float*)far_coef.ptr, alpha_node.ptr);
  launch(ctx, "bt_alpha_leaf", bt_alpha_leaf_kernel, dim3(blocks_for(nl, 256)), dim3(256), 0, nl, nk,
         (const uint32_t*)leaves.ptr, (const uint32_t*)parent.ptr, (const double*)alpha_node.ptr, alpha_leaf.ptr);
}

// Kernel names carry the channel count (timing per product kind).
#define BT_NAME(base) (C == 12 ? base "_c12" : C == 3 ? base "_c3" : base "_c")

template <int C, int H>
void BlockTreeDev::multiply(ddgrs_ctx* ctx, const float* u_hl, double w0, double w1, float* z_hl, bool extended) {
  sigma.reserve((size_t)nn * C);
  mean.reserve((size_t)nn * C);
  gamma.reserve((size_t)nn * C);
  delta.reserve((size_t)nl * C);
  TreeArgs ta{start.ptr, end.ptr, right.ptr, parent.ptr, leaves.ptr, weight.ptr, inv_mass.ptr, counters.ptr, nl};
  launch(ctx, BT_NAME("bt_upward"), bt_upward_kernel<C>, dim3(blocks_for(nl, 256)), dim3(256), 0, ta, u_hl, sigma.ptr,
         mean.ptr);
  if (nk == 2) {
    launch(ctx, BT_NAME("bt_far"), bt_far_kernel<C, H, 2>, dim3(blocks_for(nn, 4)), dim3(256), 0, nn,
           (const uint32_t*)far_ptr.ptr, (const uint32_t*)far_col.ptr, (const float*)far_coef.ptr,
           (const double*)mean.ptr, gamma.ptr);
    launch(ctx, BT_NAME("bt_leaf_far"), bt_leaf_far_kernel<C, H, 2>, dim3(blocks_for((int64_t)nl * C, 256)), dim3(256),
           0, ta, (const double*)alpha_node.ptr, (const double*)mean.ptr, (const double*)gamma.ptr, delta.ptr);
  } else {
    launch(ctx, BT_NAME("bt_far"), bt_far_kernel<C, H, 1>, dim3(blocks_for(nn, 4)), dim3(256), 0, nn,
           (const uint32_t*)far_ptr.ptr, (const uint32_t*)far_col.ptr, (const float*)far_coef.ptr,
           (const double*)mean.ptr, gamma.ptr);
    launch(ctx, BT_NAME("bt_leaf_far"), bt_leaf_far_kernel<C, H, 1>, dim3(blocks_for((int64_t)nl * C, 256)), dim3(256),
           0, ta, (const double*)alpha_node.ptr, (const double*)mean.ptr, (const double*)gamma.ptr, delta.ptr);
  }
  PanelArgs pa{panel_leaf.ptr, panel_row0.ptr, panel_rows.ptr, col_ptr.ptr,        cols.ptr, leaves.ptr,
               panel_val.ptr,  val.ptr,        alpha_leaf.ptr, (const double*)mean.ptr, (const double*)delta.ptr, np,
               val.ptr + (size_t)near_total * nk};
  SegArgs sa{seg_panel.ptr, seg_col0.ptr, seg_ptr.ptr, seg_arrive.ptr, nullptr, nullptr, seg_cols, nseg};
  const dim3 grid(blocks_for(nseg, 4));
  if (!extended) {
    partial.reserve((size_t)std::max(nseg, 1) * kPanelRows * C);
    sa.partial = partial.ptr;
    if (nk == 2)
      launch(ctx, BT_NAME("bt_panel"), bt_panel_kernel<C, H, 2, float>, grid, dim3(256), 0, pa, sa, u_hl, w0, w1, z_hl);
    else
      launch(ctx, BT_NAME("bt_panel"), bt_panel_kernel<C, H, 1, float>, grid, dim3(256), 0, pa, sa, u_hl, w0, w1, z_hl);
  } else {
    partial64.reserve((size_t)std::max(nseg, 1) * kPanelRows * C);
    sa.partial64 = partial64.ptr;
    if (nk == 2)
      launch(ctx, BT_NAME("bt_panel_x"), bt_panel_kernel<C, H, 2, double>, grid, dim3(256), 0, pa, sa, u_hl, w0, w1,
             z_hl);
    else
      launch(ctx, BT_NAME("bt_panel_x"), bt_panel_kernel<C, H, 1, double>, grid, dim3(256), 0, pa, sa, u_hl, w0, w1,
             z_hl);
  }
}

template void BlockTreeDev::multiply<12, 9>(ddgrs_ctx*, const float*, double, double, float*, bool);
template void BlockTreeDev::multiply<3, 3>(ddgrs_ctx*, const float*, double, double, float*, bool);

// ---- face metric ---------------------------------------------------------------------------

#pragma clang fp contract(off)

struct NodeGeo {
  const double *cen, *proj, *mass;  // 3, 6, 1 per node
};

// Far coefficients [ρ^{−2s}, ½(vᵀP̄_I v + vᵀP̄_J v) ρ^{−4} ρ^{−2s}] at the centroids, (min, max)
// ordered (hierarchical.rs:185 hi_lo, :206 far), times the column cluster's area.
__global__ __launch_bounds__(256) void fm_far_coef_kernel(uint32_t nfar, const uint32_t* __restrict__ far_row,
                                                          const uint32_t* __restrict__ far_col, NodeGeo g, double neg_s,
                                                          float* __restrict__ coef) {
  uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nfar) return;
  const uint32_t col = far_col[q];
  uint32_t a = far_row[q], b = col;
  if (a > b) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = g.cen[3 * a + k] - g.cen[3 * b + k];
  auto pq = [&](uint32_t n) {
    const double* p = g.proj + 6 * n;
    return p[0] * d[0] * d[0] + p[3] * d[1] * d[1] + p[5] * d[2] * d[2] +
           2.0 * (p[1] * d[0] * d[1] + p[2] * d[0] * d[2] + p[4] * d[1] * d[2]);
  };
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double hi = pow(r2, neg_s);
  double lo = 0.5 * (pq(a) + pq(b)) / (r2 * r2) * hi;
  const double m = g.mass[col];
  coef[2 * (size_t)q] = (float)(hi * m);
  coef[2 * (size_t)q + 1] = (float)(lo * m);
}

// Near values of one panel per wave: rows r, columns kk (lanes), (min, max)-ordered kernels times
// the column face's area; the diagonal is 0.
__global__ __launch_bounds__(256) void fm_panel_values_kernel(PanelArgs P, SegArgs G,
                                                              const uint32_t* __restrict__ order,
                                                              const double* __restrict__ bary,
                                                              const double* __restrict__ nrm,
                                                              const double* __restrict__ area, double neg_s,
                                                              float* __restrict__ val) {
  const uint32_t sg = xcd_wave_id();
  const int lane = threadIdx.x & 63;
  if (sg >= (uint32_t)G.n) return;
  const uint32_t p = G.panel[sg], j0 = G.col0[sg];
  const uint32_t k = P.panel_leaf[p], row0 = P.panel_row0[p], R = P.panel_rows[p];
  const uint32_t c0 = P.col_ptr[k], K = P.col_ptr[k + 1] - c0;
  const uint32_t kend = min(K, j0 + G.cols);
  float* out = val + 2 * P.panel_val[p];
  // the panel's rows (wave-uniform), then one column per lane: its data loaded once for all rows
  uint32_t srow[kPanelRows];
#pragma unroll
  for (int r = 0; r < (int)kPanelRows; ++r) srow[r] = order[row0 + ((uint32_t)r < R ? r : 0)];
  for (uint32_t kk = j0 + lane; kk < kend; kk += 64) {
    const uint32_t t0 = order[P.cols[c0 + kk]];
    const double bt[3] = {bary[3 * t0], bary[3 * t0 + 1], bary[3 * t0 + 2]};
    const double nt3[3] = {nrm[3 * t0], nrm[3 * t0 + 1], nrm[3 * t0 + 2]};
    const double at = area[t0];
#pragma unroll
    for (int r = 0; r < (int)kPanelRows; ++r) {
      if ((uint32_t)r >= R) break;
      const uint32_t s0 = srow[r];
      float* o = out + 2 * ((size_t)r * K + kk);
      if (s0 == t0) {
        o[0] = 0.f;
        o[1] = 0.f;
        continue;
      }
      const double bs[3] = {bary[3 * s0], bary[3 * s0 + 1], bary[3 * s0 + 2]};
      const double ns3[3] = {nrm[3 * s0], nrm[3 * s0 + 1], nrm[3 * s0 + 2]};
      // (min, max) face order as the reference evaluates the pair
      const bool sfirst = s0 < t0;
      const double* bi = sfirst ? bs : bt;
      const double* bj = sfirst ? bt : bs;
      const double* ni = sfirst ? ns3 : nt3;
      const double* nj = sfirst ? nt3 : ns3;
      double d[3];
      for (int x = 0; x < 3; ++x) d[x] = bi[x] - bj[x];
      double ns = ni[0] * d[0] + ni[1] * d[1] + ni[2] * d[2];
      double nt = nj[0] * d[0] + nj[1] * d[1] + nj[2] * d[2];
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double hi = pow(r2, neg_s);
      double lo = 0.5 * (ns * ns + nt * nt) / (r2 * r2) * hi;
      o[0] = (float)(hi * at);
      o[1] = (float)(lo * at);
    }
  }
}

// Per BVH position: D_f coefficients g_c = (n/|n|²) × e_opp(c) (RS/src/metric/mod.rs:120), area,
// vertices.
__global__ __launch_bounds__(256) void fm_face_kernel(int nf, const uint32_t* __restrict__ order,
                                                      const uint32_t* __restrict__ faces, const double* __restrict__ pos,
                                                      const double* __restrict__ area, float* __restrict__ df,
                                                      double* __restrict__ area_p, uint32_t* __restrict__ vert_p) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nf) return;
  uint32_t s = order[p];
  double q[3][3];
  for (int c = 0; c < 3; ++c) {
    uint32_t v = faces[3 * s + c];
    vert_p[3 * p + c] = v;
    for (int k = 0; k < 3; ++k) q[c][k] = pos[3 * v + k];
  }
  double e1[3], e2[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = q[1][k] - q[0][k];
    e2[k] = q[2][k] - q[0][k];
  }
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double inv = 1.0 / (len * len);
  double w[3] = {n[0] * inv, n[1] * inv, n[2] * inv};
  for (int c = 0; c < 3; ++c) {
    int i0 = (c + 1) % 3, i1 = (c + 2) % 3;
    double e[3] = {q[i1][0] - q[i0][0], q[i1][1] - q[i0][1], q[i1][2] - q[i0][2]};
    double g[3] = {w[1] * e[2] - w[2] * e[1], w[2] * e[0] - w[0] * e[2], w[0] * e[1] - w[1] * e[0]};
    for (int k = 0; k < 3; ++k) df[9 * p + 3 * c + k] = (float)g[k];
  }
  area_p[p] = area[s];
}

__device__ __forceinline__ void split_store(double x, float* hi, float* lo) {
  const float h = (float)x;
  *hi = h;
  *lo = (float)(x - (double)h);
}

// m = M x in FP64 (9 gradient channels 3k + r, then 3 average channels) per BVH position, stored
// as hi/lo pairs.
__global__ __launch_bounds__(256) void fm_m_kernel(int nf, const uint32_t* __restrict__ vert_p,
                                                   const float* __restrict__ df, const double* __restrict__ x,
                                                   float* __restrict__ m_hl) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nf) return;
  double xv[3][3];
  for (int c = 0; c < 3; ++c) {
    uint32_t v = vert_p[3 * p + c];
    for (int k = 0; k < 3; ++k) xv[c][k] = x[3 * v + k];
  }
  const float* g = df + 9 * p;
  float* o = m_hl + 24 * (size_t)p;
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) {
      double t = xv[0][k] * (double)g[r] + xv[1][k] * (double)g[3 + r] + xv[2][k] * (double)g[6 + r];
      split_store(t, o + 3 * k + r, o + 12 + 3 * k + r);
    }
  for (int k = 0; k < 3; ++k) split_store((xv[0][k] + xv[1][k] + xv[2][k]) / 3.0, o + 9 + k, o + 21 + k);
}

// y_v = 2 Σ_{S∋v} a_S (Mᵀ row of S at v)·z_S (faces ascending; FP64).
__global__ __launch_bounds__(256) void fm_gather_kernel(int nv, const uint32_t* __restrict__ inc_ptr,
                                                        const uint32_t* __restrict__ inc, const float* __restrict__ df,
                                                        const double* __restrict__ area_p, const float* __restrict__ z_hl,
                                                        double* __restrict__ y) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t q = inc_ptr[v]; q < inc_ptr[v + 1]; ++q) {
    uint32_t e = inc[q];
    uint32_t p = e >> 2, c = e & 3u;
    const double a = area_p[p];
    const float* g = df + 9 * (size_t)p + 3 * c;
    const float* zs = z_hl + 24 * (size_t)p;
    double z[12];
    for (int i = 0; i < 12; ++i) z[i] = (double)zs[i] + (double)zs[12 + i];
    for (int k = 0; k < 3; ++k) {
      double t = (double)g[0] * z[3 * k] + (double)g[1] * z[3 * k + 1] + (double)g[2] * z[3 * k + 2] + z[9 + k] / 3.0;
      acc[k] += a * t;
    }
  }
  for (int k = 0; k < 3; ++k) y[3 * (size_t)v + k] = 2.0 * acc[k];
}

FaceMetricPlan plan_face_metric(const std::vector<uint32_t>& faces, size_t nv, const Topology& t, const Clusters& cl,
                                double chi, size_t leaf_pairs, const std::vector<double>& face_area) {
  FaceMetricPlan P;
  const size_t nf = t.order.size();
  P.topo = t;
  P.bs = build_block_structure(t, cl, chi, leaf_pairs);
  std::vector<double> wbvh(nf);
  for (size_t p = 0; p < nf; ++p) wbvh[p] = face_area[t.order[p]];
  P.bt = plan_block_tree(t, P.bs, 2, std::move(wbvh), cl.mass);
  P.centroid = cl.centroid;
  P.projector = cl.projector;
  P.node_mass = cl.mass;
  // vertex → (BVH position of the face, corner), faces ascending
  const std::vector<uint32_t>& position = P.bt.position;
  P.inc_ptr.assign(nv + 1, 0);
  P.inc.resize(3 * nf);
  for (size_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) P.inc_ptr[faces[3 * f + c] + 1]++;
  for (size_t v = 0; v < nv; ++v) P.inc_ptr[v + 1] += P.inc_ptr[v];
  std::vector<uint32_t> fill(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
  for (size_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) P.inc[fill[faces[3 * f + c]]++] = (position[f] << 2) | (uint32_t)c;
  return P;
}

// vertex → (BVH position of the face, corner) from the surface's vertex → (face, corner) lists
__global__ __launch_bounds__(256) void fm_inc_kernel(uint32_t m, const uint32_t* __restrict__ position,
                                                     const uint32_t* __restrict__ inc, uint32_t* __restrict__ out) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q < m) out[q] = (position[inc[q] >> 2] << 2) | (inc[q] & 3u);
}

namespace {
// Far coefficients, α, near values and the per-face data of a face metric whose block tree is on
// the device.
void face_metric_values(ddgrs_ctx* ctx, ddgrs_surface* s, FaceMetricDev& F, const std::vector<double>& centroid,
                        const std::vector<double>& projector, const std::vector<double>& node_mass, double s_exp,
                        const DeviceTree* dt = nullptr) {
  BlockTreeDev& bt = F.bt;
  DevBuf<double>&cen = F.node_cen, &proj = F.node_proj, &nmass = F.node_mass;
  if (dt) {
    const size_t nn = node_mass.size();
    cen.reserve(3 * nn);
    proj.reserve(6 * nn);
    nmass.reserve(nn);
    HIPCHECK(hipMemcpyAsync(cen.ptr, dt->centroid, 3 * nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(proj.ptr, dt->projector, 6 * nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(nmass.ptr, dt->mass, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    ddgrs::upload(ctx, cen, centroid);
    ddgrs::upload(ctx, proj, projector);
    ddgrs::upload(ctx, nmass, node_mass);
  }
  const double neg_s = -s_exp;
  launch(ctx, "fm_far_coef", fm_far_coef_kernel, dim3(blocks_for(bt.n_far, 256)), dim3(256), 0, bt.n_far,
         (const uint32_t*)bt.far_row.ptr, (const uint32_t*)bt.far_col.ptr, NodeGeo{cen.ptr, proj.ptr, nmass.ptr}, neg_s,
         bt.far_coef.ptr);
  bt.finish_far(ctx);
  PanelArgs pa{bt.panel_leaf.ptr, bt.panel_row0.ptr, bt.panel_rows.ptr, bt.col_ptr.ptr, bt.cols.ptr, bt.leaves.ptr,
               bt.panel_val.ptr,  nullptr,          nullptr,           nullptr,         nullptr,     bt.np};
  const SegArgs sa{bt.seg_panel.ptr, bt.seg_col0.ptr, bt.seg_ptr.ptr, nullptr, nullptr, nullptr, bt.seg_cols, bt.nseg};
  launch(ctx, "fm_panel_values", fm_panel_values_kernel, dim3(blocks_for(bt.nseg, 4)), dim3(256), 0, pa, sa,
         (const uint32_t*)bt.order.ptr, (const double*)s->bary.ptr, (const double*)s->normal.ptr,
         (const double*)s->area.ptr, neg_s, bt.val.ptr);
  const int nf = F.nf;
  F.df.reserve(9 * (size_t)nf);
  F.area_p.reserve(nf);
  F.vert_p.reserve(3 * (size_t)nf);
  launch(ctx, "fm_face", fm_face_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, (const uint32_t*)bt.order.ptr,
         (const uint32_t*)s->faces.ptr, (const double*)s->positions.ptr, (const double*)s->area.ptr, F.df.ptr,
         F.area_p.ptr, F.vert_p.ptr);
  F.m_hl.reserve(24 * (size_t)nf);
  F.z_hl.reserve(24 * (size_t)nf);
}
}  // namespace

void FaceMetricDev::build(ddgrs_surface* s, const Topology& t, const Clusters& cl, double chi, size_t leaf_pairs,
                          double s_exp) {
  build(s->ctx, s, t, cl, chi, leaf_pairs, s_exp);
}

void FaceMetricDev::build(ddgrs_ctx* ctx, ddgrs_surface* s, const Topology& t, const Clusters& cl, double chi,
                          size_t leaf_pairs, double s_exp, const DeviceTree* dt) {
  topo = t;
  nv = (int)s->nv;
  nf = (int)s->nf;
  stats = bt.build_structure(ctx, t, cl, chi, leaf_pairs, 2, s->area.ptr, dt);
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  inc_ptr.reserve((size_t)nv + 1);
  HIPCHECK(hipMemcpyAsync(inc_ptr.ptr, s->inc_ptr.ptr, ((size_t)nv + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                          ctx->stream));
  inc.reserve(3 * (size_t)nf);
  launch(ctx, "fm_inc", fm_inc_kernel, dim3(blocks_for(3 * (int64_t)nf, 256)), dim3(256), 0, (uint32_t)(3 * nf),
         (const uint32_t*)bt.position.ptr, (const uint32_t*)s->inc.ptr, inc.ptr);
  face_metric_values(ctx, s, *this, cl.centroid, cl.projector, cl.mass, s_exp, dt);
  if (trace) {
    ctx->sync();
    std::fprintf(stderr, "blocks: %-14s %8.2f ms\n", "values",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() * 1e3);
  }
}

void FaceMetricDev::build(ddgrs_surface* s, const FaceMetricPlan& P, double s_exp) {
  ddgrs_ctx* ctx = s->ctx;
  topo = P.topo;
  stats = block_stats(P.bs);
  nv = (int)s->nv;
  nf = (int)s->nf;
  bt.upload(ctx, P.bt);
  bt.set_near_lists(ctx, P.bs);
  ddgrs::upload(ctx, inc_ptr, P.inc_ptr);
  ddgrs::upload(ctx, inc, P.inc);
  face_metric_values(ctx, s, *this, P.centroid, P.projector, P.node_mass, s_exp);
}

void FaceMetricDev::apply(ddgrs_ctx* ctx, const double* x, double* y) {
  launch(ctx, "fm_m", fm_m_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, (const uint32_t*)vert_p.ptr,
         (const float*)df.ptr, x, m_hl.ptr);
  bt.multiply<12, 9>(ctx, m_hl.ptr, w_hi, w_lo, z_hl.ptr, extended);
  launch(ctx, "fm_gather", fm_gather_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)inc_ptr.ptr,
         (const uint32_t*)inc.ptr, (const float*)df.ptr, (const double*)area_p.ptr, (const float*)z_hl.ptr, y);
}

// ---- vertex-pair fractional operator ------------------------------------------------------

// ρ^{−(τ+1)}... at the centroids, (min, max) ordered, times the column cluster's mass.
__global__ __launch_bounds__(256) void vf_far_coef_kernel(uint32_t nfar, const uint32_t* __restrict__ far_row,
                                                          const uint32_t* __restrict__ far_col,
                                                          const double* __restrict__ cen,
                                                          const double* __restrict__ nmass, double ex,
                                                          float* __restrict__ coef) {
  uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nfar) return;
  const uint32_t col = far_col[q];
  uint32_t a = far_row[q], b = col;
  if (a > b) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = cen[3 * a + k] - cen[3 * b + k];
  coef[q] = (float)(pow(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], ex) * nmass[col]);
}

__global__ __launch_bounds__(256) void vf_panel_values_kernel(PanelArgs P, SegArgs G,
                                                              const uint32_t* __restrict__ order,
                                                              const double* __restrict__ pos,
                                                              const double* __restrict__ mass, double ex,
                                                              float* __restrict__ val) {
  const uint32_t sg = xcd_wave_id();
  const int lane = threadIdx.x & 63;
  if (sg >= (uint32_t)G.n) return;
  const uint32_t p = G.panel[sg], j0 = G.col0[sg];
  const uint32_t k = P.panel_leaf[p], row0 = P.panel_row0[p], R = P.panel_rows[p];
  const uint32_t c0 = P.col_ptr[k], K = P.col_ptr[k + 1] - c0;
  const uint32_t kend = min(K, j0 + G.cols);
  float* out = val + P.panel_val[p];
  uint32_t srow[kPanelRows];
#pragma unroll
  for (int r = 0; r < (int)kPanelRows; ++r) srow[r] = order[row0 + ((uint32_t)r < R ? r : 0)];
  for (uint32_t kk = j0 + lane; kk < kend; kk += 64) {
    const uint32_t t0 = order[P.cols[c0 + kk]];
    const double pt[3] = {pos[3 * t0], pos[3 * t0 + 1], pos[3 * t0 + 2]};
    const double mt = mass[t0];
#pragma unroll
    for (int r = 0; r < (int)kPanelRows; ++r) {
      if ((uint32_t)r >= R) break;
      const uint32_t s0 = srow[r];
      float* o = out + (size_t)r * K + kk;
      if (s0 == t0) {
        *o = 0.f;
        continue;
      }
      const double ps[3] = {pos[3 * s0], pos[3 * s0 + 1], pos[3 * s0 + 2]};
      const bool sfirst = s0 < t0;
      double d[3];
      for (int x = 0; x < 3; ++x) d[x] = sfirst ? ps[x] - pt[x] : pt[x] - ps[x];
      *o = (float)(pow(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], ex) * mt);
    }
  }
}

__global__ __launch_bounds__(256) void vf_u_kernel(int n, const uint32_t* __restrict__ order,
                                                   const double* __restrict__ x, float* __restrict__ u_hl) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  uint32_t v = order[p];
  for (int k = 0; k < 3; ++k) split_store(x[3 * (size_t)v + k], u_hl + 6 * (size_t)p + k, u_hl + 6 * (size_t)p + 3 + k);
}

__global__ __launch_bounds__(256) void vf_out_kernel(int n, const uint32_t* __restrict__ position,
                                                     const double* __restrict__ mass, const float* __restrict__ z_hl,
                                                     double* __restrict__ y) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  uint32_t p = position[v];
  double a = mass[v];
  for (int k = 0; k < 3; ++k) {
    const double z = (double)z_hl[6 * (size_t)p + k] + (double)z_hl[6 * (size_t)p + 3 + k];
    y[3 * (size_t)v + k] = 2.0 * a * z;
  }
}

VertexFractionalPlan plan_vertex_fractional(const std::vector<double>& pos, const std::vector<double>& m,
                                            double tau, double chi, size_t leaf_pairs, Topology topo) {
  VertexFractionalPlan P;
  const size_t nv = m.size();
  // vertex BVH: point f_v, box the point, mass a_v (hierarchical.rs:1017-1060)
  P.topo = std::move(topo);
  P.clusters = build_clusters(P.topo, pos.data(), pos.data(), pos.data(), m.data());
  P.bs = build_block_structure(P.topo, P.clusters, chi, leaf_pairs);
  std::vector<double> wbvh(nv);
  for (size_t p = 0; p < nv; ++p) wbvh[p] = m[P.topo.order[p]];
  P.bt = plan_block_tree(P.topo, P.bs, 1, std::move(wbvh), P.clusters.mass);
  P.pos = pos;
  P.mass = m;
  P.tau = tau;
  return P;
}

namespace {
void vertex_fractional_values(ddgrs_ctx* ctx, VertexFractionalDev& V, const Clusters& cl, const std::vector<double>& pos,
                              double tau) {
  BlockTreeDev& bt = V.bt;
  const double ex = -(tau + 1.0);
  DevBuf<double>&cen = V.node_cen, &dpos = V.pos_dev, &nmass = V.node_mass;
  ddgrs::upload(ctx, cen, cl.centroid);
  ddgrs::upload(ctx, dpos, pos);
  ddgrs::upload(ctx, nmass, cl.mass);
  launch(ctx, "vf_far_coef", vf_far_coef_kernel, dim3(blocks_for(bt.n_far, 256)), dim3(256), 0, bt.n_far,
         (const uint32_t*)bt.far_row.ptr, (const uint32_t*)bt.far_col.ptr, (const double*)cen.ptr,
         (const double*)nmass.ptr, ex, bt.far_coef.ptr);
  bt.finish_far(ctx);
  PanelArgs pa{bt.panel_leaf.ptr, bt.panel_row0.ptr, bt.panel_rows.ptr, bt.col_ptr.ptr, bt.cols.ptr, bt.leaves.ptr,
               bt.panel_val.ptr,  nullptr,          nullptr,           nullptr,         nullptr,     bt.np};
  const SegArgs sa{bt.seg_panel.ptr, bt.seg_col0.ptr, bt.seg_ptr.ptr, nullptr, nullptr, nullptr, bt.seg_cols, bt.nseg};
  launch(ctx, "vf_panel_values", vf_panel_values_kernel, dim3(blocks_for(bt.nseg, 4)), dim3(256), 0, pa, sa,
         (const uint32_t*)bt.order.ptr, (const double*)dpos.ptr, (const double*)V.mass.ptr, ex, bt.val.ptr);
  V.u_hl.reserve(6 * (size_t)V.nv);
  V.z_hl.reserve(6 * (size_t)V.nv);
}
}  // namespace

void VertexFractionalDev::build(ddgrs_ctx* ctx, const VertexFractionalPlan& P) {
  nv = (int)P.mass.size();
  topo = P.topo;
  stats = block_stats(P.bs);
  bt.upload(ctx, P.bt);
  bt.set_near_lists(ctx, P.bs);
  ddgrs::upload(ctx, mass, P.mass);
  vertex_fractional_values(ctx, *this, P.clusters, P.pos, P.tau);
}

void VertexFractionalDev::build(ddgrs_ctx* ctx, const std::vector<double>& pos, const std::vector<double>& m, double tau,
                                double chi, size_t leaf_pairs, Topology t) {
  nv = (int)m.size();
  topo = std::move(t);
  // vertex BVH: point f_v, box the point, mass a_v (hierarchical.rs:1017-1060)
  const Clusters cl = build_clusters(topo, pos.data(), pos.data(), pos.data(), m.data());
  ddgrs::upload(ctx, mass, m);
  stats = bt.build_structure(ctx, topo, cl, chi, leaf_pairs, 1, mass.ptr);
  vertex_fractional_values(ctx, *this, cl, pos, tau);
}

void VertexFractionalDev::apply(ddgrs_ctx* ctx, const double* x, double* y) {
  launch(ctx, "vf_u", vf_u_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)bt.order.ptr, x,
         u_hl.ptr);
  bt.multiply<3, 3>(ctx, u_hl.ptr, 1.0, 1.0, z_hl.ptr, extended);
  launch(ctx, "vf_out", vf_out_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)bt.position.ptr,
         (const double*)mass.ptr, (const float*)z_hl.ptr, y);
}

}  // namespace ddgrs
