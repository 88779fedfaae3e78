// The reference preconditioner on the device (see precond.hpp).
//
// Host: K = W + εM (RS/src/geometry.rs:221 cotan_stiffness, RS/src/solve/dense.rs:529 ε = ε_rel
// tr W / tr M), the scalar saddle L′ of the global barycenter row (precond.rs:184-240, per-axis
// case, multipliers ordered last), its LDLᵀ factor in a nested-dissection order, and the dense
// inverses of the diagonal blocks of L over the separator tree.
//
// Device: the three axes of a saddle vector are solved together in the permuted numbering (n × 3
// row-major).  The separator tree is walked by height, one launch per height and direction, one
// workgroup per (block, chunk of 64 rows of the block): the workgroup forms the block's right-hand
// side in LDS from the rows (forward) or columns (backward) of L outside the block, then multiplies
// by the dense inverse of the block's diagonal part.  A separator's rows only meet its descendants
// (forward) and ancestors (backward) outside the block, so the blocks of one height are
// independent.  FP64 throughout: L′ has a near-null space of relative size ε = 1e-10.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <future>
#include <numeric>

#include "bh.hpp"
#include "host_parallel.hpp"
#include "lprime_gpu.hpp"
#include "precond.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

namespace {
constexpr uint32_t kNone = UINT32_MAX;
constexpr uint32_t kChunk = 64;      // rows of a block per workgroup (fused heights)
constexpr uint32_t kLongParts = 128; // workgroups summing the multiplier row's forward residual

HostCsr csr_from_triplets(size_t n, std::vector<std::tuple<uint32_t, uint32_t, double>>& t) {
  std::stable_sort(t.begin(), t.end(), [](const auto& a, const auto& b) {
    return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
  });
  HostCsr m;
  m.n = n;
  m.ptr.assign(n + 1, 0);
  for (size_t q = 0; q < t.size();) {
    uint32_t r = std::get<0>(t[q]), c = std::get<1>(t[q]);
    double v = 0.0;
    while (q < t.size() && std::get<0>(t[q]) == r && std::get<1>(t[q]) == c) v += std::get<2>(t[q++]);
    m.idx.push_back(c);
    m.val.push_back(v);
    m.ptr[r + 1]++;
  }
  for (size_t i = 0; i < n; ++i) m.ptr[i + 1] += m.ptr[i];
  return m;
}
}  // namespace

NdBlocks nested_dissection(const HostCsr& k, const std::vector<double>& pos) {
  const size_t n = k.n;
  NdBlocks nd;
  nd.order.reserve(n);
  std::vector<int32_t> side(n, -1);
  struct Task {
    std::vector<uint32_t> verts;
    bool emit_only;  // a separator, appended after its two domains
    uint32_t id, parent;
  };
  // Iterative post-order: a domain task expands into [domain A, domain B, separator]; the
  // separator keeps the domain's block id (it is the root of the domain's subtree).
  auto new_block = [&](uint32_t parent) {
    nd.start.push_back(0);
    nd.end.push_back(0);
    nd.parent.push_back(parent);
    return (uint32_t)nd.parent.size() - 1;
  };
  std::vector<Task> stack;
  std::vector<uint32_t> all(n);
  std::iota(all.begin(), all.end(), 0u);
  stack.push_back({std::move(all), false, new_block(kNone), kNone});
  while (!stack.empty()) {
    Task t = std::move(stack.back());
    stack.pop_back();
    if (t.emit_only || t.verts.size() <= 64) {
      nd.start[t.id] = (uint32_t)nd.order.size();
      nd.order.insert(nd.order.end(), t.verts.begin(), t.verts.end());
      nd.end[t.id] = (uint32_t)nd.order.size();
      continue;
    }
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v : t.verts)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], pos[3 * v + a]);
        hi[a] = std::max(hi[a], pos[3 * v + a]);
      }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
      if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
    std::vector<uint32_t>& v = t.verts;
    const size_t half = v.size() / 2;
    std::nth_element(v.begin(), v.begin() + half, v.end(), [&](uint32_t a, uint32_t b) {
      double xa = pos[3 * a + ax], xb = pos[3 * b + ax];
      return xa != xb ? xa < xb : a < b;
    });
    for (size_t i = 0; i < v.size(); ++i) side[v[i]] = i < half ? 0 : 1;
    std::vector<uint32_t> A, B, S;
    for (size_t i = 0; i < v.size(); ++i) {
      uint32_t x = v[i];
      if (side[x] == 1) {
        B.push_back(x);
        continue;
      }
      bool touches = false;
      for (uint32_t q = k.ptr[x]; q < k.ptr[x + 1] && !touches; ++q) touches = side[k.idx[q]] == 1;
      (touches ? S : A).push_back(x);
    }
    for (uint32_t x : v) side[x] = -1;
    std::sort(A.begin(), A.end());
    std::sort(B.begin(), B.end());
    std::sort(S.begin(), S.end());
    const uint32_t ia = new_block(t.id), ib = new_block(t.id);
    stack.push_back({std::move(S), true, t.id, t.parent});
    stack.push_back({std::move(B), false, ib, t.id});
    stack.push_back({std::move(A), false, ia, t.id});
  }
  return nd;
}

Factor factor_ldlt(const HostCsr& a, const std::vector<uint32_t>& order) {
  const size_t n = a.n;
  Factor f;
  f.n = n;
  f.perm = order;
  std::vector<uint32_t> pinv(n);
  for (size_t i = 0; i < n; ++i) pinv[order[i]] = (uint32_t)i;
  std::vector<uint32_t> parent(n, kNone), flag(n), lnz(n, 0);
  for (size_t k = 0; k < n; ++k) {
    flag[k] = (uint32_t)k;
    uint32_t kk = order[k];
    for (uint32_t q = a.ptr[kk]; q < a.ptr[kk + 1]; ++q) {
      uint32_t i = pinv[a.idx[q]];
      if (i < k)
        for (; flag[i] != k; i = parent[i]) {
          if (parent[i] == kNone) parent[i] = (uint32_t)k;
          lnz[i]++;
          flag[i] = (uint32_t)k;
        }
    }
  }
  f.cp.assign(n + 1, 0);
  for (size_t k = 0; k < n; ++k) f.cp[k + 1] = f.cp[k] + lnz[k];
  f.ci.assign(f.cp[n], 0);
  f.cx.assign(f.cp[n], 0.0);
  f.d.assign(n, 0.0);
  std::vector<double> y(n, 0.0);
  std::vector<uint32_t> pattern(n), cnt(n, 0);
  for (size_t k = 0; k < n; ++k) {
    y[k] = 0.0;
    size_t top = n;
    flag[k] = (uint32_t)k;
    uint32_t kk = order[k];
    for (uint32_t q = a.ptr[kk]; q < a.ptr[kk + 1]; ++q) {
      uint32_t i = pinv[a.idx[q]];
      if (i <= k) {
        y[i] += a.val[q];
        size_t len = 0;
        for (; flag[i] != k; i = parent[i]) {
          pattern[len++] = i;
          flag[i] = (uint32_t)k;
        }
        while (len > 0) pattern[--top] = pattern[--len];
      }
    }
    f.d[k] = y[k];
    y[k] = 0.0;
    for (; top < n; ++top) {
      uint32_t i = pattern[top];
      double yi = y[i];
      y[i] = 0.0;
      uint32_t p2 = f.cp[i] + cnt[i];
      for (uint32_t q = f.cp[i]; q < p2; ++q) y[f.ci[q]] -= f.cx[q] * yi;
      double l_ki = yi / f.d[i];
      f.d[k] -= l_ki * yi;
      f.ci[p2] = (uint32_t)k;
      f.cx[p2] = l_ki;
      cnt[i]++;
    }
    if (f.d[k] == 0.0 || !std::isfinite(f.d[k])) {
      Error e(DDGRS_SINGULAR_SADDLE, "preconditioner factorization", "zero or non-finite pivot " + std::to_string(k));
      e.a = (int64_t)k;
      throw e;
    }
  }
  // rows of L (columns increasing)
  f.rp.assign(n + 1, 0);
  for (uint32_t i : f.ci) f.rp[i + 1]++;
  for (size_t i = 0; i < n; ++i) f.rp[i + 1] += f.rp[i];
  f.rj.resize(f.ci.size());
  f.rx.resize(f.ci.size());
  std::vector<uint32_t> fill(f.rp.begin(), f.rp.end() - 1);
  for (size_t j = 0; j < n; ++j)
    for (uint32_t q = f.cp[j]; q < f.cp[j + 1]; ++q) {
      uint32_t i = f.ci[q];
      f.rj[fill[i]] = (uint32_t)j;
      f.rx[fill[i]++] = f.cx[q];
    }
  return f;
}

// ---- multifrontal factorization over the separator tree --------------------------------------

Factor factor_multifrontal(const HostCsr& a, const NdBlocks& nd) {
  const size_t nv = nd.order.size(), n = nv + 1;
  Factor F;
  F.n = n;
  F.perm = nd.order;
  F.perm.push_back((uint32_t)nv);
  std::vector<uint32_t> pinv(n);
  for (size_t i = 0; i < n; ++i) pinv[F.perm[i]] = (uint32_t)i;
  // blocks: the separator tree's, then the multiplier row as the root
  std::vector<uint32_t> bstart = nd.start, bend = nd.end, bparent = nd.parent;
  const uint32_t nb0 = (uint32_t)bstart.size(), nb = nb0 + 1;
  bstart.push_back((uint32_t)n - 1);
  bend.push_back((uint32_t)n);
  bparent.push_back(kNone);
  for (uint32_t b = 0; b < nb0; ++b)
    if (bparent[b] == kNone) bparent[b] = nb0;
  std::vector<std::vector<uint32_t>> children(nb);
  for (uint32_t b = 0; b < nb0; ++b) children[bparent[b]].push_back(b);
  std::vector<uint32_t> height(nb, 0);
  for (uint32_t b = nb0; b-- > 0;) height[bparent[b]] = std::max(height[bparent[b]], height[b] + 1);
  const uint32_t H = height[nb0] + 1;
  // symbolic: front rows = the block's rows, then struct(J) = rows ≥ end(J) of A's columns in J and
  // of the children's structs (children before parents: decreasing ids, the root last)
  std::vector<std::vector<uint32_t>> front(nb);
  auto symbolic = [&](uint32_t b) {
    const uint32_t j0 = bstart[b], j1 = bend[b];
    std::vector<uint32_t> st;
    for (uint32_t j = j0; j < j1; ++j) {
      const uint32_t o = F.perm[j];
      for (uint32_t q = a.ptr[o]; q < a.ptr[o + 1]; ++q) {
        const uint32_t i = pinv[a.idx[q]];
        if (i >= j1) st.push_back(i);
      }
    }
    for (uint32_t c : children[b])
      for (size_t r = bend[c] - bstart[c]; r < front[c].size(); ++r) {
        if (front[c][r] < j0) throw Error(DDGRS_GPU, "preconditioner", "separator tree: a front row precedes its parent");
        if (front[c][r] >= j1) st.push_back(front[c][r]);
      }
    std::sort(st.begin(), st.end());
    st.erase(std::unique(st.begin(), st.end()), st.end());
    auto& fr = front[b];
    fr.resize(j1 - j0 + st.size());
    for (uint32_t j = j0; j < j1; ++j) fr[j - j0] = j;
    std::copy(st.begin(), st.end(), fr.begin() + (j1 - j0));
  };
  for (uint32_t b = nb0; b-- > 0;) symbolic(b);
  symbolic(nb0);
  // numeric, by height; update matrices live until the parent has assembled them
  std::vector<std::vector<double>> upd(nb);
  std::vector<std::vector<std::pair<uint32_t, double>>> lcol(n);  // column entries (row, value), rows ascending
  F.d.assign(n, 0.0);
  std::vector<std::vector<uint32_t>> by_height(H);
  for (uint32_t b = 0; b < nb; ++b) by_height[height[b]].push_back(b);
  std::atomic<int64_t> bad{-1};
  auto eliminate = [&](uint32_t b, bool inner_parallel) {
    const uint32_t j0 = bstart[b], k = bend[b] - bstart[b];
    const auto& fr = front[b];
    const size_t f = fr.size();
    auto local = [&](uint32_t i) -> size_t {
      if (i < bend[b]) return i - j0;
      return k + (size_t)(std::lower_bound(fr.begin() + k, fr.end(), i) - (fr.begin() + k));
    };
    std::vector<double> M(f * f, 0.0);  // row-major, lower triangle used
    // assemble A's columns of the block (entries on or below the diagonal)
    for (uint32_t c = 0; c < k; ++c) {
      const uint32_t j = j0 + c, o = F.perm[j];
      for (uint32_t q = a.ptr[o]; q < a.ptr[o + 1]; ++q) {
        const uint32_t i = pinv[a.idx[q]];
        if (i >= j) M[local(i) * f + c] += a.val[q];
      }
    }
    // extend-add the children's update matrices
    for (uint32_t cb : children[b]) {
      const auto& cf = front[cb];
      const size_t ck = bend[cb] - bstart[cb], cs = cf.size() - ck;
      std::vector<size_t> map(cs);
      for (size_t r = 0; r < cs; ++r) map[r] = local(cf[ck + r]);
      const auto& U = upd[cb];
      for (size_t r = 0; r < cs; ++r)
        for (size_t c = 0; c <= r; ++c) M[map[r] * f + map[c]] += U[r * cs + c];
      std::vector<double>().swap(upd[cb]);
    }
    // blocked right-looking LDLᵀ of the first k columns (rank-NB trailing updates)
    constexpr size_t NB = 32;
    std::vector<double> W(f * NB);  // unscaled panel columns l·d
    for (size_t c0 = 0; c0 < k; c0 += NB) {
      const size_t c1 = std::min<size_t>(c0 + NB, k), nbk = c1 - c0;
      for (size_t c = c0; c < c1; ++c) {
        const double dc = M[c * f + c];
        if (dc == 0.0 || !std::isfinite(dc)) {
          bad = (int64_t)(j0 + c);
          return;
        }
        for (size_t i = c + 1; i < f; ++i) {
          const double w = M[i * f + c];
          W[i * NB + (c - c0)] = w;
          M[i * f + c] = w / dc;
        }
        // update the rest of the panel's columns
        for (size_t i = c + 1; i < f; ++i) {
          const double li = M[i * f + c];
          const size_t mend = std::min(i + 1, c1);
          for (size_t m = c + 1; m < mend; ++m) M[i * f + m] -= li * W[m * NB + (c - c0)];
        }
      }
      // trailing update of columns ≥ c1: M[i][m] −= Σ_c L[i][c] W[m][c]
      auto rows = [&](size_t i) {
        double* mi = &M[i * f];
        const double* li = &M[i * f + c0];
        for (size_t m = c1; m <= i; ++m) {
          const double* wm = &W[m * NB];
          double acc = 0.0;
          for (size_t t = 0; t < nbk; ++t) acc += li[t] * wm[t];
          mi[m] -= acc;
        }
      };
      if (inner_parallel && f - c1 >= 256)
        parallel(f - c1, [&](size_t r) { rows(c1 + r); });
      else
        for (size_t i = c1; i < f; ++i) rows(i);
    }
    for (uint32_t c = 0; c < k; ++c) {
      F.d[j0 + c] = M[c * f + c];
      auto& col = lcol[j0 + c];
      for (size_t i = c + 1; i < f; ++i) {
        const double v = M[i * f + c];
        if (v != 0.0) col.push_back({fr[i], v});
      }
    }
    const size_t s2 = f - k;
    std::vector<double> U(s2 * s2, 0.0);
    for (size_t r = 0; r < s2; ++r)
      for (size_t c = 0; c <= r; ++c) U[r * s2 + c] = M[(k + r) * f + (k + c)];
    upd[b] = std::move(U);
  };
  static const bool trace = std::getenv("DDGRS_TRACE_FACTOR") != nullptr;
  auto t_num = std::chrono::steady_clock::now();
  for (uint32_t h = 0; h < H; ++h) {
    auto t_h = std::chrono::steady_clock::now();
    const auto& bl = by_height[h];
    if (bl.size() >= 8)
      parallel(bl.size(), [&](size_t i) { eliminate(bl[i], false); });
    else
      for (uint32_t b : bl) eliminate(b, true);
    if (trace) {
      size_t fmax = 0;
      for (uint32_t b : bl) fmax = std::max(fmax, front[b].size());
      std::fprintf(stderr, "factor height %u: %zu blocks, largest front %zu, %.2f ms\n", h, bl.size(), fmax,
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t_h).count() * 1e3);
    }
    if (bad >= 0) {
      Error e(DDGRS_SINGULAR_SADDLE, "preconditioner factorization", "zero or non-finite pivot " + std::to_string(bad.load()));
      e.a = bad;
      throw e;
    }
  }
  if (trace)
    std::fprintf(stderr, "factor numeric %.2f ms\n",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t_num).count() * 1e3);
  // column and row storage of L
  F.cp.assign(n + 1, 0);
  for (size_t j = 0; j < n; ++j) F.cp[j + 1] = F.cp[j] + (uint32_t)lcol[j].size();
  F.ci.resize(F.cp[n]);
  F.cx.resize(F.cp[n]);
  parallel(n, [&](size_t j) {
    uint32_t q = F.cp[j];
    for (auto& e : lcol[j]) {
      F.ci[q] = e.first;
      F.cx[q++] = e.second;
    }
  });
  F.rp.assign(n + 1, 0);
  for (uint32_t i : F.ci) F.rp[i + 1]++;
  for (size_t i = 0; i < n; ++i) F.rp[i + 1] += F.rp[i];
  F.rj.resize(F.ci.size());
  F.rx.resize(F.ci.size());
  std::vector<uint32_t> fill(F.rp.begin(), F.rp.end() - 1);
  for (size_t j = 0; j < n; ++j)
    for (uint32_t q = F.cp[j]; q < F.cp[j + 1]; ++q) {
      const uint32_t i = F.ci[q];
      F.rj[fill[i]] = (uint32_t)j;
      F.rx[fill[i]++] = F.cx[q];
    }
  return F;
}

// ---- device solves ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void lp_gather_kernel(uint32_t n, uint32_t nv, const uint32_t* __restrict__ perm,
                                                        const double* __restrict__ r, double* __restrict__ xb) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * n) return;
  uint32_t j = t / 3, a = t % 3;
  uint32_t old = perm[j];
  xb[t] = old < nv ? r[3 * (size_t)old + a] : r[3 * (size_t)nv + a];
}

__global__ __launch_bounds__(256) void lp_scatter_kernel(uint32_t n, uint32_t nv, const uint32_t* __restrict__ perm,
                                                         const double* __restrict__ xb, double* __restrict__ x) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * n) return;
  uint32_t j = t / 3, a = t % 3;
  uint32_t old = perm[j];
  x[old < nv ? 3 * (size_t)old + a : 3 * (size_t)nv + a] = xb[t];
}

__device__ __forceinline__ double wsum64d(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

struct LpArgs {
  const uint32_t *item_block, *item_row0, *blk_start, *blk_rows;
  const uint64_t* blk_inv;
  const double* inv;
  const uint32_t *ptr, *split, *idx;  // forward: rows of L; backward: columns of L
  const double *val, *d;
};

// Σ over [q0, q1) of val[q] x[idx[q]] (3 axes) by one wave.
__device__ __forceinline__ void wave_dot3(const LpArgs& A, uint32_t q0, uint32_t q1, const double* __restrict__ x,
                                          int lane, double& s0, double& s1, double& s2) {
  s0 = s1 = s2 = 0.0;
  for (uint32_t q = q0 + lane; q < q1; q += 64) {
    const double v = A.val[q];
    const double* xk = x + 3 * (size_t)A.idx[q];
    s0 += v * xk[0];
    s1 += v * xk[1];
    s2 += v * xk[2];
  }
  s0 = wsum64d(s0);
  s1 = wsum64d(s1);
  s2 = wsum64d(s2);
}

// Σ over [q0, q1) by the whole workgroup (a single long row: the multiplier row).
__device__ __forceinline__ void block_dot3(const LpArgs& A, uint32_t q0, uint32_t q1, const double* __restrict__ x,
                                           double* red, double& s0, double& s1, double& s2) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (uint32_t q = q0 + t; q < q1; q += 256) {
    const double v = A.val[q];
    const double* xk = x + 3 * (size_t)A.idx[q];
    a0 += v * xk[0];
    a1 += v * xk[1];
    a2 += v * xk[2];
  }
  a0 = wsum64d(a0);
  a1 = wsum64d(a1);
  a2 = wsum64d(a2);
  if (lane == 0) {
    red[3 * wave] = a0;
    red[3 * wave + 1] = a1;
    red[3 * wave + 2] = a2;
  }
  __syncthreads();
  s0 = (red[0] + red[3]) + (red[6] + red[9]);
  s1 = (red[1] + red[4]) + (red[7] + red[10]);
  s2 = (red[2] + red[5]) + (red[8] + red[11]);
}

// Forward, one height of the separator tree: workgroup = (block J, rows [row0, row0 + 64) of J).
// r_J = b_J − L_{J, j < start} x (rows of L up to the block), then x = M_J r_J, M_J = L_JJ⁻¹.
__global__ __launch_bounds__(256) void lp_forward_kernel(LpArgs A, uint32_t item0, const double* __restrict__ b,
                                                         double* __restrict__ xf) {
  extern __shared__ double lds[];  // 12 reduction slots, then r (k × 3)
  double* red = lds;
  double* r = lds + 16;
  const uint32_t it = item0 + blockIdx.x;
  const uint32_t J = A.item_block[it], row0 = A.item_row0[it];
  const uint32_t j0 = A.blk_start[J], k = A.blk_rows[J];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (k == 1) {
    double s0, s1, s2;
    block_dot3(A, A.ptr[j0], A.split[j0], xf, red, s0, s1, s2);
    if (threadIdx.x == 0) {
      r[0] = b[3 * (size_t)j0] - s0;
      r[1] = b[3 * (size_t)j0 + 1] - s1;
      r[2] = b[3 * (size_t)j0 + 2] - s2;
    }
  } else {
    for (uint32_t ii = wave; ii < k; ii += 4) {
      const uint32_t i = j0 + ii;
      double s0, s1, s2;
      wave_dot3(A, A.ptr[i], A.split[i], xf, lane, s0, s1, s2);
      if (lane == 0) {
        r[3 * ii] = b[3 * (size_t)i] - s0;
        r[3 * ii + 1] = b[3 * (size_t)i + 1] - s1;
        r[3 * ii + 2] = b[3 * (size_t)i + 2] - s2;
      }
    }
  }
  __syncthreads();
  const double* M = A.inv + A.blk_inv[J];
  const uint32_t rend = min(row0 + kChunk, k);
  for (uint32_t ii = row0 + wave; ii < rend; ii += 4) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l = lane; l <= ii; l += 64) {
      const double m = M[(size_t)ii * k + l];
      s0 += m * r[3 * l];
      s1 += m * r[3 * l + 1];
      s2 += m * r[3 * l + 2];
    }
    s0 = wsum64d(s0);
    s1 = wsum64d(s1);
    s2 = wsum64d(s2);
    if (lane == 0) {
      const size_t i = j0 + ii;
      xf[3 * i] = s0;
      xf[3 * i + 1] = s1;
      xf[3 * i + 2] = s2;
    }
  }
}

// Backward, one height: r_J = D_J⁻¹ y_J − L_{i ≥ end, J}ᵀ x (columns of L below the block), then
// x = M_Jᵀ r_J.
__global__ __launch_bounds__(256) void lp_backward_kernel(LpArgs A, uint32_t item0, const double* __restrict__ y,
                                                          double* __restrict__ xb) {
  extern __shared__ double lds[];  // 12 reduction slots, 4 × 64 × 3 partials, then r (k × 3)
  double* red = lds;
  double* part = lds + 16;
  double* r = lds + 16 + 768;
  const uint32_t it = item0 + blockIdx.x;
  const uint32_t J = A.item_block[it], row0 = A.item_row0[it];
  const uint32_t j0 = A.blk_start[J], k = A.blk_rows[J];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (k == 1) {
    double s0, s1, s2;
    block_dot3(A, A.split[j0], A.ptr[j0 + 1], xb, red, s0, s1, s2);
    if (threadIdx.x == 0) {
      const double dj = A.d[j0];
      r[0] = y[3 * (size_t)j0] / dj - s0;
      r[1] = y[3 * (size_t)j0 + 1] / dj - s1;
      r[2] = y[3 * (size_t)j0 + 2] / dj - s2;
    }
  } else {
    for (uint32_t ii = wave; ii < k; ii += 4) {
      const uint32_t i = j0 + ii;
      double s0, s1, s2;
      wave_dot3(A, A.split[i], A.ptr[i + 1], xb, lane, s0, s1, s2);
      if (lane == 0) {
        const double di = A.d[i];
        r[3 * ii] = y[3 * (size_t)i] / di - s0;
        r[3 * ii + 1] = y[3 * (size_t)i + 1] / di - s1;
        r[3 * ii + 2] = y[3 * (size_t)i + 2] / di - s2;
      }
    }
  }
  __syncthreads();
  const double* M = A.inv + A.blk_inv[J];
  const uint32_t ii = row0 + lane;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (ii < k)
    for (uint32_t l = ii + wave; l < k; l += 4) {
      const double m = M[(size_t)l * k + ii];
      s0 += m * r[3 * l];
      s1 += m * r[3 * l + 1];
      s2 += m * r[3 * l + 2];
    }
  part[3 * threadIdx.x] = s0;
  part[3 * threadIdx.x + 1] = s1;
  part[3 * threadIdx.x + 2] = s2;
  __syncthreads();
  if (wave == 0 && ii < k) {
    const size_t i = j0 + ii;
    for (int a = 0; a < 3; ++a)
      xb[3 * i + a] = (part[3 * lane + a] + part[3 * (lane + 64) + a]) + (part[3 * (lane + 128) + a] + part[3 * (lane + 192) + a]);
  }
}

// Split heights, residual phase: one wave per row i of the height.
// forward: r_i = b_i − Σ_{j < block start} L_ij x_j;  backward: r_i = y_i / d_i − Σ_{j ≥ block end} L_ji x_j.
__global__ __launch_bounds__(256) void lp_residual_kernel(LpArgs A, const uint32_t* __restrict__ rows, uint32_t count,
                                                          int backward, const double* __restrict__ rhs,
                                                          const double* __restrict__ x, double* __restrict__ rbuf) {
  const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= count) return;
  const uint32_t i = rows[w];
  const uint32_t q0 = backward ? A.split[i] : A.ptr[i], q1 = backward ? A.ptr[i + 1] : A.split[i];
  double s0, s1, s2;
  wave_dot3(A, q0, q1, x, lane, s0, s1, s2);
  if (lane == 0) {
    if (backward) {
      rbuf[3 * (size_t)i] = rhs[3 * (size_t)i] / A.d[i] - s0;
      rbuf[3 * (size_t)i + 1] = rhs[3 * (size_t)i + 1] / A.d[i] - s1;
      rbuf[3 * (size_t)i + 2] = rhs[3 * (size_t)i + 2] / A.d[i] - s2;
    } else {
      rbuf[3 * (size_t)i] = rhs[3 * (size_t)i] - s0;
      rbuf[3 * (size_t)i + 1] = rhs[3 * (size_t)i + 1] - s1;
      rbuf[3 * (size_t)i + 2] = rhs[3 * (size_t)i + 2] - s2;
    }
  }
}

// Split heights, dense phase: one wave per row i of block J, ii = i − start(J):
// forward x_i = Σ_{l ≤ ii} M[ii][l] r_l;  backward x_i = Σ_{l ≥ ii} Mᵀ[ii][l] r_l.
__global__ __launch_bounds__(256) void lp_dense_kernel(LpArgs A, const double* __restrict__ g,
                                                       const uint32_t* __restrict__ block_of,
                                                       const uint32_t* __restrict__ rows, uint32_t count, int backward,
                                                       const double* __restrict__ rbuf, double* __restrict__ x) {
  const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= count) return;
  const uint32_t i = rows[w];
  const uint32_t J = block_of[i], j0 = A.blk_start[J], k = A.blk_rows[J], ii = i - j0;
  const double* row = g + A.blk_inv[J] + (size_t)ii * k;
  const uint32_t l0 = backward ? ii : 0, l1 = backward ? k : ii + 1;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t l = l0 + lane; l < l1; l += 64) {
    const double m = row[l];
    const double* r = rbuf + 3 * (size_t)(j0 + l);
    s0 += m * r[0];
    s1 += m * r[1];
    s2 += m * r[2];
  }
  s0 = wsum64d(s0);
  s1 = wsum64d(s1);
  s2 = wsum64d(s2);
  if (lane == 0) {
    x[3 * (size_t)i] = s0;
    x[3 * (size_t)i + 1] = s1;
    x[3 * (size_t)i + 2] = s2;
  }
}

// The multiplier row's forward residual (its row of L spans every column): kLongParts workgroups
// each sum a strided share into parts[3 × part]; the final kernel adds the parts in order and
// writes x_n = b_n − Σ (its block is 1 × 1 with inverse 1).
__global__ __launch_bounds__(256) void lp_long_partial_kernel(LpArgs A, uint32_t i, const double* __restrict__ x,
                                                              double* __restrict__ parts) {
  __shared__ double red[16];
  double s0, s1, s2;
  const uint32_t q0 = A.ptr[i], q1 = A.split[i];
  const uint32_t per = (q1 - q0 + kLongParts - 1) / kLongParts;
  const uint32_t lo = min(q1, q0 + blockIdx.x * per), hi = min(q1, lo + per);
  block_dot3(A, lo, hi, x, red, s0, s1, s2);
  if (threadIdx.x == 0) {
    parts[3 * blockIdx.x] = s0;
    parts[3 * blockIdx.x + 1] = s1;
    parts[3 * blockIdx.x + 2] = s2;
  }
}

__global__ void lp_long_final_kernel(uint32_t i, const double* __restrict__ parts, const double* __restrict__ b,
                                     double* __restrict__ x) {
  const int a = threadIdx.x;
  if (a >= 3) return;
  double s = 0.0;
  for (uint32_t p = 0; p < kLongParts; ++p) s += parts[3 * p + a];
  x[3 * (size_t)i + a] = b[3 * (size_t)i + a] - s;
}

LprimePlan make_lprime_schedule(const Factor& f, const NdBlocks& nd) {
  LprimePlan P;
  const size_t n = f.n;
  auto& bstart = P.bstart;
  auto& bend = P.bend;
  auto& bheight = P.bheight;
  auto& max_rows = P.max_rows;
  auto& level_off = P.level_off;
  // blocks: the separator tree's, then the multiplier row (the root)
  bstart = nd.start;
  bend = nd.end;
  std::vector<uint32_t> bparent = nd.parent;
  const uint32_t nb0 = (uint32_t)bstart.size();
  bstart.push_back((uint32_t)n - 1);
  bend.push_back((uint32_t)n);
  bparent.push_back(kNone);
  for (uint32_t b = 0; b < nb0; ++b)
    if (bparent[b] == kNone) bparent[b] = nb0;
  const uint32_t nb = nb0 + 1;
  // heights: a domain's two halves get ids after the domain's, so decreasing ids visit children
  // before parents; the multiplier row (id nb0) is the parent of the tree's root (id 0)
  bheight.assign(nb, 0);
  for (uint32_t b = nb0; b-- > 0;) bheight[bparent[b]] = std::max(bheight[bparent[b]], bheight[b] + 1);
  std::vector<uint32_t> block_of(n);
  for (uint32_t b = 0; b < nb; ++b)
    for (uint32_t i = bstart[b]; i < bend[b]; ++i) block_of[i] = b;
  // off-block split points; every off-block entry must meet a block of another height on the
  // right side (descendants below, ancestors above), else the schedule would race
  auto& rs = P.rsplit;
  auto& cs = P.csplit;
  rs.resize(n);
  cs.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t b = block_of[i];
    uint32_t q = f.rp[i];
    while (q < f.rp[i + 1] && f.rj[q] < bstart[b]) {
      if (bheight[block_of[f.rj[q]]] >= bheight[b]) throw Error(DDGRS_GPU, "preconditioner", "separator tree schedule violated (row)");
      ++q;
    }
    rs[i] = q;
    q = f.cp[i];
    while (q < f.cp[i + 1] && f.ci[q] < bend[b]) ++q;
    cs[i] = q;
    for (uint32_t e = q; e < f.cp[i + 1]; ++e)
      if (bheight[block_of[f.ci[e]]] <= bheight[b]) throw Error(DDGRS_GPU, "preconditioner", "separator tree schedule violated (column)");
  }
  // offsets of the dense inverses M_J = L_JJ⁻¹ (row-major k × k)
  P.inv_off.assign(nb + 1, 0);
  max_rows = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    const uint64_t k = bend[b] - bstart[b];
    P.inv_off[b + 1] = P.inv_off[b] + k * k;
    max_rows = std::max(max_rows, (uint32_t)k);
  }
  // work items by height
  const uint32_t H = *std::max_element(bheight.begin(), bheight.end()) + 1;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> by(H);
  for (uint32_t b = 0; b < nb; ++b)
    for (uint32_t r0 = 0; r0 < bend[b] - bstart[b]; r0 += kChunk) by[bheight[b]].push_back({b, r0});
  auto& iblock = P.item_block;
  auto& irow0 = P.item_row0;
  auto& brows = P.rows;
  brows.resize(nb);
  level_off.assign(1, 0);
  for (uint32_t h = 0; h < H; ++h) {
    for (auto [b, r0] : by[h]) {
      iblock.push_back(b);
      irow0.push_back(r0);
    }
    level_off.push_back((uint32_t)iblock.size());
  }
  for (uint32_t b = 0; b < nb; ++b) brows[b] = bend[b] - bstart[b];
  // row → block, fused or split heights
  P.block_of = block_of;
  P.fused.assign(H, 1);
  for (uint32_t b = 0; b < nb; ++b)
    if (brows[b] > kChunk) P.fused[bheight[b]] = 0;
  P.split_off.assign(1, 0);
  for (uint32_t h = 0; h < H; ++h) {
    if (!P.fused[h])
      for (auto [b, r0] : by[h])
        for (uint32_t i = bstart[b] + r0; i < std::min(bstart[b] + r0 + kChunk, bend[b]); ++i) P.split_rows.push_back(i);
    P.split_off.push_back((uint32_t)P.split_rows.size());
  }
  return P;
}

// Dense inverses of the diagonal blocks (and their transposes) from the rows of L.
void host_block_inverses(const Factor& f, LprimePlan& P) {
  const uint32_t nb = (uint32_t)P.bstart.size();
  const auto& bstart = P.bstart;
  const auto& bend = P.bend;
  const auto& rs = P.rsplit;
  const auto& ioff = P.inv_off;
  auto& M = P.inv;
  M.assign(ioff[nb], 0.0);
  auto invert_column = [&](uint32_t b, const std::vector<double>& L, uint32_t c) {
    const uint32_t k = bend[b] - bstart[b];
    double* m = M.data() + ioff[b];
    m[(size_t)c * k + c] = 1.0;
    for (uint32_t i = c + 1; i < k; ++i) {
      double acc = 0.0;
      for (uint32_t j = c; j < i; ++j) acc += L[(size_t)i * k + j] * m[(size_t)j * k + c];
      m[(size_t)i * k + c] = -acc;
    }
  };
  auto dense_l = [&](uint32_t b) {
    const uint32_t k = bend[b] - bstart[b], j0 = bstart[b];
    std::vector<double> L((size_t)k * k, 0.0);
    for (uint32_t ii = 0; ii < k; ++ii)
      for (uint32_t q = rs[j0 + ii]; q < f.rp[j0 + ii + 1]; ++q) L[(size_t)ii * k + (f.rj[q] - j0)] = f.rx[q];
    return L;
  };
  std::vector<uint32_t> small, big;
  for (uint32_t b = 0; b < nb; ++b) (bend[b] - bstart[b] >= 256 ? big : small).push_back(b);
  parallel(small.size(), [&](size_t i) {
    const uint32_t b = small[i], k = bend[b] - bstart[b];
    auto L = dense_l(b);
    for (uint32_t c = 0; c < k; ++c) invert_column(b, L, c);
  });
  for (uint32_t b : big) {
    auto L = dense_l(b);
    parallel(bend[b] - bstart[b], [&](size_t c) { invert_column(b, L, (uint32_t)c); });
  }
  P.inv_t.assign(M.size(), 0.0);
  for (uint32_t b = 0; b < nb; ++b) {
    const uint32_t k = bend[b] - bstart[b];
    const double* m = M.data() + ioff[b];
    double* mt = P.inv_t.data() + ioff[b];
    for (uint32_t i = 0; i < k; ++i)
      for (uint32_t j = 0; j < k; ++j) mt[(size_t)j * k + i] = m[(size_t)i * k + j];
  }
}

LprimePlan make_lprime_plan(const Factor& f, const NdBlocks& nd) {
  LprimePlan P = make_lprime_schedule(f, nd);
  host_block_inverses(f, P);
  return P;
}

void LprimeDev::setup(ddgrs_ctx* ctx, const Factor& f, const LprimePlan& P) {
  n = f.n;
  level_off = P.level_off;
  max_rows = P.max_rows;
  ddgrs::upload(ctx, perm, f.perm);
  ddgrs::upload(ctx, rp, f.rp);
  ddgrs::upload(ctx, rj, f.rj);
  ddgrs::upload(ctx, cp, f.cp);
  ddgrs::upload(ctx, ci, f.ci);
  ddgrs::upload(ctx, rsplit, P.rsplit);
  ddgrs::upload(ctx, csplit, P.csplit);
  ddgrs::upload(ctx, item_block, P.item_block);
  ddgrs::upload(ctx, item_row0, P.item_row0);
  ddgrs::upload(ctx, blk_start, P.bstart);
  ddgrs::upload(ctx, blk_rows, P.rows);
  ddgrs::upload(ctx, blk_inv, P.inv_off);
  ddgrs::upload(ctx, block_of, P.block_of);
  ddgrs::upload(ctx, split_rows, P.split_rows);
  split_off = P.split_off;
  fused = P.fused;
  rx.reserve(std::max<size_t>(f.rj.size(), 1));
  cx.reserve(std::max<size_t>(f.ci.size(), 1));
  d.reserve(n);
  inv.reserve(std::max<uint64_t>(P.inv_off.back(), 1));
  inv_t.reserve(std::max<uint64_t>(P.inv_off.back(), 1));
  xb.reserve(3 * n);
  xf.reserve(3 * n);
  rbuf.reserve(3 * n);
  parts.reserve(3 * kLongParts);
}

void LprimeDev::upload(ddgrs_ctx* ctx, const Factor& f, const LprimePlan& P) {
  setup(ctx, f, P);
  ddgrs::upload(ctx, rx, f.rx);
  ddgrs::upload(ctx, cx, f.cx);
  ddgrs::upload(ctx, d, f.d);
  ddgrs::upload(ctx, inv, P.inv);
  ddgrs::upload(ctx, inv_t, P.inv_t);
}

void LprimeDev::solve(ddgrs_ctx* ctx, const double* r, double* x) {
  const uint32_t un = (uint32_t)n, unv = (uint32_t)nv;
  launch(ctx, "lp_gather", lp_gather_kernel, dim3(blocks_for(3 * n, 256)), dim3(256), 0, un, unv,
         (const uint32_t*)perm.ptr, r, xb.ptr);
  LpArgs fa{item_block.ptr, item_row0.ptr, blk_start.ptr, blk_rows.ptr, blk_inv.ptr, inv.ptr,
            rp.ptr,         rsplit.ptr,    rj.ptr,        rx.ptr,       d.ptr};
  LpArgs ba{item_block.ptr, item_row0.ptr, blk_start.ptr, blk_rows.ptr, blk_inv.ptr, inv.ptr,
            cp.ptr,         csplit.ptr,    ci.ptr,        cx.ptr,       d.ptr};
  const uint32_t H = (uint32_t)level_off.size() - 1;
  const size_t lds_f = (16 + 3 * (size_t)kChunk) * sizeof(double);
  const size_t lds_b = (16 + 768 + 3 * (size_t)kChunk) * sizeof(double);
  // forward: heights bottom-up; the top height is the multiplier row (row n − 1)
  for (uint32_t h = 0; h < H; ++h) {
    if (h == H - 1) {
      launch(ctx, "lp_long_partial", lp_long_partial_kernel, dim3(kLongParts), dim3(256), 0, fa, un - 1,
             (const double*)xf.ptr, parts.ptr);
      launch(ctx, "lp_long_final", lp_long_final_kernel, dim3(1), dim3(64), 0, un - 1, (const double*)parts.ptr,
             (const double*)xb.ptr, xf.ptr);
    } else if (fused[h]) {
      launch(ctx, "lp_forward", lp_forward_kernel, dim3(level_off[h + 1] - level_off[h]), dim3(256), lds_f, fa,
             level_off[h], (const double*)xb.ptr, xf.ptr);
    } else {
      const uint32_t cnt = split_off[h + 1] - split_off[h];
      const uint32_t* rows = split_rows.ptr + split_off[h];
      launch(ctx, "lp_residual", lp_residual_kernel, dim3(blocks_for(cnt, 4)), dim3(256), 0, fa, rows, cnt, 0,
             (const double*)xb.ptr, (const double*)xf.ptr, rbuf.ptr);
      launch(ctx, "lp_dense", lp_dense_kernel, dim3(blocks_for(cnt, 4)), dim3(256), 0, fa, (const double*)inv.ptr,
             (const uint32_t*)block_of.ptr, rows, cnt, 0, (const double*)rbuf.ptr, xf.ptr);
    }
  }
  // backward: heights top-down (the multiplier row is a fused 1 × 1 block)
  for (uint32_t h = H; h-- > 0;) {
    if (fused[h] || h == H - 1) {
      launch(ctx, "lp_backward", lp_backward_kernel, dim3(level_off[h + 1] - level_off[h]), dim3(256), lds_b, ba,
             level_off[h], (const double*)xf.ptr, xb.ptr);
    } else {
      const uint32_t cnt = split_off[h + 1] - split_off[h];
      const uint32_t* rows = split_rows.ptr + split_off[h];
      launch(ctx, "lp_residual", lp_residual_kernel, dim3(blocks_for(cnt, 4)), dim3(256), 0, ba, rows, cnt, 1,
             (const double*)xf.ptr, (const double*)xb.ptr, rbuf.ptr);
      launch(ctx, "lp_dense", lp_dense_kernel, dim3(blocks_for(cnt, 4)), dim3(256), 0, ba, (const double*)inv_t.ptr,
             (const uint32_t*)block_of.ptr, rows, cnt, 1, (const double*)rbuf.ptr, xb.ptr);
    }
  }
  launch(ctx, "lp_scatter", lp_scatter_kernel, dim3(blocks_for(3 * n, 256)), dim3(256), 0, un, unv,
         (const uint32_t*)perm.ptr, (const double*)xb.ptr, x);
}

// Augmented form: v = s·v − (K ⊗ I₃) y on the positions, 0 on the multipliers (precond.rs:1046-1066);
// congruence form: only the multipliers' zeros.
__global__ __launch_bounds__(256) void pc_augment_kernel(uint32_t nv, const uint32_t* __restrict__ kp,
                                                         const uint32_t* __restrict__ kj, const double* __restrict__ kx,
                                                         double scale, const double* __restrict__ y,
                                                         double* __restrict__ v, int augmented) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * nv + 3) return;
  if (t >= 3 * nv) {
    v[t] = 0.0;
    return;
  }
  if (!augmented) return;  // congruence form: L^V y as it stands
  uint32_t j = t / 3, a = t % 3;
  double acc = 0.0;
  for (uint32_t q = kp[j]; q < kp[j + 1]; ++q) acc += kx[q] * y[3 * (size_t)kj[q] + a];
  v[t] = scale * v[t] - acc;
}

__global__ __launch_bounds__(256) void pc_add_kernel(uint32_t n, const double* __restrict__ y, double* __restrict__ z) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t < n) z[t] += y[t];
}

__device__ inline double pc_sum256(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// The refinement residual r − L′ x, L′ = [[K ⊗ I₃, cᵀ], [c, 0]] per axis: the position rows
// r − (K x + c_v λ) here, and per workgroup partial sums of c · x per axis for the multiplier rows.
__global__ __launch_bounds__(256) void pc_lres_kernel(uint32_t nv, const uint32_t* __restrict__ kp,
                                                      const uint32_t* __restrict__ kj, const double* __restrict__ kx,
                                                      const double* __restrict__ crow, const double* __restrict__ r,
                                                      const double* __restrict__ x, double* __restrict__ res,
                                                      double* __restrict__ partial) {
  __shared__ double sh[256];
  const size_t nd = 3 * (size_t)nv;
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  if (j < nv) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t q = kp[j]; q < kp[j + 1]; ++q) {
      const double k = kx[q];
      const size_t c = 3 * (size_t)kj[q];
      for (int a = 0; a < 3; ++a) acc[a] += k * x[c + a];
    }
    const double c = crow[j];
    for (int a = 0; a < 3; ++a) {
      res[3 * (size_t)j + a] = r[3 * (size_t)j + a] - (acc[a] + c * x[nd + a]);
      s[a] = c * x[3 * (size_t)j + a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    const double t = pc_sum256(s[a], sh);
    if (threadIdx.x == 0) partial[3 * (size_t)blockIdx.x + a] = t;
  }
}

__global__ void pc_lres_final_kernel(uint32_t nblocks, uint32_t nv, const double* __restrict__ partial,
                                     const double* __restrict__ r, double* __restrict__ res) {
  const int a = threadIdx.x;
  if (a >= 3) return;
  double s = 0.0;
  for (uint32_t b = 0; b < nblocks; ++b) s += partial[3 * (size_t)b + a];
  res[3 * (size_t)nv + a] = r[3 * (size_t)nv + a] - s;
}

HostCsr cotan_stiffness(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                        const std::vector<double>& mass, double eps_rel, double* eps) {
  const size_t nf = faces.size() / 3;
  // vertex → incident faces, ascending
  std::vector<uint32_t> vptr(nv + 1, 0), vinc(3 * nf);
  for (size_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) vptr[faces[3 * f + c] + 1]++;
  for (size_t v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  {
    std::vector<uint32_t> fill(vptr.begin(), vptr.end() - 1);
    for (size_t f = 0; f < nf; ++f)
      for (int c = 0; c < 3; ++c) vinc[fill[faces[3 * f + c]]++] = (uint32_t)f;
  }
  // row patterns: the vertex and its neighbours, ascending
  HostCsr k;
  k.n = nv;
  k.ptr.assign(nv + 1, 0);
  std::vector<std::vector<uint32_t>> pat(nv);
  parallel(nv, [&](size_t a) {
    auto& p = pat[a];
    p.push_back((uint32_t)a);
    for (uint32_t q = vptr[a]; q < vptr[a + 1]; ++q)
      for (int c = 0; c < 3; ++c) p.push_back(faces[3 * (size_t)vinc[q] + c]);
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
  });
  for (size_t a = 0; a < nv; ++a) k.ptr[a + 1] = k.ptr[a] + (uint32_t)pat[a].size();
  k.idx.resize(k.ptr[nv]);
  k.val.assign(k.ptr[nv], 0.0);
  // values: for each incident face (ascending) and corner c, the edge opposite c gets
  // −½ cot(angle at c) off the diagonal and +½ cot on it, summed in the triplets' order
  parallel(nv, [&](size_t a) {
    const uint32_t r0 = k.ptr[a];
    std::copy(pat[a].begin(), pat[a].end(), k.idx.begin() + r0);
    auto at = [&](uint32_t col) -> double& {
      const uint32_t* b = k.idx.data() + r0;
      return k.val[r0 + (std::lower_bound(b, b + pat[a].size(), col) - b)];
    };
    for (uint32_t q = vptr[a]; q < vptr[a + 1]; ++q) {
      const uint32_t* t = &faces[3 * (size_t)vinc[q]];
      for (int c = 0; c < 3; ++c) {
        const uint32_t x = t[(c + 1) % 3], y = t[(c + 2) % 3];
        if (x != a && y != a) continue;
        const double* p0 = &pos[3 * (size_t)t[c]];
        const double* p1 = &pos[3 * (size_t)x];
        const double* p2 = &pos[3 * (size_t)y];
        double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        double cot = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        double h = 0.5 * cot;
        at(x == a ? y : x) += -h;
        at((uint32_t)a) += h;
      }
    }
  });
  double trw = 0.0, trm = 0.0;
  std::vector<uint32_t> diag(nv);
  for (size_t a = 0; a < nv; ++a) {
    const uint32_t* b = k.idx.data() + k.ptr[a];
    diag[a] = k.ptr[a] + (uint32_t)(std::lower_bound(b, b + (k.ptr[a + 1] - k.ptr[a]), (uint32_t)a) - b);
    trw += k.val[diag[a]];
  }
  for (size_t v = 0; v < nv; ++v) trm += mass[v];
  *eps = eps_rel * trw / trm;
  for (size_t v = 0; v < nv; ++v) k.val[diag[v]] += *eps * mass[v];
  return k;
}

PrecondPlan plan_preconditioner(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                                const std::vector<double>& mass, double s_exp, double w_tp, double eps_rel) {
  PrecondPlan P;
  P.nv = nv;
  P.k = cotan_stiffness(faces, nv, pos, mass, eps_rel, &P.eps);
  const HostCsr& k = P.k;
  P.crow.resize(nv);
  double tot = 0.0;
  for (size_t v = 0; v < nv; ++v) tot += mass[v];
  for (size_t v = 0; v < nv; ++v) P.crow[v] = mass[v] / tot;
  // L′ = [[K, c], [cᵀ, 0]] (both triangles), the multiplier last
  HostCsr lpm;
  lpm.n = nv + 1;
  lpm.ptr.assign(nv + 2, 0);
  lpm.idx.reserve(k.idx.size() + 2 * nv);
  lpm.val.reserve(k.idx.size() + 2 * nv);
  for (size_t v = 0; v < nv; ++v) {
    for (uint32_t q = k.ptr[v]; q < k.ptr[v + 1]; ++q) {
      lpm.idx.push_back(k.idx[q]);
      lpm.val.push_back(k.val[q]);
    }
    lpm.idx.push_back((uint32_t)nv);
    lpm.val.push_back(P.crow[v]);
    lpm.ptr[v + 1] = (uint32_t)lpm.idx.size();
  }
  for (size_t v = 0; v < nv; ++v) {
    lpm.idx.push_back((uint32_t)v);
    lpm.val.push_back(P.crow[v]);
  }
  lpm.ptr[nv + 1] = (uint32_t)lpm.idx.size();
  NdBlocks nd = nested_dissection(k, pos);
  P.f = factor_multifrontal(lpm, nd);
  P.lp = make_lprime_plan(P.f, nd);
  // the scale of L̃^V_{2−s}: 1/(w_TP 4 I(σ) I(2−s)) (precond.rs:528-564, :893-906)
  auto I = [](double t) { return M_PI * std::tgamma(1.0 - t) / (t * std::pow(4.0, t) * std::tgamma(1.0 + t)); };
  P.frac_scale = 1.0 / (4.0 * I(s_exp - 1.0) * I(2.0 - s_exp)) / w_tp;
  return P;
}

Preconditioner::Preconditioner() = default;
Preconditioner::~Preconditioner() = default;

namespace {
double fractional_scale(double s_exp, double w_tp) {
  // the scale of L̃^V_{2−s}: 1/(w_TP 4 I(σ) I(2−s)) (precond.rs:528-564, :893-906)
  auto I = [](double t) { return M_PI * std::tgamma(1.0 - t) / (t * std::pow(4.0, t) * std::tgamma(1.0 + t)); };
  return 1.0 / (4.0 * I(s_exp - 1.0) * I(2.0 - s_exp)) / w_tp;
}
bool lprime_on_host() {
  const char* e = std::getenv("DDGRS_LPRIME_ON_HOST");
  return e && e[0] == '1';
}
}  // namespace

void Preconditioner::build(ddgrs_surface* s, double s_exp, double w_tp, double eps_rel, double chi, size_t leaf_size,
                           size_t leaf_pairs) {
  std::vector<double> pos, mass;
  vertex_areas_host(s, &pos, &mass);
  GpuBvhBuilder b;
  Topology vt = build_vertex_tree(s, leaf_size, b);
  if (lprime_on_host()) {
    build(s->ctx, plan_preconditioner(s->faces_h, (size_t)s->nv, pos, mass, s_exp, w_tp, eps_rel));
  } else {
    setup_device(s->ctx, lprime_symbolic(s->faces_h, (size_t)s->nv, pos, s->ctx->prop.sharedMemPerBlock));
    factor_device(s->ctx, s, s_exp, w_tp, eps_rel);
    s->ctx->sync();
    check_factor(s->ctx);
  }
  build_fractional(s->ctx, pos, mass, 2.0 - s_exp, chi, leaf_pairs, std::move(vt));
}

void Preconditioner::build(ddgrs_ctx* ctx, const PrecondPlan& P) {
  nv = P.nv;
  n = 3 * nv + 3;
  eps = P.eps;
  frac_scale = P.frac_scale;
  crow = P.crow;
  lp.nv = nv;
  lp.upload(ctx, P.f, P.lp);
  ddgrs::upload(ctx, kp, P.k.ptr);
  ddgrs::upload(ctx, kj, P.k.idx);
  ddgrs::upload(ctx, kx, P.k.val);
  ddgrs::upload(ctx, crow_d, crow);
  k_ptr = kp.ptr;
  k_idx = kj.ptr;
  k_val = kx.ptr;
  crow_dev = crow_d.ptr;
  fac.reset();
  y.reserve(n);
  v.reserve(n);
  ctx->sync();  // the plan's host buffers may be released after return
}

void Preconditioner::setup_device(ddgrs_ctx* ctx, const LprimeSymbolic& sym) {
  nv = sym.nv;
  n = 3 * nv + 3;
  fac = std::make_unique<LprimeFactorDev>();
  fac->setup(ctx, sym);
  k_ptr = fac->kp.ptr;
  k_idx = fac->kj.ptr;
  k_val = fac->kx.ptr;
  crow_dev = fac->crow.ptr;
  y.reserve(n);
  v.reserve(n);
}

void Preconditioner::factor_device(ddgrs_ctx* ctx, ddgrs_surface* s, double s_exp, double w_tp, double eps_rel) {
  if (!fac) throw Error(DDGRS_GPU, "preconditioner", "device factorization used before its symbolic set-up");
  frac_scale = fractional_scale(s_exp, w_tp);
  fac->factor(ctx, s, eps_rel);
  eps = -1.0;  // read on demand
}

void Preconditioner::check_factor(ddgrs_ctx* ctx) {
  if (fac) fac->check(ctx);
}

double Preconditioner::eps_value(ddgrs_ctx* ctx) {
  if (fac && eps < 0.0) eps = fac->eps(ctx);
  return eps;
}

void Preconditioner::build_fractional(ddgrs_ctx* ctx, const std::vector<double>& pos, const std::vector<double>& mass,
                                      double tau, double chi, size_t leaf_pairs, Topology vertex_tree) {
  vf.build(ctx, pos, mass, tau, chi, leaf_pairs, std::move(vertex_tree));
}

void Preconditioner::lprime_solve(ddgrs_ctx* ctx, const double* r, double* x) {
  if (fac)
    fac->solve(ctx, r, x);
  else
    lp.solve(ctx, r, x);
}

void Preconditioner::lprime_solve_refined(ddgrs_ctx* ctx, const double* r, double* x) {
  lprime_solve(ctx, r, x);
  if (!refine) return;
  // x ← x + L′⁻¹ (r − L′ x) (precond.rs:1031 solve_many_on with refine)
  const unsigned nb = blocks_for((int64_t)nv, 256);
  res.reserve(n);
  dx.reserve(n);
  part.reserve(3 * (size_t)nb);
  launch(ctx, "pc_lres", pc_lres_kernel, dim3(nb), dim3(256), 0, (uint32_t)nv, k_ptr, k_idx, k_val, crow_dev, r,
         (const double*)x, res.ptr, part.ptr);
  launch(ctx, "pc_lres_final", pc_lres_final_kernel, dim3(1), dim3(64), 0, (uint32_t)nb, (uint32_t)nv,
         (const double*)part.ptr, r, res.ptr);
  lprime_solve(ctx, res.ptr, dx.ptr);
  launch(ctx, "pc_add", pc_add_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (uint32_t)n, (const double*)dx.ptr, x);
}

void Preconditioner::apply(ddgrs_ctx* ctx, const double* r, double* z) {
  lprime_solve_refined(ctx, r, y.ptr);
  vf.apply(ctx, y.ptr, v.ptr);
  launch(ctx, "pc_augment", pc_augment_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (uint32_t)nv, k_ptr, k_idx,
         k_val, scale(), (const double*)y.ptr, v.ptr, (int)(form == kFormAugmented));
  lprime_solve_refined(ctx, v.ptr, z);
  if (form == kFormAugmented)
    launch(ctx, "pc_add", pc_add_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (uint32_t)n, (const double*)y.ptr, z);
}

}  // namespace ddgrs
