// The reference preconditioner on the device (see precond.hpp).
//
// Host: K = W + εM (RS/src/geometry.rs:221 cotan_stiffness, RS/src/solve/dense.rs:529 ε = ε_rel
// tr W / tr M), the scalar saddle L′ of the global barycenter row (precond.rs:184-240, per-axis
// case, multipliers ordered last), its LDLᵀ factor in a nested-dissection order, and the
// elimination-tree levels that schedule the device solves.
//
// Device: the three axes of a saddle vector are solved together in the permuted numbering (n × 3
// row-major).  Forward substitution gathers a row of L (x_j = b_j − Σ_k L_jk x_k, the k are
// descendants of j), backward substitution gathers a column (x_j −= Σ_i L_ij x_i, the i are
// ancestors), so a level of the elimination tree is a set of independent rows.  Wide levels (the
// bottom of the tree) are one grid launch each, one thread per (row, axis); the narrow top of the
// tree (separators and the dense multiplier row) is solved by one workgroup that walks its levels
// with a barrier between them, one wave per row.  FP64 throughout: L′ has a near-null space of
// relative size ε = 1e-10.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "bh.hpp"
#include "precond.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

namespace {
constexpr uint32_t kNone = UINT32_MAX;
constexpr size_t kWideMin = 2048;  // rows of a level solved by a whole-grid launch

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

std::vector<uint32_t> nested_dissection_order(const HostCsr& k, const std::vector<double>& pos) {
  const size_t n = k.n;
  std::vector<uint32_t> out;
  out.reserve(n);
  std::vector<int32_t> side(n, -1);
  struct Task {
    std::vector<uint32_t> verts;
    bool emit_only;  // a separator, appended after its two domains
  };
  // Iterative post-order: a task expands into [domain A, domain B, separator].
  std::vector<Task> stack;
  std::vector<uint32_t> all(n);
  std::iota(all.begin(), all.end(), 0u);
  stack.push_back({std::move(all), false});
  while (!stack.empty()) {
    Task t = std::move(stack.back());
    stack.pop_back();
    if (t.emit_only || t.verts.size() <= 64) {
      out.insert(out.end(), t.verts.begin(), t.verts.end());
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
    stack.push_back({std::move(S), true});
    stack.push_back({std::move(B), false});
    stack.push_back({std::move(A), false});
  }
  return out;
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
  f.parent = parent;
  f.level.assign(n, 0);
  for (size_t j = 0; j < n; ++j)
    if (parent[j] != kNone) f.level[parent[j]] = std::max(f.level[parent[j]], f.level[j] + 1);
  f.n_levels = n ? *std::max_element(f.level.begin(), f.level.end()) + 1 : 0;
  return f;
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

// One wide level: thread per (row, axis); forward uses the row of L, backward the column.
__global__ __launch_bounds__(256) void lp_level_kernel(uint32_t count, const uint32_t* __restrict__ rows,
                                                       const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ idx,
                                                       const double* __restrict__ val, double* __restrict__ xb) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * count) return;
  uint32_t j = rows[t / 3], a = t % 3;
  double s = 0.0;
  for (uint32_t q = ptr[j]; q < ptr[j + 1]; ++q) s += val[q] * xb[3 * (size_t)idx[q] + a];
  xb[3 * (size_t)j + a] -= s;
}

// The narrow levels in one workgroup: levels in order (forward: bottom-up; backward: top-down), a
// barrier between levels, one wave per row (lanes split the terms, fixed-order reduction).
__global__ __launch_bounds__(1024) void lp_narrow_kernel(uint32_t n_levels, const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ rows, int backward,
                                                         const uint32_t* __restrict__ ptr,
                                                         const uint32_t* __restrict__ idx,
                                                         const double* __restrict__ val, double* __restrict__ xb) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t l = 0; l < n_levels; ++l) {
    const uint32_t lv = backward ? n_levels - 1 - l : l;
    for (uint32_t q = off[lv] + wave; q < off[lv + 1]; q += 16) {
      const uint32_t j = rows[q];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (uint32_t e = ptr[j] + lane; e < ptr[j + 1]; e += 64) {
        const double v = val[e];
        const double* xk = xb + 3 * (size_t)idx[e];
        s0 += v * xk[0];
        s1 += v * xk[1];
        s2 += v * xk[2];
      }
      for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
      }
      if (lane == 0) {
        xb[3 * (size_t)j] -= s0;
        xb[3 * (size_t)j + 1] -= s1;
        xb[3 * (size_t)j + 2] -= s2;
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void lp_diag_kernel(uint32_t n, const double* __restrict__ d, double* __restrict__ xb) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t < 3 * n) xb[t] /= d[t / 3];
}

void LprimeDev::upload(ddgrs_ctx* ctx) {
  n = f.n;
  ddgrs::upload(ctx, perm, f.perm);
  ddgrs::upload(ctx, rp, f.rp);
  ddgrs::upload(ctx, rj, f.rj);
  ddgrs::upload(ctx, rx, f.rx);
  ddgrs::upload(ctx, cp, f.cp);
  ddgrs::upload(ctx, ci, f.ci);
  ddgrs::upload(ctx, cx, f.cx);
  ddgrs::upload(ctx, d, f.d);
  std::vector<std::vector<uint32_t>> by(f.n_levels);
  for (size_t j = 0; j < n; ++j) by[f.level[j]].push_back((uint32_t)j);
  uint32_t first_narrow = 0;
  while (first_narrow < f.n_levels && by[first_narrow].size() >= kWideMin) ++first_narrow;
  std::vector<uint32_t> wide, narrow;
  wide_off.assign(1, 0);
  narrow_off.assign(1, 0);
  for (uint32_t l = 0; l < f.n_levels; ++l) {
    auto& dst = l < first_narrow ? wide : narrow;
    dst.insert(dst.end(), by[l].begin(), by[l].end());
    (l < first_narrow ? wide_off : narrow_off).push_back((uint32_t)dst.size());
  }
  ddgrs::upload(ctx, wide_rows, wide);
  ddgrs::upload(ctx, narrow_rows, narrow);
  ddgrs::upload(ctx, narrow_off_d, narrow_off);
  xb.reserve(3 * n);
}

void LprimeDev::solve(ddgrs_ctx* ctx, const double* r, double* x) {
  const uint32_t un = (uint32_t)n, unv = (uint32_t)nv;
  launch(ctx, "lp_gather", lp_gather_kernel, dim3(blocks_for(3 * n, 256)), dim3(256), 0, un, unv,
         (const uint32_t*)perm.ptr, r, xb.ptr);
  const uint32_t nw = (uint32_t)wide_off.size() - 1, nn = (uint32_t)narrow_off.size() - 1;
  for (uint32_t l = 0; l < nw; ++l) {
    uint32_t cnt = wide_off[l + 1] - wide_off[l];
    launch(ctx, "lp_forward_level", lp_level_kernel, dim3(blocks_for(3 * (int64_t)cnt, 256)), dim3(256), 0, cnt,
           (const uint32_t*)(wide_rows.ptr + wide_off[l]), (const uint32_t*)rp.ptr, (const uint32_t*)rj.ptr,
           (const double*)rx.ptr, xb.ptr);
  }
  if (nn > 0)
    launch(ctx, "lp_forward_narrow", lp_narrow_kernel, dim3(1), dim3(1024), 0, nn, (const uint32_t*)narrow_off_d.ptr,
           (const uint32_t*)narrow_rows.ptr, 0, (const uint32_t*)rp.ptr, (const uint32_t*)rj.ptr, (const double*)rx.ptr,
           xb.ptr);
  launch(ctx, "lp_diag", lp_diag_kernel, dim3(blocks_for(3 * n, 256)), dim3(256), 0, un, (const double*)d.ptr, xb.ptr);
  if (nn > 0)
    launch(ctx, "lp_backward_narrow", lp_narrow_kernel, dim3(1), dim3(1024), 0, nn, (const uint32_t*)narrow_off_d.ptr,
           (const uint32_t*)narrow_rows.ptr, 1, (const uint32_t*)cp.ptr, (const uint32_t*)ci.ptr, (const double*)cx.ptr,
           xb.ptr);
  for (uint32_t l = nw; l-- > 0;) {
    uint32_t cnt = wide_off[l + 1] - wide_off[l];
    launch(ctx, "lp_backward_level", lp_level_kernel, dim3(blocks_for(3 * (int64_t)cnt, 256)), dim3(256), 0, cnt,
           (const uint32_t*)(wide_rows.ptr + wide_off[l]), (const uint32_t*)cp.ptr, (const uint32_t*)ci.ptr,
           (const double*)cx.ptr, xb.ptr);
  }
  launch(ctx, "lp_scatter", lp_scatter_kernel, dim3(blocks_for(3 * n, 256)), dim3(256), 0, un, unv,
         (const uint32_t*)perm.ptr, (const double*)xb.ptr, x);
}

// v = s·v − (K ⊗ I₃) y on the positions, 0 on the multipliers (precond.rs:1046-1066).
__global__ __launch_bounds__(256) void pc_augment_kernel(uint32_t nv, const uint32_t* __restrict__ kp,
                                                         const uint32_t* __restrict__ kj, const double* __restrict__ kx,
                                                         double scale, const double* __restrict__ y,
                                                         double* __restrict__ v) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * nv + 3) return;
  if (t >= 3 * nv) {
    v[t] = 0.0;
    return;
  }
  uint32_t j = t / 3, a = t % 3;
  double acc = 0.0;
  for (uint32_t q = kp[j]; q < kp[j + 1]; ++q) acc += kx[q] * y[3 * (size_t)kj[q] + a];
  v[t] = scale * v[t] - acc;
}

__global__ __launch_bounds__(256) void pc_add_kernel(uint32_t n, const double* __restrict__ y, double* __restrict__ z) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t < n) z[t] += y[t];
}

void Preconditioner::build(ddgrs_surface* s, double s_exp, double w_tp, double eps_rel, double chi, size_t leaf_size,
                           size_t leaf_pairs) {
  auto t0 = std::chrono::steady_clock::now();
  ddgrs_ctx* ctx = s->ctx;
  std::vector<double> pos, mass;
  vertex_areas_host(s, &pos, &mass);
  nv = (size_t)s->nv;
  n = 3 * nv + 3;
  // K = W + ε M (cotan stiffness from the face list)
  std::vector<std::tuple<uint32_t, uint32_t, double>> trip;
  trip.reserve(12 * (size_t)s->nf);
  for (int64_t f = 0; f < s->nf; ++f) {
    const uint32_t* t = &s->faces_h[3 * f];
    for (int c = 0; c < 3; ++c) {
      uint32_t a = t[(c + 1) % 3], b = t[(c + 2) % 3];
      const double* p0 = &pos[3 * (size_t)t[c]];
      const double* p1 = &pos[3 * (size_t)a];
      const double* p2 = &pos[3 * (size_t)b];
      double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      double cot = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      double h = 0.5 * cot;
      trip.emplace_back(a, b, -h);
      trip.emplace_back(b, a, -h);
      trip.emplace_back(a, a, h);
      trip.emplace_back(b, b, h);
    }
  }
  double trw = 0.0, trm = 0.0;
  {
    HostCsr w = csr_from_triplets(nv, trip);
    for (size_t i = 0; i < nv; ++i)
      for (uint32_t q = w.ptr[i]; q < w.ptr[i + 1]; ++q)
        if (w.idx[q] == i) trw += w.val[q];
  }
  for (size_t v = 0; v < nv; ++v) trm += mass[v];
  eps = eps_rel * trw / trm;
  for (size_t v = 0; v < nv; ++v) trip.emplace_back((uint32_t)v, (uint32_t)v, eps * mass[v]);
  auto ktrip = trip;
  HostCsr k = csr_from_triplets(nv, ktrip);
  crow.resize(nv);
  double tot = 0.0;
  for (size_t v = 0; v < nv; ++v) tot += mass[v];
  for (size_t v = 0; v < nv; ++v) crow[v] = mass[v] / tot;
  for (size_t v = 0; v < nv; ++v) {
    trip.emplace_back((uint32_t)nv, (uint32_t)v, crow[v]);
    trip.emplace_back((uint32_t)v, (uint32_t)nv, crow[v]);
  }
  HostCsr lpm = csr_from_triplets(nv + 1, trip);
  std::vector<uint32_t> order = nested_dissection_order(k, pos);
  order.push_back((uint32_t)nv);
  lp.f = factor_ldlt(lpm, order);
  lp.nv = nv;
  lp.upload(ctx);
  ddgrs::upload(ctx, kp, k.ptr);
  ddgrs::upload(ctx, kj, k.idx);
  ddgrs::upload(ctx, kx, k.val);
  ddgrs::upload(ctx, crow_d, crow);
  // L̃^V_{2−s} and its scale 1/(w_TP 4 I(σ) I(2−s)) (precond.rs:528-564, :893-906)
  auto I = [](double t) { return M_PI * std::tgamma(1.0 - t) / (t * std::pow(4.0, t) * std::tgamma(1.0 + t)); };
  frac_scale = 1.0 / (4.0 * I(s_exp - 1.0) * I(2.0 - s_exp)) / w_tp;
  vf.build(ctx, pos, mass, 2.0 - s_exp, chi, leaf_size, leaf_pairs);
  y.reserve(n);
  v.reserve(n);
  factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void Preconditioner::apply(ddgrs_ctx* ctx, const double* r, double* z) {
  lp.solve(ctx, r, y.ptr);
  vf.apply(ctx, y.ptr, v.ptr);
  launch(ctx, "pc_augment", pc_augment_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (uint32_t)nv,
         (const uint32_t*)kp.ptr, (const uint32_t*)kj.ptr, (const double*)kx.ptr, frac_scale, (const double*)y.ptr, v.ptr);
  lp.solve(ctx, v.ptr, z);
  launch(ctx, "pc_add", pc_add_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (uint32_t)n, (const double*)y.ptr, z);
}

}  // namespace ddgrs
