// L′ factored on the device (see lprime_gpu.hpp).
//
// Symbolic analysis (host, once per mesh): K's pattern (RS/src/geometry.rs:221 cotan_stiffness:
// each vertex and its neighbours), a nested-dissection order and separator tree, the rows of every
// front (the block's rows, then the later rows its columns or its children's update matrices
// reach), and the positions that the numeric phase reads and writes: where each entry of L′ lands
// in its front, where each child's update rows land in the parent's front, where each entry of L
// and D is found after the elimination.
//
// Numeric phase (device, per step): K's values in the triplet order of the host assembly, ε and the
// barycenter row by fixed-order sums, the fronts zeroed and L′'s lower triangle scattered into
// them, then per height of the tree: the fronts whose lower triangle fits in LDS are loaded,
// extended by their children's update matrices, eliminated (right-looking, one column at a time)
// and written back by one workgroup each; the larger fronts receive their children's updates in
// one launch per child slot and are eliminated by panels of kPanelCols columns (one workgroup per
// front) and tiled trailing updates (one workgroup per 64 × 64 tile).  Finally L's rows and
// columns and D are gathered from the fronts and the dense inverses of the diagonal blocks are
// formed by forward substitution, one wave per column.  FP64 throughout.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "host_parallel.hpp"
#include "kernels.hpp"
#include "lprime_gpu.hpp"

namespace ddgrs {

namespace {
constexpr uint32_t kNone = UINT32_MAX;
constexpr uint32_t kEaRows = 32;  // update rows per extend-add work item
}  // namespace

// ---- symbolic analysis --------------------------------------------------------------------------

LprimeSymbolic lprime_symbolic(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                               size_t lds_bytes) {
  LprimeSymbolic S;
  const size_t nf = faces.size() / 3, n = nv + 1;
  S.nv = nv;
  S.n = n;
  // K's pattern: the vertex and the vertices of its faces, ascending
  std::vector<uint32_t> vptr(nv + 1, 0), vinc(3 * nf);
  for (size_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) vptr[faces[3 * f + c] + 1]++;
  for (size_t v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  {
    std::vector<uint32_t> fill(vptr.begin(), vptr.end() - 1);
    for (size_t f = 0; f < nf; ++f)
      for (int c = 0; c < 3; ++c) vinc[fill[faces[3 * f + c]]++] = (uint32_t)f;
  }
  std::vector<std::vector<uint32_t>> pat(nv);
  parallel(nv, [&](size_t a) {
    auto& p = pat[a];
    p.push_back((uint32_t)a);
    for (uint32_t q = vptr[a]; q < vptr[a + 1]; ++q)
      for (int c = 0; c < 3; ++c) p.push_back(faces[3 * (size_t)vinc[q] + c]);
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
  });
  HostCsr& K = S.kpat;
  K.n = nv;
  K.ptr.assign(nv + 1, 0);
  for (size_t a = 0; a < nv; ++a) K.ptr[a + 1] = K.ptr[a] + (uint32_t)pat[a].size();
  K.idx.resize(K.ptr[nv]);
  S.kdiag.resize(nv);
  for (size_t a = 0; a < nv; ++a) {
    std::copy(pat[a].begin(), pat[a].end(), K.idx.begin() + K.ptr[a]);
    S.kdiag[a] = K.ptr[a] + (uint32_t)(std::lower_bound(pat[a].begin(), pat[a].end(), (uint32_t)a) - pat[a].begin());
  }
  S.nd = nested_dissection(K, pos);
  const NdBlocks& nd = S.nd;
  // order, the multiplier row last
  Factor& F = S.pattern;
  F.n = n;
  F.perm = nd.order;
  F.perm.push_back((uint32_t)nv);
  std::vector<uint32_t> pinv(n);
  for (size_t i = 0; i < n; ++i) pinv[F.perm[i]] = (uint32_t)i;
  // blocks: the separator tree's, then the multiplier row as the root
  const uint32_t nb0 = (uint32_t)nd.start.size(), nb = nb0 + 1;
  S.nb = nb;
  std::vector<uint32_t> bend = nd.end;
  S.bstart = nd.start;
  S.bstart.push_back((uint32_t)n - 1);
  bend.push_back((uint32_t)n);
  S.bparent = nd.parent;
  S.bparent.push_back(kNone);
  for (uint32_t b = 0; b < nb0; ++b)
    if (S.bparent[b] == kNone) S.bparent[b] = nb0;
  std::vector<std::vector<uint32_t>> children(nb);
  for (uint32_t b = 0; b < nb0; ++b) children[S.bparent[b]].push_back(b);
  S.bheight.assign(nb, 0);
  for (uint32_t b = nb0; b-- > 0;) S.bheight[S.bparent[b]] = std::max(S.bheight[S.bparent[b]], S.bheight[b] + 1);
  S.H = S.bheight[nb0] + 1;
  std::vector<uint32_t> block_of(n);
  for (uint32_t b = 0; b < nb; ++b)
    for (uint32_t i = S.bstart[b]; i < bend[b]; ++i) block_of[i] = b;
  // front rows: the block's rows, then the later rows of its columns and of its children's fronts
  // (children first: decreasing ids, the root last)
  std::vector<std::vector<uint32_t>> front(nb);
  auto symbolic = [&](uint32_t b) {
    const uint32_t j0 = S.bstart[b], j1 = bend[b];
    std::vector<uint32_t> st;
    for (uint32_t j = j0; j < j1; ++j) {
      const uint32_t o = F.perm[j];
      if (o == nv) continue;  // the multiplier column has no entries below the diagonal
      for (uint32_t q = K.ptr[o]; q < K.ptr[o + 1]; ++q) {
        const uint32_t i = pinv[K.idx[q]];
        if (i >= j1) st.push_back(i);
      }
      if ((uint32_t)(n - 1) >= j1) st.push_back((uint32_t)n - 1);
    }
    for (uint32_t c : children[b])
      for (size_t r = bend[c] - S.bstart[c]; r < front[c].size(); ++r) {
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
  S.bk.resize(nb);
  S.bf.resize(nb);
  S.foff.assign(nb + 1, 0);
  for (uint32_t b = 0; b < nb; ++b) {
    S.bk[b] = bend[b] - S.bstart[b];
    S.bf[b] = (uint32_t)front[b].size();
    S.foff[b + 1] = S.foff[b] + (uint64_t)S.bf[b] * S.bf[b];
    S.max_front = std::max(S.max_front, S.bf[b]);
    S.max_rows = std::max(S.max_rows, S.bk[b]);
  }
  S.front_total = S.foff[nb];
  S.child_ptr.assign(nb + 1, 0);
  for (uint32_t b = 0; b < nb; ++b) {
    S.child_ptr[b + 1] = S.child_ptr[b] + (uint32_t)children[b].size();
    S.child.insert(S.child.end(), children[b].begin(), children[b].end());
    S.max_children = std::max(S.max_children, (uint32_t)children[b].size());
  }
  auto local = [&](uint32_t b, uint32_t i) -> uint64_t {
    const auto& fr = front[b];
    const uint32_t k = S.bk[b];
    if (i < bend[b]) return i - S.bstart[b];
    auto it = std::lower_bound(fr.begin() + k, fr.end(), i);
    if (it == fr.end() || *it != i) throw Error(DDGRS_GPU, "preconditioner", "front row missing from the symbolic fronts");
    return (uint64_t)(it - fr.begin());
  };
  // L′'s lower triangle → front positions
  S.amap_k.assign(K.idx.size(), LprimeSymbolic::kNoPos);
  parallel(nv, [&](size_t o) {
    const uint32_t i = pinv[o];
    for (uint32_t q = K.ptr[o]; q < K.ptr[o + 1]; ++q) {
      const uint32_t j = pinv[K.idx[q]];
      if (i < j) continue;
      const uint32_t b = block_of[j];
      S.amap_k[q] = S.foff[b] + local(b, i) * S.bf[b] + (j - S.bstart[b]);
    }
  });
  S.amap_c.resize(nv);
  for (size_t v = 0; v < nv; ++v) {
    const uint32_t j = pinv[v], b = block_of[j];
    S.amap_c[v] = S.foff[b] + local(b, (uint32_t)n - 1) * S.bf[b] + (j - S.bstart[b]);
  }
  // children's update rows → parent front rows
  S.ea_off.assign(nb + 1, 0);
  for (uint32_t b = 0; b < nb; ++b) S.ea_off[b + 1] = S.ea_off[b] + (S.bf[b] - S.bk[b]);
  S.ea_map.resize(S.ea_off[nb]);
  for (uint32_t b = 0; b < nb0; ++b) {
    const uint32_t p = S.bparent[b];
    for (uint32_t r = 0; r < S.bf[b] - S.bk[b]; ++r) S.ea_map[S.ea_off[b] + r] = (uint32_t)local(p, front[b][S.bk[b] + r]);
  }
  // L's columns (every front entry below the diagonal), its rows, D
  F.cp.assign(n + 1, 0);
  S.dpos.resize(n);
  for (uint32_t b = 0; b < nb; ++b)
    for (uint32_t c = 0; c < S.bk[b]; ++c) {
      const uint32_t j = S.bstart[b] + c;
      F.cp[j + 1] = S.bf[b] - c - 1;
      S.dpos[j] = S.foff[b] + (uint64_t)c * S.bf[b] + c;
    }
  for (size_t j = 0; j < n; ++j) F.cp[j + 1] += F.cp[j];
  F.ci.resize(F.cp[n]);
  S.gmap_c.resize(F.cp[n]);
  parallel(nb, [&](size_t b) {
    const uint32_t f = S.bf[b];
    for (uint32_t c = 0; c < S.bk[b]; ++c) {
      uint32_t q = F.cp[S.bstart[b] + c];
      for (uint32_t r = c + 1; r < f; ++r, ++q) {
        F.ci[q] = front[b][r];
        S.gmap_c[q] = S.foff[b] + (uint64_t)r * f + c;
      }
    }
  });
  F.rp.assign(n + 1, 0);
  for (uint32_t i : F.ci) F.rp[i + 1]++;
  for (size_t i = 0; i < n; ++i) F.rp[i + 1] += F.rp[i];
  F.rj.resize(F.ci.size());
  S.gmap_r.resize(F.ci.size());
  {
    std::vector<uint32_t> fill(F.rp.begin(), F.rp.end() - 1);
    for (size_t j = 0; j < n; ++j)
      for (uint32_t q = F.cp[j]; q < F.cp[j + 1]; ++q) {
        const uint32_t i = F.ci[q];
        F.rj[fill[i]] = (uint32_t)j;
        S.gmap_r[fill[i]++] = S.gmap_c[q];
      }
  }
  S.plan = make_lprime_schedule(F, nd);
  // work lists per height
  const size_t lds_doubles = lds_bytes / sizeof(double);
  auto fits_lds = [&](uint32_t f) { return (uint64_t)f * (f + 1) / 2 + f <= lds_doubles; };
  const uint32_t H = S.H, MC = std::max<uint32_t>(S.max_children, 1);
  std::vector<std::vector<uint32_t>> lds_h(H), big_h(H);
  for (uint32_t b = 0; b < nb; ++b) {
    if (S.bf[b] == 0) continue;
    (fits_lds(S.bf[b]) ? lds_h : big_h)[S.bheight[b]].push_back(b);
  }
  S.lds_off.assign(1, 0);
  S.lds_rows.assign(H, 0);
  std::vector<std::vector<FrontItem>> ea(H * MC);
  S.step_off.assign(1, 0);
  for (uint32_t h = 0; h < H; ++h) {
    for (uint32_t b : lds_h[h]) {
      S.lds_blocks.push_back(b);
      S.lds_rows[h] = std::max(S.lds_rows[h], S.bf[b]);
    }
    S.lds_off.push_back((uint32_t)S.lds_blocks.size());
    uint32_t kmax = 0;
    for (uint32_t b : big_h[h]) {
      kmax = std::max(kmax, S.bk[b]);
      for (uint32_t s = S.child_ptr[b]; s < S.child_ptr[b + 1]; ++s) {
        const uint32_t c = S.child[s], rows = S.bf[c] - S.bk[c];
        for (uint32_t r0 = 0; r0 < rows; r0 += kEaRows)
          ea[h * MC + (s - S.child_ptr[b])].push_back({c, r0, std::min(rows, r0 + kEaRows)});
      }
    }
    for (uint32_t c0 = 0; c0 < kmax; c0 += kPanelCols) {
      S.step_c0.push_back(c0);
      S.panel_off.push_back((uint32_t)S.panels.size());
      S.tile_off.push_back((uint32_t)S.tiles.size());
      for (uint32_t b : big_h[h]) {
        const uint32_t k = S.bk[b], f = S.bf[b];
        if (k <= c0) continue;
        const uint32_t c1 = std::min(c0 + kPanelCols, k);
        for (uint32_t r0 = 0; r0 == 0 || c1 + r0 < f; r0 += kPanelItemRows) S.panels.push_back({b, r0});
        const uint32_t T = (f - c1 + kTile - 1) / kTile;
        for (uint32_t I = 0; I < T; ++I)
          for (uint32_t J = 0; J <= I; ++J) S.tiles.push_back({b, c1 + kTile * I, c1 + kTile * J});
      }
    }
    S.step_off.push_back((uint32_t)S.step_c0.size());
  }
  S.panel_off.push_back((uint32_t)S.panels.size());
  S.tile_off.push_back((uint32_t)S.tiles.size());
  S.max_children = MC;
  S.ea_item_off.assign(1, 0);
  for (auto& v : ea) {
    S.ea_items.insert(S.ea_items.end(), v.begin(), v.end());
    S.ea_item_off.push_back((uint32_t)S.ea_items.size());
  }
  return S;
}

// ---- numeric kernels ---------------------------------------------------------------------------

struct FrontArgs {
  const uint32_t *bstart, *bk, *bf, *bparent, *child_ptr, *child, *ea_off, *ea_map;
  const uint64_t* foff;
};

__device__ __forceinline__ size_t tri(uint32_t r) { return (size_t)r * (r + 1) / 2; }

__device__ __forceinline__ double wsum64(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// a_v = Σ_{S∋v} a_S / 3, faces ascending (RS/src/geometry.rs:194)
__global__ __launch_bounds__(256) void lf_mass_kernel(uint32_t nv, const uint32_t* __restrict__ inc_ptr,
                                                      const uint32_t* __restrict__ inc, const double* __restrict__ area,
                                                      double* __restrict__ mass) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  double a = 0.0;
  for (uint32_t q = inc_ptr[v]; q < inc_ptr[v + 1]; ++q) a += area[inc[q] >> 2] / 3.0;
  mass[v] = a;
}

#pragma clang fp contract(off)
// Row a of W (RS/src/geometry.rs:221): for each incident face (ascending) and corner c whose
// opposite edge contains a, −½ cot(angle at c) on the edge's other end and +½ cot on the diagonal,
// in the order of the host assembly.
__global__ __launch_bounds__(256) void lf_cotan_kernel(uint32_t nv, const uint32_t* __restrict__ kp,
                                                       const uint32_t* __restrict__ kj,
                                                       const uint32_t* __restrict__ inc_ptr,
                                                       const uint32_t* __restrict__ inc,
                                                       const uint32_t* __restrict__ faces,
                                                       const double* __restrict__ pos, double* __restrict__ kx) {
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a >= nv) return;
  const uint32_t r0 = kp[a], r1 = kp[a + 1];
  for (uint32_t q = r0; q < r1; ++q) kx[q] = 0.0;
  auto at = [&](uint32_t col) -> uint32_t {
    uint32_t lo = r0, hi = r1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (kj[mid] < col)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  };
  const uint32_t qa = at(a);
  for (uint32_t e = inc_ptr[a]; e < inc_ptr[a + 1]; ++e) {
    const uint32_t* t = faces + 3 * (size_t)(inc[e] >> 2);
    for (int c = 0; c < 3; ++c) {
      const uint32_t x = t[(c + 1) % 3], y = t[(c + 2) % 3];
      if (x != a && y != a) continue;
      const double* p0 = pos + 3 * (size_t)t[c];
      const double* p1 = pos + 3 * (size_t)x;
      const double* p2 = pos + 3 * (size_t)y;
      const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      const double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      const double cot = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      const double h = 0.5 * cot;
      const uint32_t qo = at(x == a ? y : x);
      kx[qo] += -h;
      kx[qa] += h;
    }
  }
}

__global__ __launch_bounds__(256) void lf_diag_kernel(uint32_t nv, const uint32_t* __restrict__ kdiag,
                                                      const double* __restrict__ kx, double* __restrict__ out) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v < nv) out[v] = kx[kdiag[v]];
}

// ε = ε_rel tr W / tr M (sums[0], sums[1]) into sums[2]; K = W + εM; c = a / Σa.
__global__ __launch_bounds__(256) void lf_finish_k_kernel(uint32_t nv, double eps_rel, const uint32_t* __restrict__ kdiag,
                                                          const double* __restrict__ mass, double* __restrict__ sums,
                                                          double* __restrict__ kx, double* __restrict__ crow) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  const double trw = sums[0], trm = sums[1];
  const double eps = eps_rel * trw / trm;
  if (v == 0) sums[2] = eps;
  if (v >= nv) return;
  kx[kdiag[v]] += eps * mass[v];
  crow[v] = mass[v] / trm;
}
#pragma clang fp contract(on)

__global__ __launch_bounds__(256) void lf_fill_kernel(size_t n, double* __restrict__ x) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) x[i] = 0.0;
}

__global__ void lf_reset_kernel(uint32_t* bad) { *bad = UINT32_MAX; }

// L′'s lower triangle into the fronts: K's entries, then the barycenter row.
__global__ __launch_bounds__(256) void lf_assemble_kernel(uint32_t nnz, uint32_t nv, const uint64_t* __restrict__ amap_k,
                                                          const double* __restrict__ kx,
                                                          const uint64_t* __restrict__ amap_c,
                                                          const double* __restrict__ crow, double* __restrict__ front) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q < nnz && amap_k[q] != LprimeSymbolic::kNoPos) front[amap_k[q]] = kx[q];
  if (q < nv) front[amap_c[q]] = crow[q];
}

// One workgroup per front that fits LDS: load its lower triangle (packed by rows), add the
// children's update matrices (one child after another), eliminate its k columns, write back.
__global__ __launch_bounds__(256) void lf_lds_front_kernel(FrontArgs A, const uint32_t* __restrict__ list,
                                                           double* __restrict__ front, uint32_t* __restrict__ bad) {
  extern __shared__ double T[];
  const uint32_t b = list[blockIdx.x];
  const uint32_t f = A.bf[b], k = A.bk[b], j0 = A.bstart[b];
  double* G = front + A.foff[b];
  double* wv = T + tri(f);
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t r = wave; r < f; r += 4) {
    const size_t rs = tri(r);
    for (uint32_t c = lane; c <= r; c += 64) T[rs + c] = G[(size_t)r * f + c];
  }
  __syncthreads();
  for (uint32_t s = A.child_ptr[b]; s < A.child_ptr[b + 1]; ++s) {
    const uint32_t ch = A.child[s], cf = A.bf[ch], ck = A.bk[ch], cs = cf - ck;
    const uint32_t* map = A.ea_map + A.ea_off[ch];
    const double* Gc = front + A.foff[ch];
    for (uint32_t r = wave; r < cs; r += 4) {
      const size_t rs = tri(map[r]);
      const double* src = Gc + (size_t)(ck + r) * cf + ck;
      for (uint32_t c = lane; c <= r; c += 64) T[rs + map[c]] += src[c];
    }
    __syncthreads();
  }
  for (uint32_t c = 0; c < k; ++c) {
    const double d = T[tri(c) + c];
    if (threadIdx.x == 0 && !(d != 0.0 && isfinite(d))) atomicMin(bad, j0 + c);
    for (uint32_t r = c + 1 + threadIdx.x; r < f; r += 256) {
      const size_t rs = tri(r);
      const double w = T[rs + c];
      wv[r] = w;
      T[rs + c] = w / d;
    }
    __syncthreads();
    for (uint32_t r = c + 1 + wave; r < f; r += 4) {
      const size_t rs = tri(r);
      const double l = T[rs + c];
      for (uint32_t m = c + 1 + lane; m <= r; m += 64) T[rs + m] = fma(-l, wv[m], T[rs + m]);
    }
    __syncthreads();
  }
  for (uint32_t r = wave; r < f; r += 4) {
    const size_t rs = tri(r);
    for (uint32_t c = lane; c <= r; c += 64) G[(size_t)r * f + c] = T[rs + c];
  }
}

// One workgroup per item: rows [r0, r1) of a child's update matrix added into its parent's front.
__global__ __launch_bounds__(256) void lf_extend_add_kernel(FrontArgs A, const FrontItem* __restrict__ items,
                                                            double* __restrict__ front) {
  const FrontItem it = items[blockIdx.x];
  const uint32_t ch = it.child, p = A.bparent[ch];
  const uint32_t cf = A.bf[ch], ck = A.bk[ch], fp = A.bf[p];
  const uint32_t* map = A.ea_map + A.ea_off[ch];
  const double* Gc = front + A.foff[ch];
  double* Gp = front + A.foff[p];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t r = it.r0 + wave; r < it.r1; r += 4) {
    const double* src = Gc + (size_t)(ck + r) * cf + ck;
    double* dst = Gp + (size_t)map[r] * fp;
    for (uint32_t c = lane; c <= r; c += 64) dst[map[c]] += src[c];
  }
}

// Diagonal block of panel [c0, c1) of a large front, factored in place: D = L_D diag(d) L_Dᵀ
// (one wave, one row per lane; items with r0 ≠ 0 have nothing to do).
__global__ __launch_bounds__(64) void lf_panel_diag_kernel(FrontArgs A, const FrontPanel* __restrict__ items,
                                                           uint32_t c0, double* __restrict__ front,
                                                           uint32_t* __restrict__ bad) {
  __shared__ double Ds[kPanelCols][kPanelCols + 1];
  __shared__ double wv[kPanelCols];
  const FrontPanel it = items[blockIdx.x];
  if (it.r0 != 0) return;
  const uint32_t b = it.block, f = A.bf[b], k = A.bk[b], j0 = A.bstart[b];
  const uint32_t c1 = min(c0 + kPanelCols, k), nbk = c1 - c0;
  double* G = front + A.foff[b];
  const uint32_t lane = threadIdx.x;
  if (lane < nbk)
    for (uint32_t j = 0; j <= lane; ++j) Ds[lane][j] = G[(size_t)(c0 + lane) * f + c0 + j];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t t = 0; t < nbk; ++t) {
    const double d = Ds[t][t];
    if (lane > t && lane < nbk) {
      const double w = Ds[lane][t];
      wv[lane] = w;
      Ds[lane][t] = w / d;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane > t && lane < nbk) {
      const double l = Ds[lane][t];
      for (uint32_t m = t + 1; m <= lane; ++m) Ds[lane][m] = fma(-l, wv[m], Ds[lane][m]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (lane < nbk) {
    for (uint32_t j = 0; j <= lane; ++j) G[(size_t)(c0 + lane) * f + c0 + j] = Ds[lane][j];
    const double d = Ds[lane][lane];
    if (!(d != 0.0 && isfinite(d))) atomicMin(bad, j0 + c0 + lane);
  }
}

// Rows below the diagonal block of panel [c0, c1) (the block factored): one thread per row r,
// w_t = a_rt − Σ_{s<t} w_s (L_D)_ts, l_rt = w_t / d_t; the rows are independent.
__global__ __launch_bounds__(kPanelItemRows) void lf_panel_kernel(FrontArgs A, const FrontPanel* __restrict__ items,
                                                                  uint32_t c0, double* __restrict__ front) {
  __shared__ double Ds[kPanelCols][kPanelCols + 1];
  __shared__ double R[kPanelItemRows][kPanelCols + 1];
  const FrontPanel it = items[blockIdx.x];
  const uint32_t b = it.block, f = A.bf[b], k = A.bk[b];
  const uint32_t c1 = min(c0 + kPanelCols, k), nbk = c1 - c0;
  const uint32_t rbase = c1 + it.r0;
  if (rbase >= f) return;  // uniform per workgroup
  const uint32_t nrows = min(kPanelItemRows, f - rbase);
  double* G = front + A.foff[b];
  const uint32_t tid = threadIdx.x;
  for (uint32_t e = tid; e < kPanelCols * kPanelCols; e += kPanelItemRows) {
    const uint32_t i = e / kPanelCols, j = e % kPanelCols;
    Ds[i][j] = (i < nbk && j <= i) ? G[(size_t)(c0 + i) * f + c0 + j] : (i == j ? 1.0 : 0.0);
  }
  for (uint32_t e = tid; e < nrows * kPanelCols; e += kPanelItemRows) {
    const uint32_t i = e / kPanelCols, t = e % kPanelCols;
    R[i][t] = t < nbk ? G[(size_t)(rbase + i) * f + c0 + t] : 0.0;
  }
  __syncthreads();
  if (tid < nrows) {
    double w[kPanelCols];
#pragma unroll
    for (uint32_t t = 0; t < kPanelCols; ++t) {
      double x = R[tid][t];
#pragma unroll
      for (uint32_t s2 = 0; s2 < t; ++s2) x = fma(-w[s2], Ds[t][s2], x);
      w[t] = x;
    }
#pragma unroll
    for (uint32_t t = 0; t < kPanelCols; ++t) R[tid][t] = w[t] / Ds[t][t];
  }
  __syncthreads();
  for (uint32_t e = tid; e < nrows * kPanelCols; e += kPanelItemRows) {
    const uint32_t i = e / kPanelCols, t = e % kPanelCols;
    if (t < nbk) G[(size_t)(rbase + i) * f + c0 + t] = R[i][t];
  }
}

// One workgroup per 64 × 64 tile of a large front's trailing part:
// G[i][m] −= Σ_{t ∈ [c0, c1)} L[i][t] d_t L[m][t]  for m ≤ i.
__global__ __launch_bounds__(256) void lf_update_kernel(FrontArgs A, const FrontTile* __restrict__ tiles, uint32_t c0,
                                                        double* __restrict__ front) {
  __shared__ double Ls[kTile][kPanelCols + 1], Ws[kTile][kPanelCols + 1];
  const FrontTile t = tiles[blockIdx.x];
  const uint32_t f = A.bf[t.block], k = A.bk[t.block];
  const uint32_t c1 = min(c0 + kPanelCols, k), nbk = c1 - c0;
  double* G = front + A.foff[t.block];
  for (uint32_t e = threadIdx.x; e < kTile * kPanelCols; e += 256) {
    const uint32_t rr = e / kPanelCols, tt = e % kPanelCols;
    const uint32_t i = t.i0 + rr, m = t.j0 + rr;
    const bool ok = tt < nbk;
    Ls[rr][tt] = (ok && i < f) ? G[(size_t)i * f + c0 + tt] : 0.0;
    Ws[rr][tt] = (ok && m < f) ? G[(size_t)m * f + c0 + tt] * G[(size_t)(c0 + tt) * f + c0 + tt] : 0.0;
  }
  __syncthreads();
  const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (uint32_t tt = 0; tt < nbk; ++tt) {
    double a[4], w[4];
    for (int p = 0; p < 4; ++p) {
      a[p] = Ls[ty * 4 + p][tt];
      w[p] = Ws[tx * 4 + p][tt];
    }
    for (int p = 0; p < 4; ++p)
      for (int q = 0; q < 4; ++q) acc[p][q] = fma(a[p], w[q], acc[p][q]);
  }
  for (int p = 0; p < 4; ++p) {
    const uint32_t i = t.i0 + ty * 4 + p;
    if (i >= f) break;
    for (int q = 0; q < 4; ++q) {
      const uint32_t m = t.j0 + tx * 4 + q;
      if (m <= i) G[(size_t)i * f + m] -= acc[p][q];
    }
  }
}

__global__ __launch_bounds__(256) void lf_gather_kernel(size_t nnz, uint32_t n, const uint64_t* __restrict__ gmap_r,
                                                        const uint64_t* __restrict__ gmap_c,
                                                        const uint64_t* __restrict__ dpos,
                                                        const double* __restrict__ front, double* __restrict__ rx,
                                                        double* __restrict__ cx, double* __restrict__ d) {
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nnz || q < n; q += (size_t)gridDim.x * 256) {
    if (q < nnz) {
      rx[q] = front[gmap_r[q]];
      cx[q] = front[gmap_c[q]];
    }
    if (q < n) d[q] = front[dpos[q]];
  }
}

// One wave per row j of the order: column c = j − start(J) of M_J = L_JJ⁻¹ (unit lower) by forward
// substitution, m_i = −Σ_{c ≤ l < i} L_il m_l; written to M_J (row-major) and M_Jᵀ.
__global__ __launch_bounds__(256) void lf_inverse_kernel(uint32_t n, uint32_t max_rows, FrontArgs A,
                                                         const uint32_t* __restrict__ block_of,
                                                         const uint64_t* __restrict__ inv_off,
                                                         const double* __restrict__ front, double* __restrict__ inv,
                                                         double* __restrict__ inv_t) {
  extern __shared__ double lds[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t j = blockIdx.x * 4 + wave;
  if (j >= n) return;
  const uint32_t J = block_of[j], j0 = A.bstart[J], k = A.bk[J], f = A.bf[J], c = j - j0;
  const double* L = front + A.foff[J];
  double* m = lds + (size_t)wave * max_rows;
  double* M = inv + inv_off[J];
  double* Mt = inv_t + inv_off[J];
  if (lane == 0) {
    m[c] = 1.0;
    M[(size_t)c * k + c] = 1.0;
    Mt[(size_t)c * k + c] = 1.0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = c + 1; i < k; ++i) {
    const double* Li = L + (size_t)i * f;
    double s = 0.0;
    for (uint32_t l = c + lane; l < i; l += 64) s = fma(Li[l], m[l], s);
    s = wsum64(s);
    if (lane == 0) {
      m[i] = -s;
      M[(size_t)i * k + c] = -s;
      Mt[(size_t)c * k + i] = -s;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// ---- host side -------------------------------------------------------------------------------------

LprimeFactorDev::~LprimeFactorDev() {
  if (bad_host) (void)hipHostFree(bad_host);
}

void LprimeFactorDev::setup(ddgrs_ctx* ctx, const LprimeSymbolic& S, LprimeDev& lp) {
  nv = S.nv;
  n = S.n;
  H = S.H;
  max_children = S.max_children;
  max_rows = S.max_rows;
  front_total = S.front_total;
  lds_off = S.lds_off;
  lds_rows = S.lds_rows;
  ea_item_off = S.ea_item_off;
  step_off = S.step_off;
  step_c0 = S.step_c0;
  panel_off = S.panel_off;
  tile_off = S.tile_off;
  nnz_l = S.pattern.ci.size();
  nnz_k = S.kpat.idx.size();
  ddgrs::upload(ctx, kp, S.kpat.ptr);
  ddgrs::upload(ctx, kj, S.kpat.idx);
  ddgrs::upload(ctx, kdiag, S.kdiag);
  ddgrs::upload(ctx, bstart, S.bstart);
  ddgrs::upload(ctx, bk, S.bk);
  ddgrs::upload(ctx, bf, S.bf);
  ddgrs::upload(ctx, bparent, S.bparent);
  ddgrs::upload(ctx, child_ptr, S.child_ptr);
  ddgrs::upload(ctx, child, S.child);
  ddgrs::upload(ctx, ea_off, S.ea_off);
  ddgrs::upload(ctx, ea_map, S.ea_map);
  ddgrs::upload(ctx, lds_blocks, S.lds_blocks);
  ddgrs::upload(ctx, panels, S.panels);
  ddgrs::upload(ctx, foff, S.foff);
  ddgrs::upload(ctx, amap_k, S.amap_k);
  ddgrs::upload(ctx, amap_c, S.amap_c);
  ddgrs::upload(ctx, gmap_r, S.gmap_r);
  ddgrs::upload(ctx, gmap_c, S.gmap_c);
  ddgrs::upload(ctx, dpos, S.dpos);
  ddgrs::upload(ctx, ea_items, S.ea_items);
  ddgrs::upload(ctx, tiles, S.tiles);
  front.reserve(std::max<uint64_t>(front_total, 1));
  kx.reserve(std::max<size_t>(S.kpat.idx.size(), 1));
  mass.reserve(nv);
  crow.reserve(nv);
  diagv.reserve(nv);
  sums.reserve(4);
  bad.reserve(1);
  if (!bad_host) HIPCHECK(hipHostMalloc((void**)&bad_host, sizeof(uint32_t), hipHostMallocDefault));
  *bad_host = UINT32_MAX;
  size_t lds_max = 0;
  for (uint32_t r : lds_rows) lds_max = std::max(lds_max, ((size_t)r * (r + 1) / 2 + r) * sizeof(double));
  if (lds_max > 0)
    HIPCHECK(hipFuncSetAttribute((const void*)lf_lds_front_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds_max));
  lp.nv = nv;
  lp.setup(ctx, S.pattern, S.plan);
  ctx->sync();  // the symbolic structure may be released after return
}

void LprimeFactorDev::factor(ddgrs_surface* s, double eps_rel, LprimeDev& lp) {
  ddgrs_ctx* ctx = s->ctx;
  const uint32_t unv = (uint32_t)nv;
  const unsigned gv = blocks_for((int64_t)nv, 256);
  // K = W + εM and the barycenter row
  launch(ctx, "lf_mass", lf_mass_kernel, dim3(gv), dim3(256), 0, unv, (const uint32_t*)s->inc_ptr.ptr,
         (const uint32_t*)s->inc.ptr, (const double*)s->area.ptr, mass.ptr);
  launch(ctx, "lf_cotan", lf_cotan_kernel, dim3(gv), dim3(256), 0, unv, (const uint32_t*)kp.ptr, (const uint32_t*)kj.ptr,
         (const uint32_t*)s->inc_ptr.ptr, (const uint32_t*)s->inc.ptr, (const uint32_t*)s->faces.ptr,
         (const double*)s->positions.ptr, kx.ptr);
  launch(ctx, "lf_diag", lf_diag_kernel, dim3(gv), dim3(256), 0, unv, (const uint32_t*)kdiag.ptr, (const double*)kx.ptr,
         diagv.ptr);
  sum_f64(ctx, (int64_t)nv, diagv.ptr, sums.ptr);
  sum_f64(ctx, (int64_t)nv, mass.ptr, sums.ptr + 1);
  launch(ctx, "lf_finish_k", lf_finish_k_kernel, dim3(gv), dim3(256), 0, unv, eps_rel, (const uint32_t*)kdiag.ptr,
         (const double*)mass.ptr, sums.ptr, kx.ptr, crow.ptr);
  // fronts
  launch(ctx, "lf_fill", lf_fill_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for((int64_t)front_total, 256), 4096)),
         dim3(256), 0, (size_t)front_total, front.ptr);
  launch(ctx, "lf_reset", lf_reset_kernel, dim3(1), dim3(1), 0, bad.ptr);
  launch(ctx, "lf_assemble", lf_assemble_kernel, dim3(blocks_for(std::max<size_t>(nnz_k, nv), 256)), dim3(256), 0,
         (uint32_t)nnz_k, unv, (const uint64_t*)amap_k.ptr, (const double*)kx.ptr, (const uint64_t*)amap_c.ptr,
         (const double*)crow.ptr, front.ptr);
  const FrontArgs A{bstart.ptr, bk.ptr, bf.ptr, bparent.ptr, child_ptr.ptr, child.ptr, ea_off.ptr, ea_map.ptr, foff.ptr};
  for (uint32_t h = 0; h < H; ++h) {
    const uint32_t nl = lds_off[h + 1] - lds_off[h];
    if (nl) {
      const size_t r = lds_rows[h];
      launch(ctx, "lf_lds_front", lf_lds_front_kernel, dim3(nl), dim3(256), (r * (r + 1) / 2 + r) * sizeof(double), A,
             (const uint32_t*)lds_blocks.ptr + lds_off[h], front.ptr, bad.ptr);
    }
    for (uint32_t slot = 0; slot < max_children; ++slot) {
      const uint32_t q = h * max_children + slot, cnt = ea_item_off[q + 1] - ea_item_off[q];
      launch(ctx, "lf_extend_add", lf_extend_add_kernel, dim3(cnt), dim3(256), 0, A,
             (const FrontItem*)ea_items.ptr + ea_item_off[q], front.ptr);
    }
    for (uint32_t st = step_off[h]; st < step_off[h + 1]; ++st) {
      const uint32_t np = panel_off[st + 1] - panel_off[st];
      const FrontPanel* items = (const FrontPanel*)panels.ptr + panel_off[st];
      launch(ctx, "lf_panel_diag", lf_panel_diag_kernel, dim3(np), dim3(64), 0, A, items, step_c0[st], front.ptr, bad.ptr);
      launch(ctx, "lf_panel", lf_panel_kernel, dim3(np), dim3(kPanelItemRows), 0, A, items, step_c0[st], front.ptr);
      launch(ctx, "lf_update", lf_update_kernel, dim3(tile_off[st + 1] - tile_off[st]), dim3(256), 0, A,
             (const FrontTile*)tiles.ptr + tile_off[st], step_c0[st], front.ptr);
    }
  }
  // L's rows and columns, D, the block inverses
  launch(ctx, "lf_gather", lf_gather_kernel,
         dim3((unsigned)std::min<uint64_t>(blocks_for((int64_t)std::max<size_t>(nnz_l, n), 256), 8192)), dim3(256), 0,
         nnz_l, (uint32_t)n, (const uint64_t*)gmap_r.ptr, (const uint64_t*)gmap_c.ptr, (const uint64_t*)dpos.ptr,
         (const double*)front.ptr, lp.rx.ptr, lp.cx.ptr, lp.d.ptr);
  launch(ctx, "lf_inverse", lf_inverse_kernel, dim3(blocks_for((int64_t)n, 4)), dim3(256),
         4 * (size_t)max_rows * sizeof(double), (uint32_t)n, max_rows, A, (const uint32_t*)lp.block_of.ptr,
         (const uint64_t*)lp.blk_inv.ptr, (const double*)front.ptr, lp.inv.ptr, lp.inv_t.ptr);
  download(ctx, bad_host, (const uint32_t*)bad.ptr, 1);
}

void LprimeFactorDev::check(ddgrs_ctx* ctx) {
  const uint32_t b = *bad_host;
  if (b != UINT32_MAX) {
    Error e(DDGRS_SINGULAR_SADDLE, "preconditioner factorization", "zero or non-finite pivot " + std::to_string(b));
    e.a = b;
    throw e;
  }
  (void)ctx;
}

double LprimeFactorDev::eps(ddgrs_ctx* ctx) {
  double* pin = (double*)ctx->staging(8);
  download(ctx, pin, (const double*)sums.ptr + 2, 1);
  ctx->sync();
  return pin[0];
}

}  // namespace ddgrs
