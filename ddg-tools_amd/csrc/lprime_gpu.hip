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
#include <cstdlib>
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
  S.perm = nd.order;
  S.perm.push_back((uint32_t)nv);
  const std::vector<uint32_t>& perm = S.perm;
  std::vector<uint32_t> pinv(n);
  for (size_t i = 0; i < n; ++i) pinv[perm[i]] = (uint32_t)i;
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
      const uint32_t o = perm[j];
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
  // solves: blocks with a front by height, later front rows, inverse offsets
  S.block_of.resize(n);
  for (uint32_t b = 0; b < nb; ++b)
    for (uint32_t i = S.bstart[b]; i < bend[b]; ++i) S.block_of[i] = b;
  S.srows.resize(S.ea_off[nb]);
  S.moff.assign(nb + 1, 0);
  for (uint32_t b = 0; b < nb; ++b) {
    std::copy(front[b].begin() + S.bk[b], front[b].end(), S.srows.begin() + S.ea_off[b]);
    S.moff[b + 1] = S.moff[b] + (uint64_t)S.bk[b] * S.bk[b];
  }
  S.hoff.assign(1, 0);
  S.hsplit.assign(S.H, 0);
  S.hlds.assign(S.H, 0);
  S.krows_off.assign(1, 0);
  S.urows_off.assign(1, 0);
  S.ltoff.assign(nb + 1, 0);
  // DDGRS_LPRIME_LARGE_PATH=1 solves every height on the large-front path (to test it on small meshes)
  const char* force = std::getenv("DDGRS_LPRIME_LARGE_PATH");
  const size_t small_lds = (force && force[0] == '1') ? 0 : std::min(lds_bytes, kSolveSmallLds);
  for (uint32_t h = 0; h < S.H; ++h) {
    const uint32_t h0 = (uint32_t)S.hblocks.size();
    size_t need = 0;
    for (uint32_t b = 0; b < nb; ++b)
      if (S.bheight[b] == h && S.bf[b] > 0) {
        S.hblocks.push_back(b);
        const size_t f = S.bf[b], k = S.bk[b];
        need = std::max(need, ((k + 1) * f + 3 * f + 3 * k) * sizeof(double));
      }
    S.hoff.push_back((uint32_t)S.hblocks.size());
    S.hsplit[h] = need > small_lds;
    S.hlds[h] = (uint32_t)need;
    if (S.hsplit[h])
      for (uint32_t q = h0; q < S.hblocks.size(); ++q) {
        const uint32_t b = S.hblocks[q], k = S.bk[b], ns = S.bf[b] - k;
        for (uint32_t r0 = 0; r0 < k; r0 += kSolveRows) S.krows.push_back({b, r0});
        for (uint32_t r0 = 0; r0 < ns; r0 += kSolveRows) S.urows.push_back({b, r0});
        for (uint32_t r0 = 0; r0 < ns; r0 += 32)
          for (uint32_t i0 = 0; i0 < k; i0 += 32) S.lt_tiles.push_back({b, r0, i0});
      }
    S.krows_off.push_back((uint32_t)S.krows.size());
    S.urows_off.push_back((uint32_t)S.urows.size());
  }
  for (uint32_t b = 0; b < nb; ++b)
    S.ltoff[b + 1] = S.ltoff[b] + (S.hsplit[S.bheight[b]] ? (uint64_t)S.bk[b] * (S.bf[b] - S.bk[b]) : 0);
  // subtrees of the lowest small-front heights whose fronts fit a wave's LDS slice, at most
  // kSubLevels of them (DDGRS_LPRIME_SUBTREES=<levels> sets that bound; 0: every height by its own
  // launches)
  const char* sub_env = std::getenv("DDGRS_LPRIME_SUBTREES");
  const uint32_t max_levels = sub_env ? (uint32_t)std::strtoul(sub_env, nullptr, 10) : kSubLevels;
  for (uint32_t h = 0; h < std::min(S.H, max_levels) && !S.hsplit[h]; ++h) {
    uint32_t need = 0;
    for (uint32_t q = S.hoff[h]; q < S.hoff[h + 1]; ++q)
      need = std::max(need, 3 * (S.bf[S.hblocks[q]] + S.bk[S.hblocks[q]]));
    if (need > kSubWaveDoubles) break;
    S.sub_levels = h + 1;
    S.sub_wave_doubles = std::max(S.sub_wave_doubles, need);
  }
  if (S.sub_levels > 0) {
    const uint32_t L = S.sub_levels;
    std::vector<std::vector<uint32_t>> lv(L);
    std::vector<uint32_t> stack;
    S.sub_lvl_off.assign(1, 0);
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t p = S.bparent[b];
      if (S.bheight[b] >= L || (p != kNone && S.bheight[p] < L)) continue;
      for (auto& v : lv) v.clear();
      stack.assign(1, b);
      while (!stack.empty()) {
        const uint32_t c = stack.back();
        stack.pop_back();
        if (S.bf[c] > 0) lv[S.bheight[c]].push_back(c);
        for (uint32_t q = S.child_ptr[c]; q < S.child_ptr[c + 1]; ++q) stack.push_back(S.child[q]);
      }
      for (auto& v : lv) {
        std::sort(v.begin(), v.end());
        S.sub_fronts.insert(S.sub_fronts.end(), v.begin(), v.end());
        S.sub_lvl_off.push_back((uint32_t)S.sub_fronts.size());
      }
    }
  }
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

__global__ void lf_reset_kernel(uint32_t* bad, uint32_t* solve_failed) {
  *bad = UINT32_MAX;
  *solve_failed = 0u;
}

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

// One wave per row j of the order: column c = j − start(J) of M_J = L_JJ⁻¹ (unit lower) by forward
// substitution, m_i = −Σ_{c ≤ l < i} L_il m_l (L_JJ: the top-left k × k of J's front); M_J and
// M_Jᵀ, both row-major.
__global__ __launch_bounds__(256) void lf_inverse_kernel(uint32_t n, uint32_t max_rows, FrontArgs A,
                                                         const uint32_t* __restrict__ block_of,
                                                         const uint64_t* __restrict__ moff,
                                                         const double* __restrict__ front, double* __restrict__ minv,
                                                         double* __restrict__ minv_t) {
  extern __shared__ double lds[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t j = blockIdx.x * 4 + wave;
  if (j >= n) return;
  const uint32_t J = block_of[j], j0 = A.bstart[J], k = A.bk[J], f = A.bf[J], c = j - j0;
  const double* L = front + A.foff[J];
  double* m = lds + (size_t)wave * max_rows;
  double* M = minv + moff[J];
  double* Mt = minv_t + moff[J];
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

// ---- solves with the fronts -------------------------------------------------------------------
//
// The solve's items hand vectors to one another: update vectors u (forward, child → parent), y_J
// (forward → backward of the same block), t (backward, between the two phases of a large block)
// and x (backward, parent → children, which read their later rows).  In the one-launch solve
// (Coherent) these loads and stores are device-scope atomics, coherent across the XCDs' L2s
// without fences; the per-height launches use plain ones (the launch boundaries order them).
template <bool Coherent>
__device__ __forceinline__ double ld_v(const double* p) {
  if constexpr (Coherent)
    return collect(const_cast<double*>(p));
  else
    return *p;
}
template <bool Coherent>
__device__ __forceinline__ void st_v(double* p, double v) {
  if constexpr (Coherent)
    publish(p, v);
  else
    *p = v;
}

// b (order numbering, 3 per row) from a saddle vector (positions interleaved, multipliers last);
// the first `nzero` threads also zero the one-launch solve's counters.
__global__ __launch_bounds__(256) void lf_gather_rhs_kernel(uint32_t n, uint32_t nv, const uint32_t* __restrict__ perm,
                                                            const double* __restrict__ r, double* __restrict__ b,
                                                            uint32_t nzero, uint32_t* __restrict__ zero) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t < nzero) zero[t] = 0u;
  if (t >= 3 * n) return;
  const uint32_t j = t / 3, a = t % 3, o = perm[j];
  b[t] = o < nv ? r[3 * (size_t)o + a] : r[3 * (size_t)nv + a];
}

// x from xb; NaN everywhere when the one-launch solve gave up waiting (`failed` ≠ 0), so that the
// caller sees a non-finite result instead of a wrong one.
__global__ __launch_bounds__(256) void lf_scatter_sol_kernel(uint32_t n, uint32_t nv, const uint32_t* __restrict__ perm,
                                                             const double* __restrict__ xb,
                                                             const uint32_t* __restrict__ failed,
                                                             double* __restrict__ x) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * n) return;
  const uint32_t j = t / 3, a = t % 3, o = perm[j];
  x[o < nv ? 3 * (size_t)o + a : 3 * (size_t)nv + a] = *failed ? __builtin_nan("") : xb[t];
}

// w = [b_J; 0] (3 per row, f rows) plus the children's update vectors, one child after another
// (the whole workgroup; ends synchronized).
template <bool Coherent>
__device__ void assemble_front_vector(const FrontArgs& A, uint32_t J, const double* __restrict__ b,
                                      const double* __restrict__ ubuf, double* w) {
  const uint32_t f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], tid = threadIdx.x;
  for (uint32_t i = tid; i < f; i += 256)
    for (int a = 0; a < 3; ++a) w[3 * i + a] = i < k ? b[3 * (size_t)(j0 + i) + a] : 0.0;
  __syncthreads();
  for (uint32_t q = A.child_ptr[J]; q < A.child_ptr[J + 1]; ++q) {
    const uint32_t ch = A.child[q], cs = A.bf[ch] - A.bk[ch];
    const uint32_t* map = A.ea_map + A.ea_off[ch];
    const double* uc = ubuf + 3 * (size_t)A.ea_off[ch];
    for (uint32_t r = tid; r < cs; r += 256)
      for (int a = 0; a < 3; ++a) w[3 * map[r] + a] += ld_v<Coherent>(uc + 3 * r + a);
    __syncthreads();
  }
}

// ---- small fronts: one workgroup per block, its matrices loaded into LDS first ------------------

// M_J (k × k, contiguous) and L_SJ (rows k.. of the f × f front G, its first k columns) into LDS,
// both with row stride k + 1 (L_SJ after M_J).  Each thread issues eight loads before it stores any:
// a load-then-store loop would wait a full memory round trip per element.
__device__ __forceinline__ void stage_block(const double* __restrict__ M, const double* __restrict__ G, uint32_t k,
                                            uint32_t f, double* __restrict__ sh) {
  constexpr int kBatch = 8;
  const uint32_t nm = k * k, n = nm + (f - k) * k, ks = k + 1;
  for (uint32_t e0 = threadIdx.x; e0 < n; e0 += kBatch * 256) {
    double v[kBatch];
    uint32_t dst[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const uint32_t e = min(e0 + 256u * j, n - 1);  // in bounds; the duplicates are not stored
      const bool in_m = e < nm;
      const uint32_t q = in_m ? e : e - nm, r = q / k, l = q - r * k;
      dst[j] = (in_m ? 0u : k * ks) + r * ks + l;
      v[j] = in_m ? M[e] : G[(size_t)(k + r) * f + l];
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j)
      if (e0 + 256u * j < n) sh[dst[j]] = v[j];
  }
}

// Forward: w = [b_J; 0] + children's update vectors, y_J = M_J w_J, u_J = w_S − L_SJ y_J.
// LDS: M_J (k × (k+1)), L_SJ ((f−k) × (k+1)), w (3 f), y_J (3 k).
template <bool Coherent>
__device__ void forward_small(const FrontArgs& A, uint32_t J, const uint64_t* __restrict__ moff,
                              const double* __restrict__ minv, const double* __restrict__ front,
                              const double* __restrict__ b, double* __restrict__ y, double* __restrict__ ubuf,
                              double* sh) {
  const uint32_t f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], ns = f - k, ks = k + 1, tid = threadIdx.x;
  double* Ms = sh;
  double* Ls = Ms + (size_t)k * ks;
  double* w = Ls + (size_t)ns * ks;
  double* yj = w + 3 * (size_t)f;
  const double* M = minv + moff[J];
  const double* G = front + A.foff[J];
  stage_block(M, G, k, f, sh);
  assemble_front_vector<Coherent>(A, J, b, ubuf, w);
  for (uint32_t i = tid; i < k; i += 256) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l = 0; l <= i; ++l) {
      const double m = Ms[i * ks + l];
      s0 = fma(m, w[3 * l], s0);
      s1 = fma(m, w[3 * l + 1], s1);
      s2 = fma(m, w[3 * l + 2], s2);
    }
    yj[3 * i] = s0;
    yj[3 * i + 1] = s1;
    yj[3 * i + 2] = s2;
    st_v<Coherent>(y + 3 * (size_t)(j0 + i), s0);
    st_v<Coherent>(y + 3 * (size_t)(j0 + i) + 1, s1);
    st_v<Coherent>(y + 3 * (size_t)(j0 + i) + 2, s2);
  }
  __syncthreads();
  double* u = ubuf + 3 * (size_t)A.ea_off[J];
  for (uint32_t r = tid; r < ns; r += 256) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l = 0; l < k; ++l) {
      const double v = Ls[r * ks + l];
      s0 = fma(v, yj[3 * l], s0);
      s1 = fma(v, yj[3 * l + 1], s1);
      s2 = fma(v, yj[3 * l + 2], s2);
    }
    st_v<Coherent>(u + 3 * r, w[3 * (k + r)] - s0);
    st_v<Coherent>(u + 3 * r + 1, w[3 * (k + r) + 1] - s1);
    st_v<Coherent>(u + 3 * r + 2, w[3 * (k + r) + 2] - s2);
  }
}

__global__ __launch_bounds__(256) void lf_forward_small_kernel(FrontArgs A, const uint32_t* __restrict__ list,
                                                               const uint64_t* __restrict__ moff,
                                                               const double* __restrict__ minv,
                                                               const double* __restrict__ front,
                                                               const double* __restrict__ b, double* __restrict__ y,
                                                               double* __restrict__ ubuf) {
  extern __shared__ double sh[];
  forward_small<false>(A, list[blockIdx.x], moff, minv, front, b, y, ubuf, sh);
}

// Backward: t = D_J⁻¹ y_J − L_SJᵀ x_S, x_J = M_Jᵀ t.  LDS: M_J, L_SJ (as forward), x_S (3 (f−k)), t (3 k).
template <bool Coherent>
__device__ void backward_small(const FrontArgs& A, uint32_t J, const uint32_t* __restrict__ srows,
                               const uint64_t* __restrict__ moff, const double* __restrict__ minv,
                               const double* __restrict__ front, const double* __restrict__ y, double* __restrict__ x,
                               double* sh) {
  const uint32_t f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], ns = f - k, ks = k + 1, tid = threadIdx.x;
  if (k == 0) return;  // uniform per workgroup
  double* Ms = sh;
  double* Ls = Ms + (size_t)k * ks;
  double* xs = Ls + (size_t)ns * ks;
  double* t = xs + 3 * (size_t)ns;
  const double* M = minv + moff[J];
  const double* G = front + A.foff[J];
  const uint32_t* S = srows + A.ea_off[J];
  stage_block(M, G, k, f, sh);
  for (uint32_t r = tid; r < ns; r += 256)
    for (int a = 0; a < 3; ++a) xs[3 * r + a] = ld_v<Coherent>(x + 3 * (size_t)S[r] + a);
  __syncthreads();
  for (uint32_t i = tid; i < k; i += 256) {
    const double d = G[(size_t)i * f + i];
    double s0 = ld_v<Coherent>(y + 3 * (size_t)(j0 + i)) / d, s1 = ld_v<Coherent>(y + 3 * (size_t)(j0 + i) + 1) / d,
           s2 = ld_v<Coherent>(y + 3 * (size_t)(j0 + i) + 2) / d;
    for (uint32_t r = 0; r < ns; ++r) {
      const double l = Ls[r * ks + i];
      s0 = fma(-l, xs[3 * r], s0);
      s1 = fma(-l, xs[3 * r + 1], s1);
      s2 = fma(-l, xs[3 * r + 2], s2);
    }
    t[3 * i] = s0;
    t[3 * i + 1] = s1;
    t[3 * i + 2] = s2;
  }
  __syncthreads();
  for (uint32_t i = tid; i < k; i += 256) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l = i; l < k; ++l) {
      const double m = Ms[l * ks + i];
      s0 = fma(m, t[3 * l], s0);
      s1 = fma(m, t[3 * l + 1], s1);
      s2 = fma(m, t[3 * l + 2], s2);
    }
    st_v<Coherent>(x + 3 * (size_t)(j0 + i), s0);
    st_v<Coherent>(x + 3 * (size_t)(j0 + i) + 1, s1);
    st_v<Coherent>(x + 3 * (size_t)(j0 + i) + 2, s2);
  }
}

__global__ __launch_bounds__(256) void lf_backward_small_kernel(FrontArgs A, const uint32_t* __restrict__ list,
                                                                const uint32_t* __restrict__ srows,
                                                                const uint64_t* __restrict__ moff,
                                                                const double* __restrict__ minv,
                                                                const double* __restrict__ front,
                                                                const double* __restrict__ y, double* __restrict__ x) {
  extern __shared__ double sh[];
  backward_small<false>(A, list[blockIdx.x], srows, moff, minv, front, y, x, sh);
}

// ---- subtrees of the lowest heights: one workgroup per subtree, one front per wave ---------------
//
// A workgroup takes its subtree's blocks level by level (forward from the leaves up, backward from
// the subtree's root down), each wave one block at a time, the levels separated by workgroup
// barriers: a block's update vector (forward) or its later rows' x (backward) come from the same
// workgroup's previous level, or from launches before this one.  The matrices are read where they
// lie, each lane one row of the block (M_J through M_Jᵀ, so a wave's loads of one column are
// contiguous); every sum runs over the same terms in the same order as in the small-front kernels
// above, so the results are bitwise those of one launch per height.

// Loads a lane issues before it uses the first of them.
constexpr uint32_t kSubBatch = 16;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward of block J by one wave.  LDS slice: w (3 f), y_J (3 k).
__device__ void forward_wave(const FrontArgs& A, uint32_t J, const uint64_t* __restrict__ moff,
                             const double* __restrict__ minv_t, const double* __restrict__ front,
                             const double* __restrict__ b, double* __restrict__ y, double* __restrict__ ubuf,
                             double* w, uint32_t lane) {
  const uint32_t f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], ns = f - k;
  double* yj = w + 3 * (size_t)f;
  for (uint32_t i = lane; i < f; i += 64)
    for (int a = 0; a < 3; ++a) w[3 * i + a] = i < k ? b[3 * (size_t)(j0 + i) + a] : 0.0;
  wave_sync();
  for (uint32_t q = A.child_ptr[J]; q < A.child_ptr[J + 1]; ++q) {
    const uint32_t ch = A.child[q], cs = A.bf[ch] - A.bk[ch];
    const uint32_t* map = A.ea_map + A.ea_off[ch];
    const double* uc = ubuf + 3 * (size_t)A.ea_off[ch];
    for (uint32_t r = lane; r < cs; r += 64)
      for (int a = 0; a < 3; ++a) w[3 * map[r] + a] += uc[3 * r + a];
    wave_sync();
  }
  const double* Mt = minv_t + moff[J];
  for (uint32_t i = lane; i < k; i += 64) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l0 = 0; l0 <= i; l0 += kSubBatch) {
      double m[kSubBatch];
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) m[j] = l0 + j <= i ? Mt[(size_t)(l0 + j) * k + i] : 0.0;
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) {
        const uint32_t l = l0 + j;
        if (l <= i) {
          s0 = fma(m[j], w[3 * l], s0);
          s1 = fma(m[j], w[3 * l + 1], s1);
          s2 = fma(m[j], w[3 * l + 2], s2);
        }
      }
    }
    yj[3 * i] = s0;
    yj[3 * i + 1] = s1;
    yj[3 * i + 2] = s2;
    y[3 * (size_t)(j0 + i)] = s0;
    y[3 * (size_t)(j0 + i) + 1] = s1;
    y[3 * (size_t)(j0 + i) + 2] = s2;
  }
  wave_sync();
  const double* G = front + A.foff[J];
  double* u = ubuf + 3 * (size_t)A.ea_off[J];
  for (uint32_t r = lane; r < ns; r += 64) {
    const double* row = G + (size_t)(k + r) * f;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l0 = 0; l0 < k; l0 += kSubBatch) {
      double v[kSubBatch];
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) v[j] = l0 + j < k ? row[l0 + j] : 0.0;
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) {
        const uint32_t l = l0 + j;
        if (l < k) {
          s0 = fma(v[j], yj[3 * l], s0);
          s1 = fma(v[j], yj[3 * l + 1], s1);
          s2 = fma(v[j], yj[3 * l + 2], s2);
        }
      }
    }
    u[3 * r] = w[3 * (k + r)] - s0;
    u[3 * r + 1] = w[3 * (k + r) + 1] - s1;
    u[3 * r + 2] = w[3 * (k + r) + 2] - s2;
  }
  wave_sync();  // the slice is reused by the wave's next block
}

// Backward of block J by one wave.  LDS slice: x_S (3 (f − k)), t (3 k).
__device__ void backward_wave(const FrontArgs& A, uint32_t J, const uint32_t* __restrict__ srows,
                              const uint64_t* __restrict__ moff, const double* __restrict__ minv,
                              const double* __restrict__ front, const double* __restrict__ y, double* __restrict__ x,
                              double* xs, uint32_t lane) {
  const uint32_t f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], ns = f - k;
  if (k == 0) return;  // uniform per wave
  double* t = xs + 3 * (size_t)ns;
  const double* G = front + A.foff[J];
  const uint32_t* S = srows + A.ea_off[J];
  for (uint32_t r = lane; r < ns; r += 64)
    for (int a = 0; a < 3; ++a) xs[3 * r + a] = x[3 * (size_t)S[r] + a];
  wave_sync();
  for (uint32_t i = lane; i < k; i += 64) {
    const double d = G[(size_t)i * f + i];
    double s0 = y[3 * (size_t)(j0 + i)] / d, s1 = y[3 * (size_t)(j0 + i) + 1] / d, s2 = y[3 * (size_t)(j0 + i) + 2] / d;
    for (uint32_t r0 = 0; r0 < ns; r0 += kSubBatch) {
      double v[kSubBatch];
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) v[j] = r0 + j < ns ? G[(size_t)(k + r0 + j) * f + i] : 0.0;
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) {
        const uint32_t r = r0 + j;
        if (r < ns) {
          s0 = fma(-v[j], xs[3 * r], s0);
          s1 = fma(-v[j], xs[3 * r + 1], s1);
          s2 = fma(-v[j], xs[3 * r + 2], s2);
        }
      }
    }
    t[3 * i] = s0;
    t[3 * i + 1] = s1;
    t[3 * i + 2] = s2;
  }
  wave_sync();
  const double* M = minv + moff[J];
  for (uint32_t i = lane; i < k; i += 64) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t l0 = i; l0 < k; l0 += kSubBatch) {
      double m[kSubBatch];
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) m[j] = l0 + j < k ? M[(size_t)(l0 + j) * k + i] : 0.0;
#pragma unroll
      for (uint32_t j = 0; j < kSubBatch; ++j) {
        const uint32_t l = l0 + j;
        if (l < k) {
          s0 = fma(m[j], t[3 * l], s0);
          s1 = fma(m[j], t[3 * l + 1], s1);
          s2 = fma(m[j], t[3 * l + 2], s2);
        }
      }
    }
    x[3 * (size_t)(j0 + i)] = s0;
    x[3 * (size_t)(j0 + i) + 1] = s1;
    x[3 * (size_t)(j0 + i) + 2] = s2;
  }
  wave_sync();
}

// One workgroup per subtree: `levels` level ranges of `fronts` at lvl[blockIdx.x · levels ...].
__global__ __launch_bounds__(64 * kSubWaves) void lf_forward_sub_kernel(
    FrontArgs A, uint32_t levels, uint32_t wave_doubles, const uint32_t* __restrict__ fronts,
    const uint32_t* __restrict__ lvl, const uint64_t* __restrict__ moff, const double* __restrict__ minv_t,
    const double* __restrict__ front, const double* __restrict__ b, double* __restrict__ y, double* __restrict__ ubuf) {
  extern __shared__ double sh[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* w = sh + (size_t)wave * wave_doubles;
  const uint32_t* L = lvl + (size_t)blockIdx.x * levels;
  for (uint32_t h = 0; h < levels; ++h) {
    for (uint32_t q = L[h] + wave; q < L[h + 1]; q += kSubWaves)
      forward_wave(A, fronts[q], moff, minv_t, front, b, y, ubuf, w, lane);
    __syncthreads();
  }
}

__global__ __launch_bounds__(64 * kSubWaves) void lf_backward_sub_kernel(
    FrontArgs A, uint32_t levels, uint32_t wave_doubles, const uint32_t* __restrict__ fronts,
    const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ srows, const uint64_t* __restrict__ moff,
    const double* __restrict__ minv, const double* __restrict__ front, const double* __restrict__ y,
    double* __restrict__ x) {
  extern __shared__ double sh[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* xs = sh + (size_t)wave * wave_doubles;
  const uint32_t* L = lvl + (size_t)blockIdx.x * levels;
  for (uint32_t h = levels; h-- > 0;) {
    for (uint32_t q = L[h] + wave; q < L[h + 1]; q += kSubWaves)
      backward_wave(A, fronts[q], srows, moff, minv, front, y, x, xs, lane);
    __syncthreads();
  }
}

// ---- large fronts: rows spread over many workgroups, four rows per wave -----------------------

// Σ_l row_q[l] v[l] (3 axes) for the four rows of a wave: rows q with length len[q], reduced over
// the wave (every lane gets the sums).
__device__ __forceinline__ void wave_rows4(const double* const row[4], const uint32_t lo[4], const uint32_t hi[4],
                                           const double* v, uint32_t lane, double out[4][3]) {
  double acc[4][3] = {};
  uint32_t lmin = lo[0], lmax = hi[0];
  for (int q = 1; q < 4; ++q) {
    lmin = min(lmin, lo[q]);
    lmax = max(lmax, hi[q]);
  }
  for (uint32_t l0 = lmin; l0 < lmax; l0 += 64) {
    const uint32_t l = l0 + lane;
    double m[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) m[q] = (l >= lo[q] && l < hi[q]) ? row[q][l] : 0.0;
    if (l < lmax) {
      const double v0 = v[3 * l], v1 = v[3 * l + 1], v2 = v[3 * l + 2];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[q][0] = fma(m[q], v0, acc[q][0]);
        acc[q][1] = fma(m[q], v1, acc[q][1]);
        acc[q][2] = fma(m[q], v2, acc[q][2]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    for (int a = 0; a < 3; ++a) out[q][a] = wsum64(acc[q][a]);
}

// Forward y_J = M_J w_J, rows [r0, r0 + 16) (w assembled by each workgroup).  LDS: w (3 f).
template <bool Coherent>
__device__ void forward_y(const FrontArgs& A, SolveItem it, const uint64_t* __restrict__ moff,
                          const double* __restrict__ minv, const double* __restrict__ b, double* __restrict__ y,
                          const double* __restrict__ ubuf, double* w) {
  const uint32_t J = it.block, k = A.bk[J], j0 = A.bstart[J], lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  assemble_front_vector<Coherent>(A, J, b, ubuf, w);
  const double* M = minv + moff[J];
  const double* row[4];
  uint32_t lo[4], hi[4], ri[4];
  for (int q = 0; q < 4; ++q) {
    ri[q] = it.r0 + 4 * wave + q;
    const bool ok = ri[q] < k;
    row[q] = M + (size_t)(ok ? ri[q] : 0) * k;
    lo[q] = 0;
    hi[q] = ok ? ri[q] + 1 : 0;
  }
  double out[4][3];
  wave_rows4(row, lo, hi, w, lane, out);
  if (lane == 0)
    for (int q = 0; q < 4; ++q)
      if (ri[q] < k)
        for (int a = 0; a < 3; ++a) st_v<Coherent>(y + 3 * (size_t)(j0 + ri[q]) + a, out[q][a]);
}

__global__ __launch_bounds__(256) void lf_forward_y_kernel(FrontArgs A, const SolveItem* __restrict__ items,
                                                           const uint64_t* __restrict__ moff,
                                                           const double* __restrict__ minv,
                                                           const double* __restrict__ b, double* __restrict__ y,
                                                           const double* __restrict__ ubuf) {
  extern __shared__ double w[];
  forward_y<false>(A, items[blockIdx.x], moff, minv, b, y, ubuf, w);
}

// Forward u_J = w_S − L_SJ y_J, later rows [r0, r0 + 16).  LDS: w (3 f), then y_J (3 k).
template <bool Coherent>
__device__ void forward_u(const FrontArgs& A, SolveItem it, const double* __restrict__ front,
                          const double* __restrict__ b, const double* __restrict__ y, double* __restrict__ ubuf,
                          double* w) {
  const uint32_t J = it.block, f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], ns = f - k;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* yj = w + 3 * (size_t)f;
  for (uint32_t i = threadIdx.x; i < 3 * k; i += 256) yj[i] = ld_v<Coherent>(y + 3 * (size_t)j0 + i);
  assemble_front_vector<Coherent>(A, J, b, ubuf, w);
  const double* G = front + A.foff[J];
  const double* row[4];
  uint32_t lo[4], hi[4], ri[4];
  for (int q = 0; q < 4; ++q) {
    ri[q] = it.r0 + 4 * wave + q;
    const bool ok = ri[q] < ns;
    row[q] = G + (size_t)(k + (ok ? ri[q] : 0)) * f;
    lo[q] = 0;
    hi[q] = ok ? k : 0;
  }
  double out[4][3];
  wave_rows4(row, lo, hi, yj, lane, out);
  double* u = ubuf + 3 * (size_t)A.ea_off[J];
  if (lane == 0)
    for (int q = 0; q < 4; ++q)
      if (ri[q] < ns)
        for (int a = 0; a < 3; ++a) st_v<Coherent>(u + 3 * ri[q] + a, w[3 * (k + ri[q]) + a] - out[q][a]);
}

__global__ __launch_bounds__(256) void lf_forward_u_kernel(FrontArgs A, const SolveItem* __restrict__ items,
                                                           const double* __restrict__ front,
                                                           const double* __restrict__ b, const double* __restrict__ y,
                                                           double* __restrict__ ubuf) {
  extern __shared__ double w[];
  forward_u<false>(A, items[blockIdx.x], front, b, y, ubuf, w);
}

// Backward t = D_J⁻¹ y_J − L_SJᵀ x_S, rows [r0, r0 + 16) (rows of the transposed L_SJ).
// LDS: x_S (3 (f − k)).
template <bool Coherent>
__device__ void backward_t(const FrontArgs& A, SolveItem it, const uint32_t* __restrict__ srows,
                           const uint64_t* __restrict__ ltoff, const double* __restrict__ lt,
                           const double* __restrict__ front, const double* __restrict__ y,
                           const double* __restrict__ x, double* __restrict__ tbuf, double* xs) {
  const uint32_t J = it.block, f = A.bf[J], k = A.bk[J], j0 = A.bstart[J], ns = f - k;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t* S = srows + A.ea_off[J];
  for (uint32_t r = threadIdx.x; r < ns; r += 256)
    for (int a = 0; a < 3; ++a) xs[3 * r + a] = ld_v<Coherent>(x + 3 * (size_t)S[r] + a);
  __syncthreads();
  const double* T = lt + ltoff[J];
  const double* row[4];
  uint32_t lo[4], hi[4], ri[4];
  for (int q = 0; q < 4; ++q) {
    ri[q] = it.r0 + 4 * wave + q;
    const bool ok = ri[q] < k;
    row[q] = T + (size_t)(ok ? ri[q] : 0) * ns;
    lo[q] = 0;
    hi[q] = ok ? ns : 0;
  }
  double out[4][3];
  wave_rows4(row, lo, hi, xs, lane, out);
  const double* G = front + A.foff[J];
  if (lane == 0)
    for (int q = 0; q < 4; ++q)
      if (ri[q] < k) {
        const size_t i = j0 + ri[q];
        const double d = G[(size_t)ri[q] * f + ri[q]];
        for (int a = 0; a < 3; ++a) st_v<Coherent>(tbuf + 3 * i + a, ld_v<Coherent>(y + 3 * i + a) / d - out[q][a]);
      }
}

__global__ __launch_bounds__(256) void lf_backward_t_kernel(FrontArgs A, const SolveItem* __restrict__ items,
                                                            const uint32_t* __restrict__ srows,
                                                            const uint64_t* __restrict__ ltoff,
                                                            const double* __restrict__ lt,
                                                            const double* __restrict__ front,
                                                            const double* __restrict__ y, const double* __restrict__ x,
                                                            double* __restrict__ tbuf) {
  extern __shared__ double xs[];
  backward_t<false>(A, items[blockIdx.x], srows, ltoff, lt, front, y, x, tbuf, xs);
}

// Backward x_J = M_Jᵀ t, rows [r0, r0 + 16) (rows of M_Jᵀ).  LDS: t (3 k).
template <bool Coherent>
__device__ void backward_x(const FrontArgs& A, SolveItem it, const uint64_t* __restrict__ moff,
                           const double* __restrict__ minv_t, const double* __restrict__ tbuf, double* __restrict__ x,
                           double* t) {
  const uint32_t J = it.block, k = A.bk[J], j0 = A.bstart[J], lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < 3 * k; i += 256) t[i] = ld_v<Coherent>(tbuf + 3 * (size_t)j0 + i);
  __syncthreads();
  const double* Mt = minv_t + moff[J];
  const double* row[4];
  uint32_t lo[4], hi[4], ri[4];
  for (int q = 0; q < 4; ++q) {
    ri[q] = it.r0 + 4 * wave + q;
    const bool ok = ri[q] < k;
    row[q] = Mt + (size_t)(ok ? ri[q] : 0) * k;
    lo[q] = ok ? ri[q] : 0;
    hi[q] = ok ? k : 0;
  }
  double out[4][3];
  wave_rows4(row, lo, hi, t, lane, out);
  if (lane == 0)
    for (int q = 0; q < 4; ++q)
      if (ri[q] < k)
        for (int a = 0; a < 3; ++a) st_v<Coherent>(x + 3 * (size_t)(j0 + ri[q]) + a, out[q][a]);
}

__global__ __launch_bounds__(256) void lf_backward_x_kernel(FrontArgs A, const SolveItem* __restrict__ items,
                                                            const uint64_t* __restrict__ moff,
                                                            const double* __restrict__ minv_t,
                                                            const double* __restrict__ tbuf, double* __restrict__ x) {
  extern __shared__ double t[];
  backward_x<false>(A, items[blockIdx.x], moff, minv_t, tbuf, x, t);
}

// ---- large fronts: the forward in one phase -----------------------------------------------------
//
// With P_J = L_SJ M_J (formed once per factorization) the update vector needs no y_J:
// u_J = w_S − P_J w_J, so the y and u rows of a height are one launch (the same quantity as
// w_S − L_SJ y_J, summed in another order, FP64).  The backward keeps its two phases: the one-phase
// form x_J = M_Jᵀ (D_J⁻¹ y_J) − P_Jᵀ x_S subtracts two large terms and raised L′'s residual 30-fold.

// P_J = L_SJ M_J (rows r < f − k, columns l < k; M_J lower triangular: m from l), one 32 × 32 tile
// per workgroup (the transpose tiles' items), m ascending.
__global__ __launch_bounds__(256) void lf_pmat_kernel(FrontArgs A, const FrontTile* __restrict__ items,
                                                      const uint64_t* __restrict__ ltoff,
                                                      const uint64_t* __restrict__ moff,
                                                      const double* __restrict__ minv,
                                                      const double* __restrict__ front, double* __restrict__ pj) {
  const FrontTile it = items[blockIdx.x];  // block, i0: first later row r, j0: first column l
  const uint32_t J = it.block, f = A.bf[J], k = A.bk[J], ns = f - k;
  const double* G = front + A.foff[J];
  const double* M = minv + moff[J];
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5, l = it.j0 + tx;
  if (l >= k) return;
  for (uint32_t q = 0; q < 4; ++q) {
    const uint32_t r = it.i0 + ty + 8 * q;
    if (r >= ns) break;
    const double* L = G + (size_t)(k + r) * f;
    double acc = 0.0;
    for (uint32_t m = l; m < k; ++m) acc = fma(L[m], M[(size_t)m * k + l], acc);
    pj[ltoff[J] + (size_t)r * k + l] = acc;
  }
}

// Forward u_J = w_S − P_J w_J, later rows [r0, r0 + 16).  LDS: w (3 f).
__device__ void forward_up(const FrontArgs& A, SolveItem it, const uint64_t* __restrict__ ltoff,
                           const double* __restrict__ pj, const double* __restrict__ b, double* __restrict__ ubuf,
                           double* w) {
  const uint32_t J = it.block, f = A.bf[J], k = A.bk[J], ns = f - k;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  assemble_front_vector<false>(A, J, b, ubuf, w);
  const double* P = pj + ltoff[J];
  const double* row[4];
  uint32_t lo[4], hi[4], ri[4];
  for (int q = 0; q < 4; ++q) {
    ri[q] = it.r0 + 4 * wave + q;
    const bool ok = ri[q] < ns;
    row[q] = P + (size_t)(ok ? ri[q] : 0) * k;
    lo[q] = 0;
    hi[q] = ok ? k : 0;
  }
  double out[4][3];
  wave_rows4(row, lo, hi, w, lane, out);
  double* u = ubuf + 3 * (size_t)A.ea_off[J];
  if (lane == 0)
    for (int q = 0; q < 4; ++q)
      if (ri[q] < ns)
        for (int a = 0; a < 3; ++a) u[3 * ri[q] + a] = w[3 * (k + ri[q]) + a] - out[q][a];
}

// One launch per large height: items [0, ny) are y rows (as the two-phase forward), the rest u rows.
__global__ __launch_bounds__(256) void lf_forward_yu_kernel(FrontArgs A, const SolveItem* __restrict__ yitems,
                                                            uint32_t ny, const SolveItem* __restrict__ uitems,
                                                            const uint64_t* __restrict__ moff,
                                                            const double* __restrict__ minv,
                                                            const uint64_t* __restrict__ ltoff,
                                                            const double* __restrict__ pj,
                                                            const double* __restrict__ b, double* __restrict__ y,
                                                            double* __restrict__ ubuf) {
  extern __shared__ double w[];
  if (blockIdx.x < ny)
    forward_y<false>(A, yitems[blockIdx.x], moff, minv, b, y, ubuf, w);
  else
    forward_up(A, uitems[blockIdx.x - ny], ltoff, pj, b, ubuf, w);
}

// ---- the whole solve in one launch ----------------------------------------------------------------
//
// Workgroups take tasks in ticket order (an atomic counter): the forward tasks by height, then the
// backward tasks from the root down; within a large block its y (t) tasks precede its u (x) tasks.
// A task waits only for tasks of smaller tickets:
//   forward small / y of J  — every forward task of J's children (their update vectors);
//   forward u of J          — the y tasks of J;
//   backward small / t of J — every forward task of J (y_J) and every backward task of J's parent
//                             (x of J's later rows: the parent's rows and, through the parent's own
//                             wait, those of every ancestor);
//   backward x of J         — the t tasks of J.
// A workgroup only waits after it has taken a ticket, so the smallest unfinished ticket always
// belongs to a running workgroup whose dependencies are finished: the launch completes whatever
// the number of resident workgroups.  A task's results are published (device-scope stores, then
// the stores' completion, then a workgroup barrier) before its block's counter is incremented.
// Counters are read with a read-modify-write atomic (a compare-and-swap that never matches), the
// kind of operation that increments them, never by a repeated load: the hand-offs elsewhere in this
// library read each published line once, and a load repeated on one line may be served by a copy
// in the XCD's own L2.  The wait gives up after about 0.2 s of the device's constant-rate clock or
// 65536 polls, whichever comes first (a broken invariant, not a slow solve: the whole solve takes
// well under a millisecond), marks the solve failed and ends the workgroup; the scatter then writes
// NaN.

constexpr uint64_t kSolveWaitTicks = 20000000;  // s_memrealtime runs at 100 MHz
constexpr uint32_t kSolveWaitPolls = 65536;

__device__ __forceinline__ uint32_t read_count(uint32_t* c) { return atomicCAS(c, 0xffffffffu, 0xffffffffu); }

__device__ bool wait_count(uint32_t* c, uint32_t need, uint32_t* failed) {
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  for (uint32_t spin = 0; read_count(c) < need; ++spin) {
    if (spin >= kSolveWaitPolls ||
        ((spin & 63u) == 63u &&
         (read_count(failed) != 0u || __builtin_amdgcn_s_memrealtime() - t0 > kSolveWaitTicks))) {
      atomicOr(failed, 1u);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

struct OneArgs {
  const SolveTask* tasks;
  const TaskCounts* counts;
  uint32_t ntasks;
  uint32_t *fdone, *bdone, *ticket, *failed;
};

__global__ __launch_bounds__(256) void lf_solve_one_kernel(FrontArgs A, OneArgs O, const uint32_t* __restrict__ srows,
                                                           const uint64_t* __restrict__ moff,
                                                           const double* __restrict__ minv,
                                                           const double* __restrict__ minv_t,
                                                           const double* __restrict__ front,
                                                           const uint64_t* __restrict__ ltoff,
                                                           const double* __restrict__ lt, double* __restrict__ xb,
                                                           double* __restrict__ yb, double* __restrict__ ubuf,
                                                           double* __restrict__ tb) {
  extern __shared__ double sh[];
  __shared__ uint32_t s_task, s_ok;
  // Thread 0's work between two tasks is one region at the top of the loop: the previous task's
  // completion, then the next ticket.  (With the completion at the bottom of the loop and the ticket
  // at the top, the compiler merged the two thread-0 regions across the back edge into a loop exit
  // of thread 0 alone; the other lanes of its wave then reached the next barrier first and the
  // workgroup re-read the old ticket forever.)  The ticket and the wait's outcome are read as
  // wave-uniform values, so that no loop exit looks divergent.
  // The loop is bounded on its own: a workgroup takes at most ntasks + 1 tickets (each iteration
  // takes a new one, so a correct run leaves after at most that many), and leaves as soon as
  // another workgroup has flagged a failed wait.
  uint32_t* done = nullptr;  // the counter of the task just finished
  for (uint32_t iter = 0; iter <= O.ntasks; ++iter) {
    if (threadIdx.x == 0) {
      if (done) atomicAdd(done, 1u);
      s_task = read_count(O.failed) != 0u ? O.ntasks : atomicAdd(O.ticket, 1u);
    }
    __syncthreads();
    const uint32_t tk = __builtin_amdgcn_readfirstlane(s_task);
    if (tk >= O.ntasks) return;
    const SolveTask task = O.tasks[tk];
    const uint32_t J = task.block;
    if (threadIdx.x == 0) {
      bool ok = true;
      switch (task.kind) {
        case kTaskFwdSmall:
        case kTaskFwdY:
          for (uint32_t q = A.child_ptr[J]; ok && q < A.child_ptr[J + 1]; ++q) {
            const uint32_t c = A.child[q];
            ok = wait_count(O.fdone + c, O.counts[c].nf, O.failed);
          }
          break;
        case kTaskFwdU:
          ok = wait_count(O.fdone + J, O.counts[J].ny, O.failed);
          break;
        case kTaskBwdSmall:
        case kTaskBwdT: {
          ok = wait_count(O.fdone + J, O.counts[J].nf, O.failed);
          const uint32_t P = A.bparent[J];
          if (ok && P != kNone) ok = wait_count(O.bdone + P, O.counts[P].nb, O.failed);
          break;
        }
        default:
          ok = wait_count(O.bdone + J, O.counts[J].nt, O.failed);
          break;
      }
      s_ok = ok ? 1u : 0u;
    }
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_ok)) return;
    const SolveItem it{J, task.r0};
    switch (task.kind) {
      case kTaskFwdSmall:
        forward_small<true>(A, J, moff, minv, front, xb, yb, ubuf, sh);
        break;
      case kTaskFwdY:
        forward_y<true>(A, it, moff, minv, xb, yb, ubuf, sh);
        break;
      case kTaskFwdU:
        forward_u<true>(A, it, front, xb, yb, ubuf, sh);
        break;
      case kTaskBwdSmall:
        backward_small<true>(A, J, srows, moff, minv, front, yb, xb, sh);
        break;
      case kTaskBwdT:
        backward_t<true>(A, it, srows, ltoff, lt, front, yb, xb, tb, sh);
        break;
      default:
        backward_x<true>(A, it, moff, minv_t, tb, xb, sh);
        break;
    }
    published();  // this thread's results have reached the coherent level
    __syncthreads();
    done = (task.kind < kTaskBwdSmall ? O.fdone : O.bdone) + J;
  }
  if (threadIdx.x == 0) atomicOr(O.failed, 1u);  // more tickets than tasks: the solve is invalid
}

// L_SJᵀ of the large fronts (k × (f − k), row-major), 32 × 32 tiles through LDS.
__global__ __launch_bounds__(256) void lf_transpose_kernel(FrontArgs A, const FrontTile* __restrict__ items,
                                                           const uint64_t* __restrict__ ltoff,
                                                           const double* __restrict__ front, double* __restrict__ lt) {
  __shared__ double tile[32][33];
  const FrontTile it = items[blockIdx.x];  // block, i0: first later row r, j0: first column i
  const uint32_t J = it.block, f = A.bf[J], k = A.bk[J], ns = f - k;
  const double* G = front + A.foff[J];
  double* T = lt + ltoff[J];
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (uint32_t yy = ty; yy < 32; yy += 8) {
    const uint32_t r = it.i0 + yy, i = it.j0 + tx;
    tile[yy][tx] = (r < ns && i < k) ? G[(size_t)(k + r) * f + i] : 0.0;
  }
  __syncthreads();
  for (uint32_t yy = ty; yy < 32; yy += 8) {
    const uint32_t i = it.j0 + yy, r = it.i0 + tx;
    if (i < k && r < ns) T[(size_t)i * ns + r] = tile[tx][yy];
  }
}

// ---- host side -------------------------------------------------------------------------------------

LprimeFactorDev::~LprimeFactorDev() {
  if (bad_host) (void)hipHostFree(bad_host);
}

void solve_tasks(const LprimeSymbolic& S, std::vector<SolveTask>& tl, std::vector<TaskCounts>& tc) {
  tl.clear();
  tc.assign(S.nb, TaskCounts{0, 0, 0, 0});
  for (uint32_t h = 0; h < S.H; ++h) {
    if (!S.hsplit[h]) {
      for (uint32_t q = S.hoff[h]; q < S.hoff[h + 1]; ++q) {
        tl.push_back({kTaskFwdSmall, S.hblocks[q], 0});
        tc[S.hblocks[q]].nf++;
      }
      continue;
    }
    for (uint32_t q = S.krows_off[h]; q < S.krows_off[h + 1]; ++q) {
      tl.push_back({kTaskFwdY, S.krows[q].block, S.krows[q].r0});
      tc[S.krows[q].block].nf++;
      tc[S.krows[q].block].ny++;
    }
    for (uint32_t q = S.urows_off[h]; q < S.urows_off[h + 1]; ++q) {
      tl.push_back({kTaskFwdU, S.urows[q].block, S.urows[q].r0});
      tc[S.urows[q].block].nf++;
    }
  }
  for (uint32_t h = S.H; h-- > 0;) {
    if (!S.hsplit[h]) {
      for (uint32_t q = S.hoff[h]; q < S.hoff[h + 1]; ++q) {
        tl.push_back({kTaskBwdSmall, S.hblocks[q], 0});
        tc[S.hblocks[q]].nb++;
      }
      continue;
    }
    for (uint32_t q = S.krows_off[h]; q < S.krows_off[h + 1]; ++q) {
      tl.push_back({kTaskBwdT, S.krows[q].block, S.krows[q].r0});
      tc[S.krows[q].block].nb++;
      tc[S.krows[q].block].nt++;
    }
    for (uint32_t q = S.krows_off[h]; q < S.krows_off[h + 1]; ++q) {
      tl.push_back({kTaskBwdX, S.krows[q].block, S.krows[q].r0});
      tc[S.krows[q].block].nb++;
    }
  }
}

void LprimeFactorDev::setup(ddgrs_ctx* ctx, const LprimeSymbolic& S) {
  nv = S.nv;
  n = S.n;
  H = S.H;
  max_children = S.max_children;
  max_rows = S.max_rows;
  max_front = S.max_front;
  hoff = S.hoff;
  hsplit = S.hsplit;
  hlds = S.hlds;
  krows_off = S.krows_off;
  urows_off = S.urows_off;
  front_total = S.front_total;
  lds_off = S.lds_off;
  lds_rows = S.lds_rows;
  ea_item_off = S.ea_item_off;
  step_off = S.step_off;
  step_c0 = S.step_c0;
  panel_off = S.panel_off;
  tile_off = S.tile_off;
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
  ddgrs::upload(ctx, perm, S.perm);
  ddgrs::upload(ctx, block_of, S.block_of);
  ddgrs::upload(ctx, hblocks, S.hblocks);
  ddgrs::upload(ctx, krows, S.krows);
  ddgrs::upload(ctx, urows, S.urows);
  ddgrs::upload(ctx, ltoff, S.ltoff);
  ddgrs::upload(ctx, lt_tiles, S.lt_tiles);
  n_lt_tiles = (uint32_t)S.lt_tiles.size();
  lt.reserve(std::max<uint64_t>(S.ltoff.back(), 1));
  pj.reserve(std::max<uint64_t>(S.ltoff.back(), 1));
  tb.reserve(3 * n);
  ddgrs::upload(ctx, srows, S.srows);
  ddgrs::upload(ctx, moff, S.moff);
  minv.reserve(std::max<uint64_t>(S.moff.back(), 1));
  minv_t.reserve(std::max<uint64_t>(S.moff.back(), 1));
  ubuf.reserve(3 * std::max<size_t>(S.ea_off.back(), 1));
  yb.reserve(3 * n);
  xb.reserve(3 * n);
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
  nb = S.nb;
  std::vector<SolveTask> tl;
  std::vector<TaskCounts> tc;
  solve_tasks(S, tl, tc);
  size_t small_lds = 0;
  bool any_split = false;
  for (uint32_t h = 0; h < H; ++h) {
    if (hsplit[h])
      any_split = true;
    else
      small_lds = std::max<size_t>(small_lds, hlds[h]);
  }
  sub_levels = S.sub_levels;
  sub_wave_doubles = S.sub_wave_doubles;
  nsub = sub_levels ? (uint32_t)(S.sub_lvl_off.size() - 1) / sub_levels : 0;
  if (nsub) {
    ddgrs::upload(ctx, sub_fronts, S.sub_fronts);
    ddgrs::upload(ctx, sub_lvl_off, S.sub_lvl_off);
  }
  ntasks = (uint32_t)tl.size();
  ddgrs::upload(ctx, tasks, tl);
  ddgrs::upload(ctx, task_counts, tc);
  task_done.reserve(2 * (size_t)nb + 1);
  solve_failed.reserve(1);
  launch(ctx, "lf_reset", lf_reset_kernel, dim3(1), dim3(1), 0, bad.ptr, solve_failed.ptr);
  one_lds = std::max(small_lds, any_split ? (3 * (size_t)max_front + 3 * (size_t)max_rows) * sizeof(double) : 0);
  one_ok = one_lds + 64 <= (size_t)ctx->prop.sharedMemPerBlock;  // and its two static words
  if (one_ok && one_lds > 0)
    HIPCHECK(hipFuncSetAttribute((const void*)lf_solve_one_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)one_lds));
  size_t lds_max = 0;
  for (uint32_t r : lds_rows) lds_max = std::max(lds_max, ((size_t)r * (r + 1) / 2 + r) * sizeof(double));
  if (lds_max > 0)
    HIPCHECK(hipFuncSetAttribute((const void*)lf_lds_front_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds_max));
  ctx->sync();  // the symbolic structure may be released after return
}

void LprimeFactorDev::factor(ddgrs_ctx* ctx, ddgrs_surface* s, double eps_rel) {
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
  launch(ctx, "lf_reset", lf_reset_kernel, dim3(1), dim3(1), 0, bad.ptr, solve_failed.ptr);
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
  // the dense inverses of the diagonal blocks
  launch(ctx, "lf_inverse", lf_inverse_kernel, dim3(blocks_for((int64_t)n, 4)), dim3(256),
         4 * (size_t)max_rows * sizeof(double), (uint32_t)n, max_rows, A, (const uint32_t*)block_of.ptr,
         (const uint64_t*)moff.ptr, (const double*)front.ptr, minv.ptr, minv_t.ptr);
  launch(ctx, "lf_transpose", lf_transpose_kernel, dim3(n_lt_tiles), dim3(256), 0, A,
         (const FrontTile*)lt_tiles.ptr, (const uint64_t*)ltoff.ptr, (const double*)front.ptr, lt.ptr);
  launch(ctx, "lf_pmat", lf_pmat_kernel, dim3(n_lt_tiles), dim3(256), 0, A, (const FrontTile*)lt_tiles.ptr,
         (const uint64_t*)ltoff.ptr, (const uint64_t*)moff.ptr, (const double*)minv.ptr, (const double*)front.ptr,
         pj.ptr);
  download(ctx, bad_host, (const uint32_t*)bad.ptr, 1);
}

namespace {
// Per-height kernel names (timings by height of the tree).
const char* height_name(const char* base, uint32_t h) {
  static const char* fwd[] = {"lf_forward_h00", "lf_forward_h01", "lf_forward_h02", "lf_forward_h03", "lf_forward_h04",
                              "lf_forward_h05", "lf_forward_h06", "lf_forward_h07", "lf_forward_h08", "lf_forward_h09",
                              "lf_forward_h10", "lf_forward_h11", "lf_forward_h12", "lf_forward_h13", "lf_forward_h14",
                              "lf_forward_h15"};
  static const char* bwd[] = {"lf_backward_h00", "lf_backward_h01", "lf_backward_h02", "lf_backward_h03",
                              "lf_backward_h04", "lf_backward_h05", "lf_backward_h06", "lf_backward_h07",
                              "lf_backward_h08", "lf_backward_h09", "lf_backward_h10", "lf_backward_h11",
                              "lf_backward_h12", "lf_backward_h13", "lf_backward_h14", "lf_backward_h15"};
  if (h > 15) return base;
  return base[3] == 'f' ? fwd[h] : bwd[h];
}
}  // namespace

void LprimeFactorDev::solve(ddgrs_ctx* ctx, const double* r, double* x) {
  const uint32_t un = (uint32_t)n, unv = (uint32_t)nv;
  const FrontArgs A{bstart.ptr, bk.ptr, bf.ptr, bparent.ptr, child_ptr.ptr, child.ptr, ea_off.ptr, ea_map.ptr, foff.ptr};
  // The one-launch solve hung on the device twice and its restructured loop has not run there
  // since; until it has, asking for it is an error rather than a launch.
  const char* e = std::getenv("DDGRS_LPRIME_SOLVE_ONE_LAUNCH");
  if (e && e[0] == '1')
    invalid("DDGRS_LPRIME_SOLVE_ONE_LAUNCH", e, "unset: the one-launch L' solve is disabled (it hung on the device)");
  const bool one = false;
  const uint32_t nzero = one ? 2 * nb + 1 : 0;
  launch(ctx, "lf_gather_rhs", lf_gather_rhs_kernel, dim3(blocks_for(std::max<int64_t>(3 * (int64_t)n, nzero), 256)),
         dim3(256), 0, un, unv, (const uint32_t*)perm.ptr, r, xb.ptr, nzero, task_done.ptr);
  if (one) {
    const OneArgs O{tasks.ptr, task_counts.ptr, ntasks, task_done.ptr, task_done.ptr + nb, task_done.ptr + 2 * nb,
                    solve_failed.ptr};
    const uint32_t grid = std::min<uint32_t>(ntasks, 2 * (uint32_t)ctx->prop.multiProcessorCount);
    if (grid > 0)
      launch(ctx, "lf_solve_one", lf_solve_one_kernel, dim3(grid), dim3(256), one_lds, A, O, (const uint32_t*)srows.ptr,
             (const uint64_t*)moff.ptr, (const double*)minv.ptr, (const double*)minv_t.ptr, (const double*)front.ptr,
             (const uint64_t*)ltoff.ptr, (const double*)lt.ptr, xb.ptr, yb.ptr, ubuf.ptr, tb.ptr);
    launch(ctx, "lf_scatter_sol", lf_scatter_sol_kernel, dim3(blocks_for(3 * (int64_t)n, 256)), dim3(256), 0, un, unv,
           (const uint32_t*)perm.ptr, (const double*)xb.ptr, (const uint32_t*)solve_failed.ptr, x);
    return;
  }
  const size_t lds = (3 * (size_t)max_front + 3 * (size_t)max_rows) * sizeof(double);
  const char* tp = std::getenv("DDGRS_LPRIME_SOLVE_TWO_PHASE");
  const bool two_phase = tp && tp[0] == '1';
  const size_t sub_lds = (size_t)kSubWaves * sub_wave_doubles * sizeof(double);
  if (nsub)
    launch(ctx, "lf_forward_sub", lf_forward_sub_kernel, dim3(nsub), dim3(64 * kSubWaves), sub_lds, A, sub_levels,
           sub_wave_doubles, (const uint32_t*)sub_fronts.ptr, (const uint32_t*)sub_lvl_off.ptr,
           (const uint64_t*)moff.ptr, (const double*)minv_t.ptr, (const double*)front.ptr, (const double*)xb.ptr,
           yb.ptr, ubuf.ptr);
  for (uint32_t h = sub_levels; h < H; ++h) {
    if (hsplit[h] && !two_phase) {
      const uint32_t ny = krows_off[h + 1] - krows_off[h], nu = urows_off[h + 1] - urows_off[h];
      launch(ctx, height_name("lf_forward", h), lf_forward_yu_kernel, dim3(ny + nu), dim3(256), lds, A,
             (const SolveItem*)krows.ptr + krows_off[h], ny, (const SolveItem*)urows.ptr + urows_off[h],
             (const uint64_t*)moff.ptr, (const double*)minv.ptr, (const uint64_t*)ltoff.ptr, (const double*)pj.ptr,
             (const double*)xb.ptr, yb.ptr, ubuf.ptr);
      continue;
    }
    if (!hsplit[h]) {
      launch(ctx, height_name("lf_forward", h), lf_forward_small_kernel, dim3(hoff[h + 1] - hoff[h]), dim3(256),
             (size_t)hlds[h], A, (const uint32_t*)hblocks.ptr + hoff[h], (const uint64_t*)moff.ptr,
             (const double*)minv.ptr, (const double*)front.ptr, (const double*)xb.ptr, yb.ptr, ubuf.ptr);
    } else {
      launch(ctx, height_name("lf_forward", h), lf_forward_y_kernel, dim3(krows_off[h + 1] - krows_off[h]), dim3(256),
             lds, A, (const SolveItem*)krows.ptr + krows_off[h], (const uint64_t*)moff.ptr, (const double*)minv.ptr,
             (const double*)xb.ptr, yb.ptr, (const double*)ubuf.ptr);
      launch(ctx, height_name("lf_forward", h), lf_forward_u_kernel, dim3(urows_off[h + 1] - urows_off[h]), dim3(256),
             lds, A, (const SolveItem*)urows.ptr + urows_off[h], (const double*)front.ptr, (const double*)xb.ptr,
             (const double*)yb.ptr, ubuf.ptr);
    }
  }
  for (uint32_t h = H; h-- > sub_levels;) {
    if (!hsplit[h]) {
      launch(ctx, height_name("lf_backward", h), lf_backward_small_kernel, dim3(hoff[h + 1] - hoff[h]), dim3(256),
             (size_t)hlds[h], A, (const uint32_t*)hblocks.ptr + hoff[h], (const uint32_t*)srows.ptr,
             (const uint64_t*)moff.ptr, (const double*)minv.ptr, (const double*)front.ptr, (const double*)yb.ptr, xb.ptr);
    } else {
      launch(ctx, height_name("lf_backward", h), lf_backward_t_kernel, dim3(krows_off[h + 1] - krows_off[h]), dim3(256),
             lds, A, (const SolveItem*)krows.ptr + krows_off[h], (const uint32_t*)srows.ptr, (const uint64_t*)ltoff.ptr,
             (const double*)lt.ptr, (const double*)front.ptr, (const double*)yb.ptr, (const double*)xb.ptr, tb.ptr);
      launch(ctx, height_name("lf_backward", h), lf_backward_x_kernel, dim3(krows_off[h + 1] - krows_off[h]), dim3(256),
             lds, A, (const SolveItem*)krows.ptr + krows_off[h], (const uint64_t*)moff.ptr, (const double*)minv_t.ptr,
             (const double*)tb.ptr, xb.ptr);
    }
  }
  if (nsub)
    launch(ctx, "lf_backward_sub", lf_backward_sub_kernel, dim3(nsub), dim3(64 * kSubWaves), sub_lds, A, sub_levels,
           sub_wave_doubles, (const uint32_t*)sub_fronts.ptr, (const uint32_t*)sub_lvl_off.ptr,
           (const uint32_t*)srows.ptr, (const uint64_t*)moff.ptr, (const double*)minv.ptr, (const double*)front.ptr,
           (const double*)yb.ptr, xb.ptr);
  launch(ctx, "lf_scatter_sol", lf_scatter_sol_kernel, dim3(blocks_for(3 * (int64_t)n, 256)), dim3(256), 0, un, unv,
         (const uint32_t*)perm.ptr, (const double*)xb.ptr, (const uint32_t*)solve_failed.ptr, x);
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
