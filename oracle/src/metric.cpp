// FP64 oracle — TEST INFRASTRUCTURE ONLY: the block cluster tree and the hierarchical products
// with B, B⁰ (face pairs) and the vertex-pair fractional operator.
//
// Restates:
//   crates/ddg-repulsive-surfaces/src/spatial/block_tree.rs
//       build_impl :129 (classification :160-169, levels :172-202, far :205-217, near offsets
//       :222-247, canonical fill :261-294, mirrors :297-333, chains :348-372), multiply :468,
//       multiply_core :536 (up :545-562, far :564-578, down :580-591, rows :592-620),
//       leaf_rows :647, accumulate :712, group_by_row :741
//   crates/ddg-repulsive-surfaces/src/metric/hierarchical.rs
//       separated :116, hi_lo :185, FaceKernels :192-236, FaceMetricTree::from_parts :323,
//       apply_batch :540, product :696, VertexFractionalTree::build :1017, apply_batch :1274,
//       product :1313
//   crates/ddg-repulsive-surfaces/src/metric/mod.rs:120 df_coefficients
#include <algorithm>

#include "oracle.hpp"

namespace orc {

BlockTree block_tree_build(const Bvh& bvh, const ClusterKernelFns& K, int N, size_t leaf_pairs, bool fill) {
  BlockTree t;
  t.N = N;
  const size_t nn = bvh.nodes.size(), n = bvh.order.size();
  t.start.resize(nn);
  t.end.resize(nn);
  t.right.resize(nn);
  for (size_t i = 0; i < nn; ++i) {
    t.start[i] = bvh.nodes[i].start;
    t.end[i] = bvh.nodes[i].end;
    t.right[i] = bvh.nodes[i].right;
  }
  t.order = bvh.order;
  t.position.resize(n);
  for (size_t p = 0; p < n; ++p) t.position[t.order[p]] = (uint32_t)p;
  t.weight.resize(n);
  for (size_t p = 0; p < n; ++p) t.weight[p] = K.weight(t.order[p]);
  auto is_leaf = [&](size_t a) { return t.right[a] == 0; };
  auto len = [&](size_t a) { return (size_t)(t.end[a] - t.start[a]); };

  std::vector<std::pair<uint32_t, uint32_t>> adm, nearb, level{{0, 0}};
  size_t depth = 0;
  while (!level.empty()) {
    depth++;
    std::vector<std::pair<uint32_t, uint32_t>> next;
    for (auto [a, b] : level) {
      int cls;
      if (a != b && K.admissible(a, b))
        cls = 0;
      else if ((is_leaf(a) && is_leaf(b)) || len(a) * len(b) <= leaf_pairs)
        cls = 1;
      else
        cls = 2;
      if (cls == 0)
        adm.push_back({a, b});
      else if (cls == 1)
        nearb.push_back({a, b});
      else {
        uint32_t ka[2], kb[2];
        int na = 1, nb = 1;
        if (is_leaf(a)) ka[0] = a; else { ka[0] = a + 1; ka[1] = t.right[a]; na = 2; }
        if (is_leaf(b)) kb[0] = b; else { kb[0] = b + 1; kb[1] = t.right[b]; nb = 2; }
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) next.push_back({ka[i], kb[j]});
      }
    }
    level = std::move(next);
  }
  auto group = [&](const std::vector<std::pair<uint32_t, uint32_t>>& pairs, std::vector<uint64_t>* ptr,
                   std::vector<uint32_t>* col) {
    ptr->assign(nn + 1, 0);
    for (auto& pr : pairs) (*ptr)[pr.first + 1]++;
    for (size_t i = 0; i < nn; ++i) (*ptr)[i + 1] += (*ptr)[i];
    std::vector<uint64_t> fill = *ptr;
    col->assign(pairs.size(), 0);
    for (auto& pr : pairs) (*col)[fill[pr.first]++] = pr.second;
  };
  group(adm, &t.far_ptr, &t.far_col);
  t.far_coef.resize(t.far_col.size() * N);
  for (size_t a = 0; a < nn; ++a)
    for (uint64_t q = t.far_ptr[a]; q < t.far_ptr[a + 1]; ++q) K.far(a, t.far_col[q], &t.far_coef[q * N]);
  t.stats.admissible_blocks = adm.size();
  t.stats.inadmissible_blocks = nearb.size();
  for (auto [a, b] : adm) t.stats.admissible_pairs += (uint64_t)len(a) * len(b);
  t.stats.depth = depth;
  t.stats.nodes = nn;

  group(nearb, &t.near_ptr, &t.near_col);
  const size_t nblk = t.near_col.size();
  t.near_off.assign(nblk, 0);
  std::vector<uint32_t> block_row(nblk);
  for (size_t a = 0; a < nn; ++a)
    for (uint64_t q = t.near_ptr[a]; q < t.near_ptr[a + 1]; ++q) block_row[q] = (uint32_t)a;
  uint64_t canonical = 0;
  for (size_t q = 0; q < nblk; ++q)
    if (block_row[q] <= t.near_col[q]) {
      t.near_off[q] = canonical;
      canonical += len(block_row[q]) * len(t.near_col[q]);
    }
  uint64_t total = canonical;
  for (size_t q = 0; q < nblk; ++q)
    if (block_row[q] > t.near_col[q]) {
      t.near_off[q] = total;
      total += len(block_row[q]) * len(t.near_col[q]);
    }
  t.near_total = total;
  for (auto [a, b] : nearb) t.stats.near_entries += (a == b) ? len(a) * (len(a) - 1) : len(a) * len(b);
  if (fill) {
    t.near_val.assign(total * N, 0.0);
    std::vector<std::pair<uint64_t, uint32_t>> keys(nblk);
    for (size_t q = 0; q < nblk; ++q) keys[q] = {((uint64_t)block_row[q] << 32) | t.near_col[q], (uint32_t)q};
    std::sort(keys.begin(), keys.end());
    auto find = [&](size_t a, size_t b) -> int64_t {
      uint64_t k = ((uint64_t)a << 32) | b;
      auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(k, (uint32_t)0));
      if (it != keys.end() && it->first == k) return it->second;
      return -1;
    };
    for (size_t q = 0; q < nblk; ++q) {
      size_t a = block_row[q], b = t.near_col[q];
      if (a > b) continue;
      size_t sa = t.start[a], sb = t.start[b], ra = len(a), cb = len(b);
      double* blk = &t.near_val[t.near_off[q] * N];
      for (size_t r = 0; r < ra; ++r) {
        uint32_t i = t.order[sa + r];
        size_t c0 = (a == b) ? r + 1 : 0;
        for (size_t c = c0; c < cb; ++c) K.near(i, t.order[sb + c], &blk[(r * cb + c) * N]);
      }
      if (a == b)
        for (size_t r = 0; r < ra; ++r)
          for (size_t c = 0; c < r; ++c)
            for (int k = 0; k < N; ++k) blk[(r * cb + c) * N + k] = blk[(c * cb + r) * N + k];
    }
    for (size_t q = 0; q < nblk; ++q) {
      size_t a = block_row[q], b = t.near_col[q];
      if (a <= b) continue;
      size_t ra = len(a), cb = len(b);
      double* blk = &t.near_val[t.near_off[q] * N];
      int64_t m = find(b, a);
      if (m >= 0) {
        const double* src = &t.near_val[t.near_off[m] * N];
        for (size_t r = 0; r < ra; ++r)
          for (size_t c = 0; c < cb; ++c)
            for (int k = 0; k < N; ++k) blk[(r * cb + c) * N + k] = src[(c * ra + r) * N + k];
      } else {
        size_t sa = t.start[a], sb = t.start[b];
        for (size_t r = 0; r < ra; ++r)
          for (size_t c = 0; c < cb; ++c) K.near(t.order[sa + r], t.order[sb + c], &blk[(r * cb + c) * N]);
      }
    }
  }
  // Leaves and their chains root → leaf.
  std::vector<uint32_t> parent(nn, UINT32_MAX);
  for (size_t i = 0; i < nn; ++i)
    if (!is_leaf(i)) {
      parent[i + 1] = (uint32_t)i;
      parent[t.right[i]] = (uint32_t)i;
    }
  t.chain_ptr.push_back(0);
  for (size_t a = 0; a < nn; ++a)
    if (is_leaf(a)) {
      t.leaves.push_back((uint32_t)a);
      size_t s0 = t.chain.size();
      uint32_t x = (uint32_t)a;
      for (;;) {
        t.chain.push_back(x);
        if (x == 0) break;
        x = parent[x];
      }
      std::reverse(t.chain.begin() + s0, t.chain.end());
      t.chain_ptr.push_back(t.chain.size());
    }
  return t;
}

// acc[c] += k[map[c]] · x[c] over the entries, in order (block_tree.rs:712 `accumulate`).
static inline void accumulate(double* acc, const double* ks, const double* xs, size_t count, int N, size_t C,
                              const std::vector<int>& map) {
  for (size_t e = 0; e < count; ++e)
    for (size_t c = 0; c < C; ++c) acc[c] += ks[e * N + map[c]] * xs[e * C + c];
}

std::vector<double> block_tree_multiply_permuted(const BlockTree& t, const std::vector<int>& map,
                                                 const std::vector<double>& psi, int threads) {
  const size_t C = map.size(), N = (size_t)t.N;
  const size_t n = t.order.size(), nn = t.start.size();
  std::vector<double> up(nn * C, 0.0);
  for (size_t a = nn; a-- > 0;) {
    size_t r = t.right[a];
    double* u = &up[a * C];
    if (r == 0) {
      for (size_t c = 0; c < C; ++c) u[c] = 0.0;
      for (uint32_t p = t.start[a]; p < t.end[a]; ++p)
        for (size_t c = 0; c < C; ++c) u[c] += t.weight[p] * psi[p * C + c];
    } else {
      for (size_t c = 0; c < C; ++c) u[c] = up[(a + 1) * C + c] + up[r * C + c];
    }
  }
  std::vector<double> acc(nn * C, 0.0);
  parallel_for((int64_t)nn, threads, [&](int64_t a) {
    double* f = &acc[a * C];
    for (uint64_t p = t.far_ptr[a]; p < t.far_ptr[a + 1]; ++p)
      accumulate(f, &t.far_coef[p * N], &up[(size_t)t.far_col[p] * C], 1, (int)N, C, map);
  });
  for (size_t a = 0; a < nn; ++a) {
    size_t r = t.right[a];
    if (r != 0)
      for (size_t child : {a + 1, r})
        for (size_t c = 0; c < C; ++c) acc[child * C + c] += acc[a * C + c];
  }
  std::vector<double> out(n * C, 0.0);
  parallel_for((int64_t)t.leaves.size(), threads, [&](int64_t li) {
    size_t l = t.leaves[li];
    size_t s = t.start[l];
    size_t k = t.end[l] - t.start[l];
    for (size_t r = 0; r < k; ++r)
      for (size_t c = 0; c < C; ++c) out[(s + r) * C + c] = t.weight[s + r] * acc[l * C + c];
    for (uint64_t ci = t.chain_ptr[li]; ci < t.chain_ptr[li + 1]; ++ci) {
      size_t a = t.chain[ci];
      size_t ra = t.end[a] - t.start[a];
      size_t r0 = s - t.start[a];
      for (uint64_t q = t.near_ptr[a]; q < t.near_ptr[a + 1]; ++q) {
        size_t b = t.near_col[q];
        size_t sb = t.start[b], cb = t.end[b] - t.start[b];
        const double* blk = &t.near_val[t.near_off[q] * N];
        const double* xs = &psi[sb * C];
        (void)ra;
        for (size_t r = 0; r < k; ++r) {
          const double* row = blk + (r0 + r) * cb * N;
          double* o = &out[(s + r) * C];
          if (a == b) {
            size_t d = r0 + r;
            accumulate(o, row, xs, d, (int)N, C, map);
            accumulate(o, row + (d + 1) * N, xs + (d + 1) * C, cb - d - 1, (int)N, C, map);
          } else {
            accumulate(o, row, xs, cb, (int)N, C, map);
          }
        }
      }
    }
  });
  return out;
}

std::vector<double> block_tree_multiply(const BlockTree& t, const std::vector<int>& map, const std::vector<double>& psi,
                                        int threads) {
  const size_t C = map.size(), n = t.order.size();
  std::vector<double> pp(n * C);
  for (size_t p = 0; p < n; ++p)
    for (size_t c = 0; c < C; ++c) pp[p * C + c] = psi[(size_t)t.order[p] * C + c];
  auto out = block_tree_multiply_permuted(t, map, pp, threads);
  std::vector<double> res(n * C);
  for (size_t i = 0; i < n; ++i)
    for (size_t c = 0; c < C; ++c) res[i * C + c] = out[(size_t)t.position[i] * C + c];
  return res;
}

// hierarchical.rs:116 `separated`.
static inline bool separated(double ra, const Aabb& ba, double rb, const Aabb& bb, double chi) {
  if (!(chi > 0.0)) return false;
  double d2 = box_distance_sq(ba, bb);
  double h = fmax_(ra, rb);
  return d2 > 0.0 && h * h <= chi * chi * d2;
}

// hierarchical.rs:185 `hi_lo`.
static inline void hi_lo(V3 d, double qa, double qb, double neg_s, double* out) {
  double r2 = dot(d, d);
  double hi = std::pow(r2, neg_s);
  out[0] = hi;
  out[1] = 0.5 * (qa + qb) / (r2 * r2) * hi;
}

// metric/mod.rs:120 `df_coefficients`.
std::array<V3, 3> df_coefficients(const std::array<V3, 3>& p) {
  V3 n = cross(sub(p[1], p[0]), sub(p[2], p[0]));
  double len = norm(n);
  V3 w = scale(n, 1.0 / (len * len));
  auto e = opposite_edges(p);
  return {cross(w, e[0]), cross(w, e[1]), cross(w, e[2])};
}

bool face_metric_build(const Mesh& m, const Positions& p, const Exponents& e, double chi, size_t leaf_size,
                       size_t leaf_pairs, int threads, FaceMetric* out, Error* err) {
  if (!validate_positions(m, p, err)) return false;
  out->mesh = m;
  out->geom = face_geometry(m, p);
  out->tree = face_tree_build(m, p, out->geom, leaf_size);
  out->df.resize(m.faces.size());
  for (size_t f = 0; f < m.faces.size(); ++f) out->df[f] = df_coefficients(corners(p, m.faces[f]));
  const double neg_s = -e.s();
  const FaceTree& ft = out->tree;
  const FaceGeometry& g = out->geom;
  ClusterKernelFns K;
  K.admissible = [&](size_t a, size_t b) {
    return separated(ft.clusters[a].radius, ft.clusters[a].bbox, ft.clusters[b].radius, ft.clusters[b].bbox, chi);
  };
  K.far = [&](size_t a, size_t b, double* o) {
    if (a > b) std::swap(a, b);
    const Cluster &ca = ft.clusters[a], &cb = ft.clusters[b];
    V3 d = sub(ca.centroid, cb.centroid);
    hi_lo(d, ca.pq(d), cb.pq(d), neg_s, o);
  };
  K.near = [&](size_t i, size_t j, double* o) {
    size_t a = std::min(i, j), b = std::max(i, j);
    V3 d = sub(g.barycenters[a], g.barycenters[b]);
    double ns = dot(g.normals[a], d), nt = dot(g.normals[b], d);
    hi_lo(d, ns * ns, nt * nt, neg_s, o);
  };
  K.weight = [](size_t) { return 1.0; };
  out->bct = block_tree_build(ft.bvh, K, 2, leaf_pairs, true);
  const size_t nf = m.faces.size();
  std::vector<double> psi(2 * nf);
  for (size_t f = 0; f < nf; ++f) psi[2 * f] = psi[2 * f + 1] = g.areas[f];
  out->ha = block_tree_multiply(out->bct, {0, 1}, psi, threads);
  for (size_t f = 0; f < nf; ++f)
    if (!(std::isfinite(out->ha[2 * f]) && std::isfinite(out->ha[2 * f + 1]))) {
      err->status = NON_FINITE_ENERGY;
      err->a = (int64_t)f;
      err->detail = "coincident barycenters";
      return false;
    }
  out->ha_perm.resize(2 * nf);
  for (size_t q = 0; q < nf; ++q)
    for (int k = 0; k < 2; ++k) out->ha_perm[2 * q + k] = out->ha[2 * (size_t)out->bct.order[q] + k];
  // incidence, faces ascending
  out->inc_ptr.assign(m.nv + 1, 0);
  for (auto& t : m.faces)
    for (int c = 0; c < 3; ++c) out->inc_ptr[t[c] + 1]++;
  for (int64_t v = 0; v < m.nv; ++v) out->inc_ptr[v + 1] += out->inc_ptr[v];
  out->inc.resize(3 * nf);
  std::vector<uint64_t> fill(out->inc_ptr.begin(), out->inc_ptr.end() - 1);
  for (size_t s = 0; s < nf; ++s)
    for (int c = 0; c < 3; ++c) out->inc[fill[m.faces[s][c]]++] = {(uint32_t)s, (uint8_t)c};
  return true;
}

// hierarchical.rs:696 `product` for nb vectors (interleaved 3nv each).
void face_metric_apply(const FaceMetric& fm, double w_hi, double w_lo, size_t nb, const double* xs, double* ys,
                       int threads) {
  const bool hi = w_hi != 0.0, lo = w_lo != 0.0;
  const size_t nf = fm.mesh.faces.size(), nv = (size_t)fm.mesh.nv;
  if (!hi && !lo) {
    std::fill(ys, ys + nb * 3 * nv, 0.0);
    return;
  }
  for (size_t start = 0; start < nb; start += 4) {
    const size_t k = std::min<size_t>(4, nb - start);
    const double* xb = xs + start * 3 * nv;
    double* yb = ys + start * 3 * nv;
    const size_t C = k * (9 * (hi ? 1 : 0) + 3 * (lo ? 1 : 0));
    const size_t off = hi ? 9 * k : 0;
    std::vector<double> mm(nf * C, 0.0), psi(nf * C);
    const auto& order = fm.bct.order;
    for (size_t p = 0; p < nf; ++p) {
      size_t s = order[p];
      const auto& t = fm.mesh.faces[s];
      double* out = &mm[p * C];
      for (size_t q = 0; q < k; ++q) {
        const double* x = xb + q * 3 * nv;
        if (hi) {
          const auto& g = fm.df[s];
          double* o = out + 9 * q;
          for (int kk = 0; kk < 3; ++kk)
            for (int c = 0; c < 3; ++c) {
              double xc = x[3 * (size_t)t[c] + kk];
              for (int r = 0; r < 3; ++r) o[3 * kk + r] += xc * g[c][r];
            }
        }
        if (lo)
          for (int kk = 0; kk < 3; ++kk)
            out[off + 3 * q + kk] = (x[3 * (size_t)t[0] + kk] + x[3 * (size_t)t[1] + kk] + x[3 * (size_t)t[2] + kk]) / 3.0;
      }
      double a = fm.geom.areas[s];
      for (size_t c = 0; c < C; ++c) psi[p * C + c] = mm[p * C + c] * a;
    }
    std::vector<int> map(C);
    for (size_t c = 0; c < C; ++c) map[c] = c < off ? 0 : 1;
    auto hx = block_tree_multiply_permuted(fm.bct, map, psi, threads);
    std::vector<double> z(nf * C);
    for (size_t p = 0; p < nf; ++p)
      for (size_t c = 0; c < C; ++c) {
        double w = c < off ? w_hi : w_lo;
        int kern = c < off ? 0 : 1;
        double d = 0.0;
        z[p * C + c] = w * (fm.ha_perm[2 * p + kern] * mm[p * C + c] - hx[p * C + c]) + d * mm[p * C + c];
      }
    for (size_t q = 0; q < k; ++q) {
      double* y = yb + q * 3 * nv;
      parallel_for((int64_t)nv, threads, [&](int64_t v) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (uint64_t e = fm.inc_ptr[v]; e < fm.inc_ptr[v + 1]; ++e) {
          size_t s = fm.inc[e].first, c = fm.inc[e].second;
          double a = fm.geom.areas[s];
          const double* zs = &z[(size_t)fm.bct.position[s] * C];
          if (hi) {
            const V3& g = fm.df[s][c];
            const double* zh = zs + 9 * q;
            for (int kk = 0; kk < 3; ++kk) acc[kk] += a * (g[0] * zh[3 * kk] + g[1] * zh[3 * kk + 1] + g[2] * zh[3 * kk + 2]);
          }
          if (lo)
            for (int kk = 0; kk < 3; ++kk) acc[kk] += a / 3.0 * zs[off + 3 * q + kk];
        }
        for (int kk = 0; kk < 3; ++kk) y[3 * v + kk] = 2.0 * acc[kk];
      });
    }
  }
}

bool vertex_fractional_build(const Positions& p, const std::vector<double>& mass, double tau, double chi,
                             size_t leaf_size, size_t leaf_pairs, int threads, VertexFractional* out, Error* err) {
  const size_t n = p.size();
  std::vector<Aabb> bnds(n);
  for (size_t v = 0; v < n; ++v) {
    bnds[v].lo = p[v];
    bnds[v].hi = p[v];
  }
  out->bvh = bvh_build(p, bnds, mass, leaf_size);
  const size_t nn = out->bvh.nodes.size();
  std::vector<V3> cen(nn);
  std::vector<double> rad(nn);
  for (size_t i = 0; i < nn; ++i) {
    cen[i] = out->bvh.nodes[i].centroid();
    rad[i] = farthest_corner_distance(out->bvh.nodes[i].bbox, cen[i]);
  }
  const double ex = -(tau + 1.0);
  const Bvh& bv = out->bvh;
  ClusterKernelFns K;
  K.admissible = [&](size_t a, size_t b) { return separated(rad[a], bv.nodes[a].bbox, rad[b], bv.nodes[b].bbox, chi); };
  K.far = [&](size_t a, size_t b, double* o) {
    if (a > b) std::swap(a, b);
    V3 d = sub(cen[a], cen[b]);
    o[0] = std::pow(dot(d, d), ex);
  };
  K.near = [&](size_t i, size_t j, double* o) {
    size_t a = std::min(i, j), b = std::max(i, j);
    V3 d = sub(p[a], p[b]);
    o[0] = std::pow(dot(d, d), ex);
  };
  K.weight = [](size_t) { return 1.0; };
  out->bct = block_tree_build(bv, K, 1, leaf_pairs, true);
  out->mass = mass;
  out->ha = block_tree_multiply(out->bct, {0}, mass, threads);
  for (size_t v = 0; v < n; ++v)
    if (!std::isfinite(out->ha[v])) {
      err->status = INVALID_OPTION;
      err->detail = "coincident vertices";
      return false;
    }
  return true;
}

// hierarchical.rs:1313 `product`, vectors in groups of 4.
void vertex_fractional_apply(const VertexFractional& vf, size_t nb, const double* xs, double* ys, int threads) {
  const size_t n = vf.mass.size();
  for (size_t start = 0; start < nb; start += 4) {
    const size_t k = std::min<size_t>(4, nb - start), C = 3 * k;
    std::vector<double> psi(n * C);
    for (size_t p = 0; p < n; ++p) {
      size_t v = vf.bct.order[p];
      double a = vf.mass[v];
      for (size_t q = 0; q < k; ++q)
        for (int kk = 0; kk < 3; ++kk) psi[p * C + 3 * q + kk] = a * xs[(start + q) * 3 * n + 3 * v + kk];
    }
    std::vector<int> map(C, 0);
    auto hx = block_tree_multiply_permuted(vf.bct, map, psi, threads);
    for (size_t q = 0; q < k; ++q) {
      const double* x = xs + (start + q) * 3 * n;
      double* y = ys + (start + q) * 3 * n;
      for (size_t v = 0; v < n; ++v) {
        double a = vf.mass[v];
        const double* h = &hx[(size_t)vf.bct.position[v] * C];
        for (int kk = 0; kk < 3; ++kk) y[3 * v + kk] = 2.0 * a * (vf.ha[v] * x[3 * v + kk] - h[3 * q + kk]);
      }
    }
  }
}

}  // namespace orc
