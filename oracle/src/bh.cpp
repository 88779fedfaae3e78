// FP64 oracle — TEST INFRASTRUCTURE ONLY: BVH, face tree and Barnes–Hut energy / exact gradient.
//
// Restates:
//   crates/ddg-repulsive-curves/src/spatial/bvh.rs   build_with :126 (split loop :156-205),
//                                                    refit :221, fold_up :259, summarize :312
//   crates/ddg-repulsive-surfaces/src/spatial/mod.rs face_elements :55, FaceTree::build :167,
//                                                    refit_faces :245, projectors :275,
//                                                    bh_admissible :141
//   crates/ddg-repulsive-surfaces/src/energy/barnes_hut.rs
//                                                    build :156, refit :204, energy :366,
//                                                    face_partials :478, leaf_energy :583,
//                                                    leaf_with_gradient :627, build_lists :743,
//                                                    transpose :811, pull_back_faces :830
#include <algorithm>
#include <cstring>

#include "oracle.hpp"

namespace orc {

// f64::total_cmp as a signed integer key.
static inline int64_t total_key(double x) {
  int64_t b;
  std::memcpy(&b, &x, 8);
  b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
  return b;
}

Bvh bvh_build(const std::vector<V3>& pts, const std::vector<Aabb>& bnds, const std::vector<double>& mass,
              size_t leaf_size, double min_fraction) {
  const size_t n = pts.size();
  Bvh b;
  b.leaf_size = leaf_size;
  b.order.resize(n);
  for (size_t i = 0; i < n; ++i) b.order[i] = (uint32_t)i;
  struct Item {
    size_t start, end, depth;
    uint32_t patch;
  };
  std::vector<Item> stack{{0, n, 0, UINT32_MAX}};
  std::vector<Aabb> suffix;
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    uint32_t id = (uint32_t)b.nodes.size();
    if (it.patch != UINT32_MAX) b.nodes[it.patch].right = id;
    BvhNode nd;
    nd.start = (uint32_t)it.start;
    nd.end = (uint32_t)it.end;
    b.nodes.push_back(nd);
    size_t cnt = it.end - it.start;
    if (cnt <= leaf_size) continue;
    int axis = (int)(it.depth % 3);
    uint32_t* slice = b.order.data() + it.start;
    std::sort(slice, slice + cnt, [&](uint32_t a, uint32_t c) {
      int64_t ka = total_key(pts[a][axis]), kc = total_key(pts[c][axis]);
      if (ka != kc) return ka < kc;
      return a < c;
    });
    suffix.assign(cnt, Aabb());
    Aabb acc;
    for (size_t k = cnt; k-- > 0;) {
      acc = acc.unite(bnds[slice[k]]);
      suffix[k] = acc;
    }
    Aabb prefix;
    double best_cost = INFINITY;
    size_t best_k = cnt / 2;
    size_t lo = (size_t)std::ceil(min_fraction * (double)cnt);
    lo = std::min(std::max(lo, (size_t)1), cnt / 2);
    for (size_t k = 1; k <= cnt - lo; ++k) {
      prefix = prefix.unite(bnds[slice[k - 1]]);
      if (k < lo) continue;
      double cost = prefix.diameter_sq() + suffix[k].diameter_sq();
      if (cost < best_cost) {
        best_cost = cost;
        best_k = k;
      }
    }
    size_t split = it.start + best_k;
    stack.push_back({split, it.end, it.depth + 1, id});
    stack.push_back({it.start, split, it.depth + 1, UINT32_MAX});
  }
  bvh_refit(b, pts, bnds, mass);
  return b;
}

void bvh_refit(Bvh& b, const std::vector<V3>& pts, const std::vector<Aabb>& bnds,
               const std::vector<double>& mass) {
  for (size_t i = b.nodes.size(); i-- > 0;) {
    BvhNode& nd = b.nodes[i];
    if (nd.is_leaf()) {
      Aabb bb;
      double ms = 0.0;
      V3 mom{0.0, 0.0, 0.0};
      for (uint32_t q = nd.start; q < nd.end; ++q) {
        uint32_t e = b.order[q];
        double m = mass[e];
        bb = bb.unite(bnds[e]);
        ms += m;
        for (int k = 0; k < 3; ++k) mom[k] += m * pts[e][k];
      }
      nd.bbox = bb;
      nd.mass = ms;
      nd.moment = mom;
    } else {
      const BvhNode& l = b.nodes[i + 1];
      const BvhNode& r = b.nodes[nd.right];
      V3 mom = l.moment;
      for (int k = 0; k < 3; ++k) mom[k] += r.moment[k];
      nd.bbox = l.bbox.unite(r.bbox);
      nd.mass = l.mass + r.mass;
      nd.moment = mom;
    }
  }
}

std::vector<Aabb> face_bounds(const Mesh& m, const Positions& p) {
  std::vector<Aabb> out(m.faces.size());
  for (size_t f = 0; f < m.faces.size(); ++f)
    for (int c = 0; c < 3; ++c) out[f].grow(p[m.faces[f][c]]);
  return out;
}

// spatial/mod.rs:275 `projectors` through bvh.rs:259 `fold_up`.
static std::vector<std::array<double, 6>> projectors(const Bvh& b, const FaceGeometry& g) {
  const size_t nn = b.nodes.size();
  std::vector<std::array<double, 7>> s(nn);
  for (size_t i = nn; i-- > 0;) {
    const BvhNode& nd = b.nodes[i];
    std::array<double, 7> acc{};
    if (nd.is_leaf()) {
      for (uint32_t q = nd.start; q < nd.end; ++q) {
        uint32_t f = b.order[q];
        double a = g.areas[f];
        V3 n = g.normals[f];
        acc[0] += a;
        acc[1] += a * n[0] * n[0];
        acc[2] += a * n[0] * n[1];
        acc[3] += a * n[0] * n[2];
        acc[4] += a * n[1] * n[1];
        acc[5] += a * n[1] * n[2];
        acc[6] += a * n[2] * n[2];
      }
    } else {
      for (int k = 0; k < 7; ++k) acc[k] = s[i + 1][k] + s[nd.right][k];
    }
    s[i] = acc;
  }
  std::vector<std::array<double, 6>> out(nn);
  for (size_t i = 0; i < nn; ++i)
    for (int k = 0; k < 6; ++k) out[i][k] = s[i][k + 1] / s[i][0];
  return out;
}

FaceTree face_tree_build(const Mesh& m, const Positions& p, const FaceGeometry& g, size_t leaf_size) {
  FaceTree t;
  std::vector<Aabb> bnds = face_bounds(m, p);
  t.bvh = bvh_build(g.barycenters, bnds, g.areas, leaf_size);
  const size_t nn = t.bvh.nodes.size();
  t.leaf_ordinal.assign(nn, UINT32_MAX);
  t.leaf_of_face.assign(m.faces.size(), 0);
  for (size_t n = 0; n < nn; ++n)
    if (t.bvh.nodes[n].is_leaf()) {
      t.leaf_ordinal[n] = (uint32_t)t.leaves.size();
      t.leaves.push_back((uint32_t)n);
      for (uint32_t q = t.bvh.nodes[n].start; q < t.bvh.nodes[n].end; ++q) t.leaf_of_face[t.bvh.order[q]] = (uint32_t)n;
    }
  auto pr = projectors(t.bvh, g);
  t.clusters.resize(nn);
  for (size_t n = 0; n < nn; ++n) {
    const BvhNode& nd = t.bvh.nodes[n];
    Cluster& c = t.clusters[n];
    c.area = nd.mass;
    c.centroid = nd.centroid();
    c.projector = pr[n];
    c.bbox = nd.bbox;
    c.radius = farthest_corner_distance(nd.bbox, c.centroid);
  }
  return t;
}

void face_tree_refit(FaceTree& t, const Mesh& m, const Positions& p, const FaceGeometry& g) {
  std::vector<Aabb> bnds = face_bounds(m, p);
  bvh_refit(t.bvh, g.barycenters, bnds, g.areas);
  auto pr = projectors(t.bvh, g);
  for (size_t n = 0; n < t.clusters.size(); ++n) {
    t.clusters[n].area = t.bvh.nodes[n].mass;
    t.clusters[n].centroid = t.bvh.nodes[n].centroid();
    t.clusters[n].projector = pr[n];
  }
}

// spatial/mod.rs:141 `bh_admissible` (strict).
static inline bool bh_admissible(const Cluster& leaf, const Cluster& c, double theta) {
  double h = fmax_(leaf.radius, c.radius);
  return h * h < theta * theta * box_distance_sq(leaf.bbox, c.bbox);
}

// barnes_hut.rs:811 `transpose`.
static void transpose(const std::vector<uint32_t>& entries, size_t m,
                      const std::function<size_t(uint32_t)>& key, std::vector<uint64_t>* off_out,
                      std::vector<uint32_t>* idx_out) {
  std::vector<uint64_t> off(m + 1, 0);
  for (uint32_t x : entries) off[key(x) + 1] += 1;
  for (size_t i = 0; i < m; ++i) off[i + 1] += off[i];
  std::vector<uint64_t> fill = off;
  std::vector<uint32_t> idx(entries.size());
  for (size_t k = 0; k < entries.size(); ++k) idx[fill[key(entries[k])]++] = (uint32_t)k;
  *off_out = std::move(off);
  *idx_out = std::move(idx);
}

InteractionLists bh_build_lists(const FaceTree& t, double theta, int threads) {
  const size_t nl = t.leaves.size(), nn = t.bvh.nodes.size();
  std::vector<std::vector<uint32_t>> lumped(nl), exact(nl);
  parallel_for((int64_t)nl, threads, [&](int64_t k) {
    const Cluster& cl = t.clusters[t.leaves[k]];
    std::vector<uint32_t> stack{0};
    while (!stack.empty()) {
      uint32_t c = stack.back();
      stack.pop_back();
      if (bh_admissible(cl, t.clusters[c], theta))
        lumped[k].push_back(c);
      else if (t.bvh.nodes[c].is_leaf())
        exact[k].push_back(c);
      else {
        stack.push_back(t.bvh.nodes[c].right);
        stack.push_back(c + 1);
      }
    }
  });
  InteractionLists L;
  L.lumped_off.push_back(0);
  L.exact_off.push_back(0);
  uint64_t poff = 0;
  for (size_t k = 0; k < nl; ++k) {
    const BvhNode& lf = t.bvh.nodes[t.leaves[k]];
    uint64_t ns = lf.end - lf.start;
    for (uint32_t c : exact[k]) {
      uint64_t nc = t.bvh.nodes[c].end - t.bvh.nodes[c].start;
      L.partner_off.push_back(poff);
      poff += nc;
      L.pairs_exact += (c == t.leaves[k]) ? ns * (ns - 1) : ns * nc;
    }
    L.lumped.insert(L.lumped.end(), lumped[k].begin(), lumped[k].end());
    L.exact.insert(L.exact.end(), exact[k].begin(), exact[k].end());
    L.lumped_off.push_back(L.lumped.size());
    L.exact_off.push_back(L.exact.size());
  }
  L.partner_off.push_back(poff);
  transpose(L.lumped, nn, [](uint32_t n) { return (size_t)n; }, &L.lumped_t_off, &L.lumped_t);
  transpose(L.exact, nl, [&](uint32_t n) { return (size_t)t.leaf_ordinal[n]; }, &L.exact_t_off, &L.exact_t);
  return L;
}

bool bh_build(const Mesh& m, const Positions& p, double theta, size_t leaf_size, int threads, BhTree* out,
              Error* err) {
  if (!(std::isfinite(theta) && theta >= 0.0)) {
    err->status = INVALID_OPTION;
    err->detail = "theta: finite and non-negative";
    return false;
  }
  if (leaf_size == 0) {
    err->status = INVALID_OPTION;
    err->detail = "leaf_size: at least 1";
    return false;
  }
  if (!validate_positions(m, p, err)) return false;
  out->theta = theta;
  out->leaf_size = leaf_size;
  out->mesh = m;
  out->positions = p;
  out->geom = face_geometry(m, p);
  out->tree = face_tree_build(m, p, out->geom, leaf_size);
  out->lists = bh_build_lists(out->tree, theta, threads);
  return true;
}

bool bh_refit(BhTree& bh, const Positions& p, Error* err) {
  if ((int64_t)p.size() != bh.mesh.nv) {
    err->status = INVALID_OPTION;
    err->detail = "position count mismatch";
    return false;
  }
  for (size_t v = 0; v < p.size(); ++v)
    if (!(std::isfinite(p[v][0]) && std::isfinite(p[v][1]) && std::isfinite(p[v][2]))) {
      err->status = NON_FINITE_POSITION;
      err->a = (int64_t)v;
      return false;
    }
  bh.positions = p;
  bh.geom = face_geometry(bh.mesh, p);
  face_tree_refit(bh.tree, bh.mesh, p, bh.geom);
  return true;
}

namespace {
struct Packed {
  V3 x, n;
  double a;
};
std::vector<Packed> pack(const BhTree& bh) {
  std::vector<Packed> pk(bh.tree.bvh.order.size());
  for (size_t q = 0; q < pk.size(); ++q) {
    uint32_t f = bh.tree.bvh.order[q];
    pk[q] = {bh.geom.barycenters[f], bh.geom.normals[f], bh.geom.areas[f]};
  }
  return pk;
}
struct LeafOut {
  double energy = 0.0;
  std::vector<double> own_a;
  std::vector<V3> own_x, own_n;
  std::vector<double> node_b;
  std::vector<V3> node_u;
  std::vector<double> partner_a;
  std::vector<V3> partner_x;
};
}  // namespace

double bh_energy(const BhTree& bh, const Exponents& e, int threads) {
  const auto& t = bh.tree;
  const auto& L = bh.lists;
  auto pk = pack(bh);
  const size_t nl = t.leaves.size();
  std::vector<double> per_leaf(nl);
  parallel_for((int64_t)nl, threads, [&](int64_t k) {
    const BvhNode& lf = t.bvh.nodes[t.leaves[k]];
    double energy = 0.0;
    for (uint64_t q = L.lumped_off[k]; q < L.lumped_off[k + 1]; ++q) {
      const Cluster& cl = t.clusters[L.lumped[q]];
      double b = 0.0;
      for (uint32_t ps = lf.start; ps < lf.end; ++ps) {
        V3 v = sub(pk[ps].x, cl.centroid);
        b += pk[ps].a * kernel_of_v(e, v, pk[ps].n);
      }
      energy += b * cl.area;
    }
    for (uint64_t q = L.exact_off[k]; q < L.exact_off[k + 1]; ++q) {
      const BvhNode& cn = t.bvh.nodes[L.exact[q]];
      for (uint32_t ps = lf.start; ps < lf.end; ++ps) {
        double row = 0.0;
        for (uint32_t pt = cn.start; pt < cn.end; ++pt) {
          if (pt == ps) continue;
          row += pk[pt].a * kernel_of_v(e, sub(pk[ps].x, pk[pt].x), pk[ps].n);
        }
        energy += pk[ps].a * row;
      }
    }
    per_leaf[k] = energy;
  });
  double s = 0.0;
  for (double x : per_leaf) s += x;
  return s;
}

FacePartials bh_face_partials(const BhTree& bh, const Exponents& e, int threads) {
  const auto& t = bh.tree;
  const auto& L = bh.lists;
  auto pk = pack(bh);
  const size_t nl = t.leaves.size(), nn = t.clusters.size(), nf = bh.mesh.faces.size();
  std::vector<LeafOut> outs(nl);
  // 1. Per leaf: energy, own partials, node terms, partner blocks (barnes_hut.rs:627).
  parallel_for((int64_t)nl, threads, [&](int64_t k) {
    const BvhNode& lf = t.bvh.nodes[t.leaves[k]];
    const size_t ns = lf.end - lf.start;
    LeafOut& o = outs[k];
    o.own_a.assign(ns, 0.0);
    o.own_x.assign(ns, V3{0, 0, 0});
    o.own_n.assign(ns, V3{0, 0, 0});
    for (uint64_t q = L.lumped_off[k]; q < L.lumped_off[k + 1]; ++q) {
      const Cluster& cl = t.clusters[L.lumped[q]];
      double ac = cl.area, b = 0.0;
      V3 u{0, 0, 0};
      for (size_t i = 0; i < ns; ++i) {
        const Packed& fs = pk[lf.start + i];
        double a_s = fs.a;
        KernelEval kv = kernel_with_gradient_of_v(e, sub(fs.x, cl.centroid), fs.n);
        b += a_s * kv.value;
        o.own_a[i] += kv.value * ac;
        o.own_x[i] = axpy(o.own_x[i], a_s * ac, kv.d_v);
        o.own_n[i] = axpy(o.own_n[i], a_s * ac, kv.d_n);
        u = axpy(u, -a_s, kv.d_v);
      }
      o.energy += b * ac;
      o.node_b.push_back(b);
      o.node_u.push_back(u);
    }
    for (uint64_t q = L.exact_off[k]; q < L.exact_off[k + 1]; ++q) {
      const BvhNode& cn = t.bvh.nodes[L.exact[q]];
      size_t base = o.partner_a.size(), nc = cn.end - cn.start;
      o.partner_a.resize(base + nc, 0.0);
      o.partner_x.resize(base + nc, V3{0, 0, 0});
      for (size_t i = 0; i < ns; ++i) {
        uint32_t ps = lf.start + (uint32_t)i;
        const Packed& fs = pk[ps];
        double a_s = fs.a, row = 0.0;
        V3 dx{0, 0, 0}, dn{0, 0, 0};
        for (size_t j = 0; j < nc; ++j) {
          uint32_t pt = cn.start + (uint32_t)j;
          if (pt == ps) continue;
          const Packed& ft = pk[pt];
          double a_t = ft.a;
          KernelEval kv = kernel_with_gradient_of_v(e, sub(fs.x, ft.x), fs.n);
          row += a_t * kv.value;
          dx = axpy(dx, a_t, kv.d_v);
          dn = axpy(dn, a_t, kv.d_n);
          o.partner_a[base + j] += a_s * kv.value;
          o.partner_x[base + j] = axpy(o.partner_x[base + j], -a_s * a_t, kv.d_v);
        }
        o.energy += a_s * row;
        o.own_a[i] += row;
        o.own_x[i] = axpy(o.own_x[i], a_s, dx);
        o.own_n[i] = axpy(o.own_n[i], a_s, dn);
      }
    }
  });
  FacePartials fp;
  fp.energy = 0.0;
  for (auto& o : outs) fp.energy += o.energy;
  std::vector<double> nb;
  std::vector<V3> nu;
  std::vector<double> pa;
  std::vector<V3> px;
  for (auto& o : outs) {
    nb.insert(nb.end(), o.node_b.begin(), o.node_b.end());
    nu.insert(nu.end(), o.node_u.begin(), o.node_u.end());
    pa.insert(pa.end(), o.partner_a.begin(), o.partner_a.end());
    px.insert(px.end(), o.partner_x.begin(), o.partner_x.end());
  }
  // 2. Node sums over the transposed lumped lists.
  std::vector<double> sb(nn, 0.0);
  std::vector<V3> su(nn, V3{0, 0, 0});
  for (size_t n = 0; n < nn; ++n) {
    double b = 0.0;
    V3 u{0, 0, 0};
    for (uint64_t q = L.lumped_t_off[n]; q < L.lumped_t_off[n + 1]; ++q) {
      uint32_t kk = L.lumped_t[q];
      b += nb[kk];
      u = axpy(u, 1.0, nu[kk]);
    }
    sb[n] = b;
    su[n] = u;
  }
  // 3. Downward pass (pre-order).
  std::vector<double> beta(nn, 0.0);
  std::vector<V3> ubig(nn, V3{0, 0, 0});
  if (nn > 0) {
    beta[0] = sb[0];
    ubig[0] = su[0];
  }
  for (size_t n = 0; n < nn; ++n) {
    const BvhNode& nd = t.bvh.nodes[n];
    if (nd.is_leaf()) continue;
    for (size_t c : {n + 1, (size_t)nd.right}) {
      V3 shift = sub(t.clusters[c].centroid, t.clusters[n].centroid);
      beta[c] = beta[n] + dot(ubig[n], shift) + sb[c];
      ubig[c] = axpy(ubig[n], 1.0, su[c]);
    }
  }
  // 4. Per leaf: own + partner + cluster part.
  fp.d_area.assign(nf, 0.0);
  fp.d_bary.assign(nf, V3{0, 0, 0});
  fp.d_normal.assign(nf, V3{0, 0, 0});
  parallel_for((int64_t)nl, threads, [&](int64_t kc) {
    uint32_t c = t.leaves[kc];
    const BvhNode& lf = t.bvh.nodes[c];
    const LeafOut& o = outs[kc];
    size_t ns = lf.end - lf.start;
    std::vector<double> ra = o.own_a;
    std::vector<V3> rx = o.own_x;
    for (uint64_t q = L.exact_t_off[kc]; q < L.exact_t_off[kc + 1]; ++q) {
      uint64_t off = L.partner_off[L.exact_t[q]];
      for (size_t i = 0; i < ns; ++i) {
        ra[i] += pa[off + i];
        rx[i] = axpy(rx[i], 1.0, px[off + i]);
      }
    }
    V3 xc = t.clusters[c].centroid;
    for (size_t i = 0; i < ns; ++i) {
      uint32_t f = t.bvh.order[lf.start + i];
      double a = bh.geom.areas[f];
      V3 x = bh.geom.barycenters[f];
      ra[i] += beta[c] + dot(ubig[c], sub(x, xc));
      rx[i] = axpy(rx[i], a, ubig[c]);
      fp.d_area[f] = ra[i];
      fp.d_bary[f] = rx[i];
      fp.d_normal[f] = o.own_n[i];
    }
  });
  return fp;
}

bool bh_energy_and_differential(const BhTree& bh, const Exponents& e, int threads, double* energy,
                                std::vector<V3>* diff, Error* err) {
  FacePartials fp = bh_face_partials(bh, e, threads);
  // pull_back_faces (barnes_hut.rs:830): out = axpy(out, 1.0, g), the same roundings as add.
  std::vector<V3> d = pull_back(bh.mesh, bh.positions, fp.d_area, fp.d_bary, fp.d_normal);
  bool finite = std::isfinite(fp.energy);
  for (auto& x : d)
    for (int c = 0; c < 3; ++c) finite = finite && std::isfinite(x[c]);
  if (!finite) {
    err->status = NON_FINITE_ENERGY;
    err->detail = "non-finite Barnes-Hut energy";
    return false;
  }
  *energy = fp.energy;
  *diff = std::move(d);
  return true;
}

}  // namespace orc
