// Continuous collision detection: the oracle's restatement (FP64) of
//   crates/ddg-repulsive-surfaces/src/ccd.rs
//       CcdOptions::validate :126, BoxTree / split3 / build_tree :268-352 (Morton-sorted blocks
//       merged pairwise), query :359 (explicit stack, the right child popped first), lipschitz
//       :480, advance :529, first_contact_with_band :781 (windows [0, t1], [t1, 2 t1], ...),
//       window :886 (vertex against triangle, then edge against edge, per query the horizon of
//       its best contact so far), orient / segment_meets_triangle / triangles_intersect
//       :1153-1191, intersecting_pairs :1229, IntersectionBand::new :1292, min_distance :1364
//   crates/ddg-repulsive-curves/src/ccd.rs
//       point_segment_distance :236, segment_segment_distance :266, point_triangle_distance :298
//   crates/ddg-repulsive-curves/src/spatial/aabb.rs  Aabb::empty / grow / union (f64::min, max)
// Mesh against mesh only: the static obstacles of the reference (CcdObstacle) are not restated.
// Test infrastructure only.
//
// Parity pins: the reference's own tests (tests/ccd.rs: head-on vertex-face and edge-edge contacts
// at t = 0.95, sliding and stationary scenes without contact, initially close and touching pairs,
// option errors, sampled distances of random soups and sphere pairs staying above the gap before
// τ_max, the band of two overlapping spheres, coplanar neighbours), closed forms of the distances
// and brute force (min_distance against all pairs).
#include <algorithm>
#include <cmath>
#include <limits>

#include "oracle.hpp"

namespace orc {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kMargin = 1.0 + 1e-12;  // ccd.rs:498 MARGIN

struct Box {
  V3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
};

void grow(Box& b, V3 p) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = fmin_(b.lo[k], p[k]);
    b.hi[k] = fmax_(b.hi[k], p[k]);
  }
}

Box unite(const Box& a, const Box& b) {
  Box r = a;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = fmin_(r.lo[k], b.lo[k]);
    r.hi[k] = fmax_(r.hi[k], b.hi[k]);
  }
  return r;
}

Box inflate(const Box& b, double h) {
  Box r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = b.lo[k] - h;
    r.hi[k] = b.hi[k] + h;
  }
  return r;
}

bool overlaps(const Box& a, const Box& b) {
  for (int k = 0; k < 3; ++k)
    if (!(a.lo[k] <= b.hi[k] && b.lo[k] <= a.hi[k])) return false;
  return true;
}

uint64_t split3(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x001f00000000ffffull;
  x = (x | x << 16) & 0x001f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// Rust's saturating `f64 as u64`: NaN and negatives to 0.
uint64_t sat_u64(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 18446744073709551615.0) return ~0ull;
  return (uint64_t)x;
}

struct BoxTree {
  std::vector<uint32_t> order;
  std::vector<Box> sboxes;
  std::vector<std::vector<Box>> levels;
  size_t leaf = 8;
};

BoxTree build_tree(const std::vector<Box>& boxes, size_t leaf) {
  BoxTree t;
  t.leaf = leaf;
  const size_t n = boxes.size();
  std::vector<V3> c(n);
  V3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) c[i][k] = 0.5 * (boxes[i].lo[k] + boxes[i].hi[k]);
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin_(lo[k], c[i][k]);
      hi[k] = fmax_(hi[k], c[i][k]);
    }
  }
  auto q = [&](const V3& p, int k) -> uint64_t {
    const double w = hi[k] - lo[k];
    return w > 0.0 ? sat_u64(((p[k] - lo[k]) / w) * 2097151.0) : 0;
  };
  std::vector<std::pair<uint64_t, uint32_t>> keys(n);
  for (size_t i = 0; i < n; ++i)
    keys[i] = {split3(q(c[i], 0)) | split3(q(c[i], 1)) << 1 | split3(q(c[i], 2)) << 2, (uint32_t)i};
  std::sort(keys.begin(), keys.end());
  t.order.resize(n);
  t.sboxes.resize(n);
  for (size_t i = 0; i < n; ++i) {
    t.order[i] = keys[i].second;
    t.sboxes[i] = boxes[keys[i].second];
  }
  std::vector<Box> l0;
  for (size_t a = 0; a < n; a += leaf) {
    Box b;
    for (size_t p = a; p < std::min(n, a + leaf); ++p) b = unite(b, t.sboxes[p]);
    l0.push_back(b);
  }
  t.levels.push_back(std::move(l0));
  while (t.levels.back().size() > 1) {
    const auto& prev = t.levels.back();
    std::vector<Box> next;
    for (size_t a = 0; a < prev.size(); a += 2) {
      Box b;
      for (size_t p = a; p < std::min(prev.size(), a + 2); ++p) b = unite(b, prev[p]);
      next.push_back(b);
    }
    t.levels.push_back(std::move(next));
  }
  return t;
}

template <class F>
void query(const BoxTree& t, const Box& q, F&& f) {
  std::pair<uint32_t, uint32_t> stack[96];
  stack[0] = {(uint32_t)t.levels.size() - 1, 0};
  int sp = 1;
  while (sp > 0) {
    --sp;
    const auto [k, i] = stack[sp];
    if (!overlaps(t.levels[k][i], q)) continue;
    if (k == 0) {
      const size_t a = (size_t)i * t.leaf, b = std::min(a + t.leaf, t.order.size());
      for (size_t p = a; p < b; ++p)
        if (overlaps(t.sboxes[p], q)) f(t.order[p]);
    } else {
      const auto& below = t.levels[k - 1];
      stack[sp++] = {k - 1, 2 * i};
      if ((size_t)(2 * i + 1) < below.size()) stack[sp++] = {k - 1, 2 * i + 1};
    }
  }
}

V3 at(V3 p, V3 u, double t) { return {p[0] + t * u[0], p[1] + t * u[1], p[2] + t * u[2]}; }

double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

// ccd.rs:480
double lipschitz(const V3* ua, int na, const V3* ub, int nb) {
  const double n = (double)(na + nb);
  V3 mean{0.0, 0.0, 0.0};
  for (int i = 0; i < na; ++i) mean = add(mean, ua[i]);
  for (int i = 0; i < nb; ++i) mean = add(mean, ub[i]);
  mean = scale(mean, 1.0 / n);
  auto l_of = [&](V3 c) {
    double la = 0.0, lb = 0.0;
    for (int i = 0; i < na; ++i) la = fmax_(la, norm(sub(ua[i], c)));
    for (int i = 0; i < nb; ++i) lb = fmax_(lb, norm(sub(ub[i], c)));
    return la + lb;
  };
  return fmin_(l_of({0.0, 0.0, 0.0}), l_of(mean)) * (1.0 + 1e-12);
}

struct Tally {
  uint64_t iterations = 0, pairs = 0, close = 0, touching = 0, exempt = 0;
  void merge(const Tally& o) {
    iterations += o.iterations;
    pairs += o.pairs;
    close += o.close;
    touching += o.touching;
    exempt += o.exempt;
  }
};

// (τ, kind, a, b) with the reference's total order: τ, then the primitive (VertexTriangle before
// EdgeEdge, then the ids).
struct Found {
  bool has = false;
  CcdContact best;
  Tally tally;
  static bool before(const CcdContact& x, const CcdContact& y) {
    if (x.tau != y.tau) return x.tau < y.tau;
    if (x.kind != y.kind) return x.kind < y.kind;
    if (x.a != y.a) return x.a < y.a;
    return x.b <= y.b;
  }
  void offer(bool ok, double tau, int kind, uint32_t a, uint32_t b) {
    if (!ok) return;
    const CcdContact c{tau, kind, a, b};
    if (!has || !before(best, c)) best = c;
    has = true;
  }
  void merge(const Found& o) {
    if (o.has) offer(true, o.best.tau, o.best.kind, o.best.a, o.best.b);
    tally.merge(o.tally);
  }
  double horizon(double hi) const { return has ? fmin_(best.tau, hi) : hi; }
};

// ccd.rs:529
template <class D, class L>
bool advance(D&& dist, double l_cheap, L&& l_refined, double gap_abs, const CcdOptions& o, double lo, double hi,
             Tally& tally, double* out) {
  const double d0 = dist(0.0);
  tally.pairs += 1;
  tally.iterations += 1;
  if (lo == 0.0) {
    if (d0 <= 1e-12 * gap_abs) {
      tally.touching += 1;
      return false;
    }
    if (d0 < gap_abs) tally.close += 1;
  } else if (d0 <= 1e-12 * gap_abs) {
    return false;
  }
  const double target = d0 > gap_abs ? gap_abs : o.close_factor * d0;
  const double slack = 1e-9 * d0;
  double l = l_cheap;
  bool refined = false;
  double tau = lo;
  double d = lo == 0.0 ? d0 : dist(lo);
  if (lo != 0.0) tally.iterations += 1;
  for (uint32_t it = 0; it < o.max_iterations; ++it) {
    if (std::isnan(d) || d - target <= slack) {
      *out = tau;
      return true;
    }
    double next = tau + (d - target) / l;
    if (next >= hi && !refined) {
      l = fmin_(l, l_refined());
      refined = true;
      next = tau + (d - target) / l;
    }
    if (next >= hi || std::isnan(next)) return false;
    tau = next;
    d = dist(tau);
    tally.iterations += 1;
  }
  *out = tau;
  return true;
}

struct Motion {
  const Mesh* m;
  const Positions* start;
  const Positions* vel;
  std::vector<double> speed;
  double gap_abs;
  const CcdOptions* o;
  const std::vector<uint8_t>* band;  // nullptr: none
  V3 pos(uint32_t v, double t) const { return at((*start)[v], (*vel)[v], t); }
  bool in_band(const uint32_t* vs, int n) const {
    if (!band) return false;
    for (int i = 0; i < n; ++i)
      if ((*band)[vs[i]]) return true;
    return false;
  }
};

Box hull2(V3 a, V3 b) {
  Box r;
  grow(r, a);
  grow(r, b);
  return r;
}

// ccd.rs:886
Found window(const Motion& M, const std::vector<std::array<uint32_t, 2>>& edges, double lo, double hi) {
  const Mesh& m = *M.m;
  const size_t nv = (size_t)m.nv, nf = m.faces.size(), ne = edges.size();
  const double h = 0.5 * M.gap_abs * (1.0 + 1e-6);
  const size_t leaf = M.o->leaf_size;
  std::vector<Box> vbox(nv), fbox(nf), ebox(ne);
  for (size_t v = 0; v < nv; ++v) vbox[v] = hull2(M.pos((uint32_t)v, lo), M.pos((uint32_t)v, hi));
  for (size_t f = 0; f < nf; ++f)
    fbox[f] = unite(unite(vbox[m.faces[f][0]], vbox[m.faces[f][1]]), vbox[m.faces[f][2]]);
  for (size_t e = 0; e < ne; ++e) ebox[e] = unite(vbox[edges[e][0]], vbox[edges[e][1]]);
  std::vector<Box> fi(nf), ei(ne);
  for (size_t f = 0; f < nf; ++f) fi[f] = inflate(fbox[f], h);
  for (size_t e = 0; e < ne; ++e) ei[e] = inflate(ebox[e], h);
  const BoxTree ft = build_tree(fi, leaf), et = build_tree(ei, leaf);
  const Positions& S = *M.start;
  const Positions& U = *M.vel;
  Found total;
  for (uint32_t v = 0; v < nv; ++v) {
    Found fnd;
    const Box q = inflate(vbox[v], h);
    query(ft, q, [&](uint32_t f) {
      const auto& tri = m.faces[f];
      if (tri[0] == v || tri[1] == v || tri[2] == v) return;
      if (M.in_band(&v, 1) && M.in_band(tri.data(), 3)) {
        fnd.tally.exempt += 1;
        return;
      }
      const V3 p = S[v], u = U[v];
      const V3 tp[3] = {at(S[tri[0]], U[tri[0]], 0.0), at(S[tri[1]], U[tri[1]], 0.0), at(S[tri[2]], U[tri[2]], 0.0)};
      const V3 tu[3] = {U[tri[0]], U[tri[1]], U[tri[2]]};
      const double ts = fmax_(fmax_(M.speed[tri[0]], M.speed[tri[1]]), M.speed[tri[2]]);
      const double lc = (M.speed[v] + ts) * kMargin;
      auto dist = [&](double t) {
        return point_triangle_distance(at(p, u, t), at(tp[0], tu[0], t), at(tp[1], tu[1], t), at(tp[2], tu[2], t));
      };
      const double hz = fnd.horizon(hi);
      double tau = 0.0;
      const bool ok = advance(dist, lc, [&] { return lipschitz(&u, 1, tu, 3); }, M.gap_abs, *M.o, lo, hz, fnd.tally, &tau);
      fnd.offer(ok, tau, 0, v, f);
    });
    total.merge(fnd);
  }
  for (uint32_t e = 0; e < ne; ++e) {
    Found fnd;
    const uint32_t a = edges[e][0], b = edges[e][1];
    const Box q = inflate(ebox[e], h);
    query(et, q, [&](uint32_t e2) {
      if (e2 <= e) return;
      const uint32_t c = edges[e2][0], d = edges[e2][1];
      if (c == a || c == b || d == a || d == b) return;
      const uint32_t ab[2] = {a, b}, cd[2] = {c, d};
      if (M.in_band(ab, 2) && M.in_band(cd, 2)) {
        fnd.tally.exempt += 1;
        return;
      }
      const V3 pp[2] = {S[a], S[b]}, uu[2] = {U[a], U[b]}, qq[2] = {S[c], S[d]}, ww[2] = {U[c], U[d]};
      const double lc = (fmax_(M.speed[a], M.speed[b]) + fmax_(M.speed[c], M.speed[d])) * kMargin;
      auto dist = [&](double t) {
        return segment_segment_distance(at(pp[0], uu[0], t), at(pp[1], uu[1], t), at(qq[0], ww[0], t),
                                        at(qq[1], ww[1], t));
      };
      const double hz = fnd.horizon(hi);
      double tau = 0.0;
      const bool ok = advance(dist, lc, [&] { return lipschitz(uu, 2, ww, 2); }, M.gap_abs, *M.o, lo, hz, fnd.tally, &tau);
      fnd.offer(ok, tau, 1, e, e2);
    });
    total.merge(fnd);
  }
  return total;
}

Box face_box(const Mesh& m, const Positions& p, size_t f) {
  Box b;
  for (int c = 0; c < 3; ++c) grow(b, p[m.faces[f][c]]);
  return b;
}

double orient(V3 a, V3 b, V3 c, V3 d) { return dot(cross(sub(b, a), sub(c, a)), sub(d, a)); }

}  // namespace

// ---- distances (curves crate) ----------------------------------------------------------------

double point_segment_distance(V3 p, V3 a, V3 b) {
  const V3 ab = sub(b, a), ap = sub(p, a);
  const double e = dot(ab, ab);
  const double t = e > 0.0 ? clamp01(dot(ap, ab) / e) : 0.0;
  return norm(sub(ap, scale(ab, t)));
}

double segment_segment_distance(V3 p1, V3 q1, V3 p2, V3 q2) {
  double best = fmin_(fmin_(fmin_(point_segment_distance(p1, p2, q2), point_segment_distance(q1, p2, q2)),
                            point_segment_distance(p2, p1, q1)),
                      point_segment_distance(q2, p1, q1));
  const V3 d1 = sub(q1, p1), d2 = sub(q2, p2);
  const double a = dot(d1, d1), e = dot(d2, d2);
  const V3 n = cross(d1, d2);
  const double nn = dot(n, n);
  if (a > 0.0 && e > 0.0 && nn > 0.0) {
    const V3 r = sub(p1, p2);
    const double b = dot(d1, d2), c = dot(d1, r), f = dot(d2, r);
    double s = clamp01(dot(cross(sub(p2, p1), d2), n) / nn);
    const double t = clamp01((b * s + f) / e);
    s = clamp01((b * t - c) / a);
    const V3 c1 = add(p1, scale(d1, s)), c2 = add(p2, scale(d2, t));
    best = fmin_(best, norm(sub(c1, c2)));
  }
  return best;
}

double point_triangle_distance(V3 p, V3 a, V3 b, V3 c) {
  const double edges =
      fmin_(fmin_(point_segment_distance(p, a, b), point_segment_distance(p, b, c)), point_segment_distance(p, c, a));
  const V3 ab = sub(b, a), ac = sub(c, a);
  const V3 n = cross(ab, ac);
  const double nn = dot(n, n);
  if (nn <= 0.0) return edges;
  const V3 ap = sub(p, a);
  const double v = dot(cross(ap, ac), n) / nn;
  const double w = dot(cross(ab, ap), n) / nn;
  if (v >= 0.0 && w >= 0.0 && v + w <= 1.0) {
    const double h = std::fabs(dot(ap, n)) / std::sqrt(nn);
    return fmin_(edges, h);
  }
  return edges;
}

// ---- static queries ------------------------------------------------------------------------

bool segment_meets_triangle(V3 p, V3 q, const std::array<V3, 3>& tri) {
  const V3 a = tri[0], b = tri[1], c = tri[2];
  const double d1 = orient(a, b, c, p), d2 = orient(a, b, c, q);
  if ((d1 > 0.0 && d2 > 0.0) || (d1 < 0.0 && d2 < 0.0)) return false;
  const double n = norm(cross(sub(b, a), sub(c, a)));
  const double len = fmax_(fmax_(fmax_(norm(sub(q, p)), norm(sub(b, a))), norm(sub(c, a))), norm(sub(c, b)));
  const double flat = 1e-12 * n * len;
  if (std::fabs(d1) <= flat && std::fabs(d2) <= flat) {
    const double eps = 1e-14 * (norm(sub(b, a)) + norm(sub(c, a)));
    return point_triangle_distance(p, a, b, c) <= eps || point_triangle_distance(q, a, b, c) <= eps ||
           segment_segment_distance(p, q, a, b) <= eps || segment_segment_distance(p, q, b, c) <= eps ||
           segment_segment_distance(p, q, c, a) <= eps;
  }
  const double s1 = orient(p, q, a, b), s2 = orient(p, q, b, c), s3 = orient(p, q, c, a);
  return (s1 >= 0.0 && s2 >= 0.0 && s3 >= 0.0) || (s1 <= 0.0 && s2 <= 0.0 && s3 <= 0.0);
}

bool triangles_intersect(const std::array<V3, 3>& t1, const std::array<V3, 3>& t2) {
  for (int k = 0; k < 3; ++k)
    if (segment_meets_triangle(t1[k], t1[(k + 1) % 3], t2)) return true;
  for (int k = 0; k < 3; ++k)
    if (segment_meets_triangle(t2[k], t2[(k + 1) % 3], t1)) return true;
  return false;
}

std::vector<std::array<uint32_t, 2>> mesh_edges(const Mesh& m) {
  std::vector<std::array<uint32_t, 2>> e;
  for (auto& t : m.faces)
    for (int c = 0; c < 3; ++c) {
      const uint32_t a = t[c], b = t[(c + 1) % 3];
      e.push_back({std::min(a, b), std::max(a, b)});
    }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  return e;
}

std::vector<std::array<uint32_t, 2>> intersecting_pairs(const Mesh& m, const Positions& p) {
  const size_t nf = m.faces.size();
  std::vector<Box> boxes(nf);
  for (size_t f = 0; f < nf; ++f) boxes[f] = face_box(m, p, f);
  const BoxTree t = build_tree(boxes, 8);
  auto tri = [&](uint32_t f) {
    return std::array<V3, 3>{p[m.faces[f][0]], p[m.faces[f][1]], p[m.faces[f][2]]};
  };
  std::vector<std::array<uint32_t, 2>> out;
  for (uint32_t f = 0; f < nf; ++f) {
    query(t, boxes[f], [&](uint32_t g) {
      if (g <= f) return;
      const auto &a = m.faces[f], &b = m.faces[g];
      int common[3], nc = 0;
      for (int i = 0; i < 3; ++i)
        if (b[0] == a[i] || b[1] == a[i] || b[2] == a[i]) common[nc++] = i;
      bool hit = false;
      if (nc == 0) {
        hit = triangles_intersect(tri(f), tri(g));
      } else if (nc == 1) {
        const int i = common[0];
        const V3 pp = p[a[(i + 1) % 3]], qq = p[a[(i + 2) % 3]];
        int j = 0;
        while (b[j] != a[i]) ++j;
        const V3 r = p[b[(j + 1) % 3]], s = p[b[(j + 2) % 3]];
        hit = segment_meets_triangle(pp, qq, tri(g)) || segment_meets_triangle(r, s, tri(f));
      }
      if (hit) out.push_back({f, g});
    });
  }
  std::sort(out.begin(), out.end());
  return out;
}

std::vector<uint8_t> intersection_band(const Mesh& m, const Positions& p, size_t* n_pairs, size_t* n_vertices) {
  const auto pairs = intersecting_pairs(m, p);
  const size_t nv = (size_t)m.nv;
  std::vector<uint8_t> core(nv, 0);
  for (auto& fg : pairs)
    for (uint32_t f : fg)
      for (uint32_t v : m.faces[f]) core[v] = 1;
  std::vector<uint8_t> band = core;
  if (!pairs.empty())
    for (auto& e : mesh_edges(m))
      if (core[e[0]] || core[e[1]]) band[e[0]] = band[e[1]] = 1;
  *n_pairs = pairs.size();
  *n_vertices = 0;
  for (uint8_t x : band) *n_vertices += x;
  return band;
}

double min_distance(const Mesh& m, const Positions& p) {
  const size_t nf = m.faces.size();
  const auto edges = mesh_edges(m);
  std::vector<Box> fb(nf), eb(edges.size());
  for (size_t f = 0; f < nf; ++f) fb[f] = face_box(m, p, f);
  for (size_t e = 0; e < edges.size(); ++e) eb[e] = hull2(p[edges[e][0]], p[edges[e][1]]);
  const BoxTree ft = build_tree(fb, 8), et = build_tree(eb, 8);
  double r = mean_edge_length(m, p);
  for (;;) {
    double best = kInf;
    for (uint32_t v = 0; v < (uint32_t)m.nv; ++v) {
      Box q;
      grow(q, p[v]);
      query(ft, inflate(q, r), [&](uint32_t f) {
        const auto& t = m.faces[f];
        if (t[0] != v && t[1] != v && t[2] != v)
          best = fmin_(best, point_triangle_distance(p[v], p[t[0]], p[t[1]], p[t[2]]));
      });
    }
    for (uint32_t e = 0; e < edges.size(); ++e) {
      const uint32_t a = edges[e][0], b = edges[e][1];
      query(et, inflate(eb[e], r), [&](uint32_t e2) {
        if (e2 <= e) return;
        const uint32_t c = edges[e2][0], d = edges[e2][1];
        if (c == a || c == b || d == a || d == b) return;
        best = fmin_(best, segment_segment_distance(p[a], p[b], p[c], p[d]));
      });
    }
    if (best <= r) return best;
    Box all;
    for (auto& b : fb) all = unite(all, b);
    const double diam = norm(sub(all.hi, all.lo));
    if (r > 2.0 * diam) return best;
    r *= 2.0;
  }
}

bool ccd_validate(const CcdOptions& o, Error* err) {
  auto bad = [&](const char* what) {
    err->status = INVALID_OPTION;
    err->detail = what;
    return false;
  };
  if (!(std::isfinite(o.gap) && o.gap > 0.0)) return bad("ccd.gap: finite and positive");
  if (o.reference_length == o.reference_length && !(std::isfinite(o.reference_length) && o.reference_length > 0.0))
    return bad("ccd.reference_length: finite and positive");
  if (o.max_iterations == 0) return bad("ccd.max_iterations: at least 1");
  if (!(o.close_factor >= 0.0 && o.close_factor < 1.0)) return bad("ccd.close_factor: in [0, 1)");
  if (o.leaf_size == 0) return bad("ccd.leaf_size: at least 1");
  return true;
}

bool first_contact(const Mesh& m, const Positions& start, const Positions& vel, double tau_end, const CcdOptions& o,
                   const std::vector<uint8_t>* band, CcdReport* rep, Error* err) {
  if (!ccd_validate(o, err)) return false;
  const size_t nv = (size_t)m.nv;
  for (auto& x : start)
    if (!(std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2]))) {
      err->status = NON_FINITE_POSITION;
      err->detail = "start";
      return false;
    }
  if (vel.size() != nv) {
    err->status = INVALID_OPTION;
    err->detail = "velocity: one vector per vertex";
    return false;
  }
  for (auto& u : vel)
    if (!(std::isfinite(u[0]) && std::isfinite(u[1]) && std::isfinite(u[2]))) {
      err->status = INVALID_OPTION;
      err->detail = "velocity: finite components";
      return false;
    }
  if (!(std::isfinite(tau_end) && tau_end >= 0.0)) {
    err->status = INVALID_OPTION;
    err->detail = "tau_end: finite and non-negative";
    return false;
  }
  bool any = false;
  if (band) {
    if (band->size() != nv) {
      err->status = INVALID_OPTION;
      err->detail = "band: one flag per vertex";
      return false;
    }
    for (uint8_t x : *band) any = any || x;
  }
  const auto edges = mesh_edges(m);
  const double mel = mean_edge_length(m, start);
  const double ell = o.reference_length == o.reference_length ? o.reference_length : mel;
  Motion M{&m, &start, &vel, std::vector<double>(nv), o.gap * ell, &o, any ? band : nullptr};
  double vmax = 0.0;
  for (size_t v = 0; v < nv; ++v) {
    M.speed[v] = norm(vel[v]);
    vmax = fmax_(vmax, M.speed[v]);
  }
  *rep = CcdReport();
  rep->tau_max = tau_end;
  rep->gap_abs = M.gap_abs;
  const double t1 = vmax > 0.0 ? mel / vmax : kInf;
  double lo = 0.0, hi = fmin_(t1, tau_end);
  for (;;) {
    if (vmax == 0.0) hi = 0.0;
    rep->windows += 1;
    const Found f = window(M, edges, lo, hi);
    rep->iterations += f.tally.iterations;
    rep->pairs_tested += f.tally.pairs;
    if (lo == 0.0) {
      rep->close_pairs = f.tally.close;
      rep->touching_pairs = f.tally.touching;
      rep->exempt_pairs = f.tally.exempt;
    }
    if (f.has) {
      rep->tau_max = f.best.tau;
      rep->has_contact = true;
      rep->contact = f.best;
      break;
    }
    if (hi >= tau_end || vmax == 0.0) break;
    lo = hi;
    hi = fmin_(fmax_(2.0 * hi, hi + t1), tau_end);
    if (hi <= lo) break;
  }
  return true;
}

}  // namespace orc
