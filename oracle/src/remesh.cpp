// FP64 oracle — TEST INFRASTRUCTURE ONLY: one remeshing pass, restated from
//   crates/ddg-repulsive-surfaces/src/remesh/mod.rs       (remesh :235, collapse_short :338,
//                                                           delaunay_flips :380, smooth :424)
//   crates/ddg-repulsive-surfaces/src/remesh/mesh.rs      (EditMesh :224, edge_faces :381,
//                                                           edges_with_counts :404, fan :475,
//                                                           vertex_ok :558, split :592,
//                                                           collapse_target :652, collapse :679,
//                                                           flip :888, finish :991, combine :1038)
//   crates/ddg-repulsive-surfaces/src/remesh/geom.rs      (folds :26, violates_delaunay :50,
//                                                           circumcenter :64, smoothing_displacement :103)
//   crates/ddg-repulsive-surfaces/src/ccd.rs              (CcdCollapseGuard :1448-1633,
//                                                           triangle_distance :1195)
// Operation for operation; compiled with -ffp-contract=off like the rest of the oracle.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <queue>
#include <set>
#include <unordered_set>

#include "oracle.hpp"

namespace orc {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
// geom.rs:32 `COS_FOLDOVER_LIMIT` = cos(π − 2) = −cos 2
constexpr double kCosFoldover = 0.4161468365471424;

inline double norm2(V3 v) { return dot(v, v); }

}  // namespace

// ---- geom.rs -------------------------------------------------------------------------------

// geom.rs:11 `angle_between`
double remesh_angle_between(V3 u, V3 v) { return std::atan2(norm(cross(u, v)), dot(u, v)); }
// geom.rs:17 `tri_normal`
V3 remesh_tri_normal(const std::array<V3, 3>& p) { return cross(sub(p[1], p[0]), sub(p[2], p[0])); }
// geom.rs:26 `folds`
bool remesh_folds(V3 before, V3 after) {
  const double a = norm(after);
  return !(a > 0.0 && std::isfinite(a)) || dot(before, after) < kCosFoldover * norm(before) * a;
}
// geom.rs:36 `corner_angle`
double remesh_corner_angle(V3 v, V3 w, V3 x) { return remesh_angle_between(sub(w, v), sub(x, v)); }
// geom.rs:50 `violates_delaunay`
bool remesh_violates_delaunay(V3 p, V3 q, V3 c, V3 d) {
  const V3 uc = sub(p, c), vc = sub(q, c), ud = sub(p, d), vd = sub(q, d);
  const double dc = dot(uc, vc), sc = norm(cross(uc, vc));
  const double dd = dot(ud, vd), sd = norm(cross(ud, vd));
  return sc > 0.0 && sd > 0.0 && dc * sd + dd * sc < 0.0;
}
// geom.rs:64 `circumcenter`
V3 remesh_circumcenter(const std::array<V3, 3>& p) {
  const double a2 = norm2(sub(p[2], p[1]));
  const double b2 = norm2(sub(p[2], p[0]));
  const double c2 = norm2(sub(p[1], p[0]));
  const double w[3] = {a2 * (b2 + c2 - a2), b2 * (c2 + a2 - b2), c2 * (a2 + b2 - c2)};
  const double s = w[0] + w[1] + w[2];
  return add(add(scale(p[0], w[0] / s), scale(p[1], w[1] / s)), scale(p[2], w[2] / s));
}
// geom.rs:82 `barycenter`
V3 remesh_barycenter(const std::array<V3, 3>& p) { return scale(add(add(p[0], p[1]), p[2]), 1.0 / 3.0); }
// geom.rs:103 `smoothing_displacement` (false: total area not positive)
bool remesh_smoothing_displacement(V3 x, double rho, const std::vector<std::pair<double, V3>>& faces, V3* u) {
  V3 sum = {0.0, 0.0, 0.0};
  double area = 0.0;
  for (auto& [a, c] : faces) {
    sum = add(sum, scale(sub(c, x), a));
    area += a;
  }
  if (area > 0.0 && std::isfinite(area)) {
    *u = scale(sum, rho / area);
    return true;
  }
  return false;
}

// ccd.rs:1195 `triangle_distance`
double remesh_triangle_distance(const std::array<V3, 3>& t1, const std::array<V3, 3>& t2) {
  if (triangles_intersect(t1, t2)) return 0.0;
  double d = INFINITY;
  for (int k = 0; k < 3; ++k) {
    d = fmin_(d, point_triangle_distance(t1[k], t2[0], t2[1], t2[2]));
    d = fmin_(d, point_triangle_distance(t2[k], t1[0], t1[1], t1[2]));
    for (int j = 0; j < 3; ++j) d = fmin_(d, segment_segment_distance(t1[k], t1[(k + 1) % 3], t2[j], t2[(j + 1) % 3]));
  }
  return d;
}

namespace {

// mesh.rs:1051 `degenerate`: zero or non-finite area or normal (as validate_positions)
bool degenerate(const std::array<V3, 3>& p) {
  double area;
  V3 n, x;
  face_quantities(p, &area, &n, &x);
  return !(area > 0.0 && std::isfinite(area) && std::isfinite(n[0]) && std::isfinite(n[1]) && std::isfinite(n[2]));
}

using Tri = std::array<uint32_t, 3>;
using Src = std::vector<std::pair<uint32_t, double>>;

// mesh.rs:243 `orient`: the side p → q of f on {a, b} and the opposite vertex c
bool orient(const Tri& f, uint32_t a, uint32_t b, uint32_t* p, uint32_t* q, uint32_t* c) {
  for (int k = 0; k < 3; ++k) {
    const uint32_t pp = f[k], qq = f[(k + 1) % 3], cc = f[(k + 2) % 3];
    if ((pp == a && qq == b) || (pp == b && qq == a)) {
      *p = pp;
      *q = qq;
      *c = cc;
      return true;
    }
  }
  return false;
}
// mesh.rs:254 `traverses`
bool traverses(const Tri& f, uint32_t a, uint32_t b) {
  for (int k = 0; k < 3; ++k)
    if (f[k] == a && f[(k + 1) % 3] == b) return true;
  return false;
}
Tri sorted3(Tri f) {
  std::sort(f.begin(), f.end());
  return f;
}

// mesh.rs:1038 `combine`
Src combine(const Src& sa, double wa, const Src& sb, double wb) {
  Src r;
  for (auto& [o, w] : sa) r.push_back({o, w * wa});
  for (auto& [o, w] : sb) {
    bool found = false;
    for (auto& e : r)
      if (e.first == o) {
        e.second += w * wb;
        found = true;
        break;
      }
    if (!found) r.push_back({o, w * wb});
  }
  return r;
}

struct Fan {
  std::vector<uint32_t> faces;
  std::vector<bool> flip;
  bool boundary = false;
};

}  // namespace

bool guard_allow(CollapseGuard* g, const EditMesh& m, uint32_t ea, uint32_t eb, uint32_t survivor, V3 position,
                 const std::vector<std::array<uint32_t, 3>>& new_star);

// mesh.rs:224 `EditMesh`
struct EditMesh {
  Positions pos;
  std::vector<Tri> faces;
  std::vector<uint8_t> face_alive;
  std::vector<uint32_t> parent;
  std::vector<std::vector<uint32_t>> vfaces;
  std::vector<uint8_t> v_alive, pinned;
  std::vector<uint32_t> pin_origin;
  std::vector<Src> sources;
  std::vector<uint32_t> pins_in;
  size_t n_vertices_in = 0;

  bool valid_vertex(uint32_t v) const { return (size_t)v < pos.size() && v_alive[v]; }
  std::array<V3, 3> corners(uint32_t f) const { return {pos[faces[f][0]], pos[faces[f][1]], pos[faces[f][2]]}; }
  bool contains(const Tri& t, uint32_t v) const { return t[0] == v || t[1] == v || t[2] == v; }

  // mesh.rs:381 `edge_faces`: live faces on {a, b}, ascending
  std::vector<uint32_t> edge_faces(uint32_t a, uint32_t b) const {
    std::vector<uint32_t> r;
    if (!valid_vertex(a) || !valid_vertex(b) || a == b) return r;
    for (uint32_t f : vfaces[a])
      if (contains(faces[f], b)) r.push_back(f);
    std::sort(r.begin(), r.end());
    return r;
  }
  // mesh.rs:404 `edges_with_counts`
  std::vector<std::pair<std::array<uint32_t, 2>, uint32_t>> edges_with_counts() const {
    std::vector<std::pair<std::array<uint32_t, 2>, uint32_t>> out;
    std::vector<uint32_t> buf;
    for (uint32_t v = 0; v < (uint32_t)pos.size(); ++v) {
      if (!v_alive[v]) continue;
      buf.clear();
      for (uint32_t f : vfaces[v])
        for (uint32_t w : faces[f])
          if (w > v) buf.push_back(w);
      std::sort(buf.begin(), buf.end());
      size_t i = 0;
      while (i < buf.size()) {
        size_t j = i + 1;
        while (j < buf.size() && buf[j] == buf[i]) ++j;
        out.push_back({{v, buf[i]}, (uint32_t)(j - i)});
        i = j;
      }
    }
    return out;
  }
  std::vector<std::array<uint32_t, 2>> edges() const {
    std::vector<std::array<uint32_t, 2>> out;
    for (auto& e : edges_with_counts()) out.push_back(e.first);
    return out;
  }
  // mesh.rs:434 `neighbors`, ascending
  std::vector<uint32_t> neighbors(uint32_t v) const {
    std::vector<uint32_t> r;
    for (uint32_t f : vfaces[v])
      for (uint32_t w : faces[f])
        if (w != v) r.push_back(w);
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
  }
  // mesh.rs:449 `is_boundary_vertex`
  bool is_boundary_vertex(uint32_t v) const {
    std::vector<uint32_t> nb;
    for (uint32_t f : vfaces[v])
      for (uint32_t w : faces[f])
        if (w != v) nb.push_back(w);
    std::sort(nb.begin(), nb.end());
    size_t i = 0;
    while (i < nb.size()) {
      size_t j = i;
      while (j < nb.size() && nb[j] == nb[i]) ++j;
      if (j - i == 1) return true;
      i = j;
    }
    return false;
  }
  // mesh.rs:475 `fan`
  bool fan(uint32_t v, Fan* out) const {
    const auto& fs = vfaces[v];
    if (fs.empty()) return false;
    auto wedge = [&](uint32_t f) -> std::array<uint32_t, 2> {
      const Tri& t = faces[f];
      int k = 0;
      for (int i = 0; i < 3; ++i)
        if (t[i] == v) {
          k = i;
          break;
        }
      return {t[(k + 1) % 3], t[(k + 2) % 3]};
    };
    std::vector<std::pair<uint32_t, uint32_t>> nbr;
    for (uint32_t f : fs)
      for (uint32_t w : wedge(f)) nbr.push_back({w, f});
    std::sort(nbr.begin(), nbr.end());
    auto faces_of = [&](uint32_t w) {
      auto lo = std::lower_bound(nbr.begin(), nbr.end(), std::make_pair(w, 0u));
      auto hi = std::upper_bound(nbr.begin(), nbr.end(), std::make_pair(w, kNone));
      return std::make_pair(lo, hi);
    };
    std::vector<std::pair<uint32_t, uint32_t>> ends;  // (face, neighbour)
    size_t i = 0;
    while (i < nbr.size()) {
      size_t j = i;
      while (j < nbr.size() && nbr[j].first == nbr[i].first) ++j;
      if (j - i == 1)
        ends.push_back({nbr[i].second, nbr[i].first});
      else if (j - i != 2)
        return false;
      i = j;
    }
    bool boundary;
    if (ends.empty())
      boundary = false;
    else if (ends.size() == 2)
      boundary = true;
    else
      return false;
    uint32_t start, entry;
    if (boundary) {
      auto m = *std::min_element(ends.begin(), ends.end());
      start = m.first;
      entry = m.second;
    } else {
      start = *std::min_element(fs.begin(), fs.end());
      entry = wedge(start)[0];
    }
    Fan o;
    o.faces.push_back(start);
    o.flip.push_back(false);
    o.boundary = boundary;
    uint32_t cur = start;
    bool cur_flip = false;
    for (;;) {
      const auto w = wedge(cur);
      const uint32_t exit = w[0] == entry ? w[1] : w[0];
      auto [lo, hi] = faces_of(exit);
      uint32_t next = kNone;
      for (auto it = lo; it != hi; ++it)
        if (it->second != cur) {
          next = it->second;
          break;
        }
      if (next == kNone) break;
      if (next == start) break;
      if (o.faces.size() > fs.size()) return false;
      const bool same = traverses(faces[cur], v, exit) == traverses(faces[next], v, exit);
      cur_flip ^= same;
      o.faces.push_back(next);
      o.flip.push_back(cur_flip);
      cur = next;
      entry = exit;
    }
    if (o.faces.size() != fs.size()) return false;
    *out = std::move(o);
    return true;
  }
  // mesh.rs:558 `vertex_ok`
  bool vertex_ok(uint32_t v) const {
    if (!v_alive[v]) return true;
    Fan f;
    if (!fan(v, &f)) return false;
    const auto& fs = vfaces[v];
    for (size_t i = 0; i < fs.size(); ++i) {
      const Tri si = sorted3(faces[fs[i]]);
      for (size_t j = i + 1; j < fs.size(); ++j)
        if (si == sorted3(faces[fs[j]])) return false;
    }
    return true;
  }
  // mesh.rs:577 `remove_vface` (swap_remove)
  void remove_vface(uint32_t v, uint32_t f) {
    auto& l = vfaces[v];
    for (size_t i = 0; i < l.size(); ++i)
      if (l[i] == f) {
        l[i] = l.back();
        l.pop_back();
        return;
      }
  }

  // mesh.rs:592 `split`
  int split(uint32_t a, uint32_t b, uint32_t* m_out) {
    const auto inc = edge_faces(a, b);
    if (inc.empty()) return REMESH_NO_SUCH_EDGE;
    const uint32_t m = (uint32_t)pos.size();
    const V3 mid = scale(add(pos[a], pos[b]), 0.5);
    for (uint32_t f : inc) {
      uint32_t p, q, c;
      if (orient(faces[f], a, b, &p, &q, &c))
        if (degenerate({pos[p], mid, pos[c]}) || degenerate({mid, pos[q], pos[c]})) return REMESH_DEGENERATE_FACE;
    }
    pos.push_back(mid);
    v_alive.push_back(1);
    pinned.push_back(0);
    pin_origin.push_back(kNone);
    sources.push_back(combine(sources[a], 0.5, sources[b], 0.5));
    vfaces.emplace_back();
    for (uint32_t f : inc) {
      uint32_t p, q, c;
      if (!orient(faces[f], a, b, &p, &q, &c)) continue;
      const uint32_t nf = (uint32_t)faces.size();
      faces[f] = {p, m, c};
      faces.push_back({m, q, c});
      face_alive.push_back(1);
      parent.push_back(parent[f]);
      remove_vface(q, f);
      vfaces[q].push_back(nf);
      vfaces[c].push_back(nf);
      vfaces[m].push_back(f);
      vfaces[m].push_back(nf);
    }
    *m_out = m;
    return REMESH_OK;
  }

  // mesh.rs:652 `collapse_target`
  void collapse_target(uint32_t a, uint32_t b, V3* x, double* wa, double* wb) const {
    const bool pa = pinned[a], pb = pinned[b];
    const bool ba = is_boundary_vertex(a), bb = is_boundary_vertex(b);
    if (pa) {
      *x = pos[a], *wa = 1.0, *wb = 0.0;
    } else if (pb) {
      *x = pos[b], *wa = 0.0, *wb = 1.0;
    } else if (ba && !bb) {
      *x = pos[a], *wa = 1.0, *wb = 0.0;
    } else if (bb && !ba) {
      *x = pos[b], *wa = 0.0, *wb = 1.0;
    } else {
      *x = scale(add(pos[a], pos[b]), 0.5), *wa = 0.5, *wb = 0.5;
    }
  }

  // mesh.rs:679 `collapse`
  int collapse(uint32_t a, uint32_t b, bool foldover, CollapseGuard* guard, uint32_t* keep_out) {
    const auto inc = edge_faces(a, b);
    if (inc.empty()) return REMESH_NO_SUCH_EDGE;
    if (pinned[a] && pinned[b]) return REMESH_BOTH_PINNED;
    const auto na = neighbors(a), nb = neighbors(b);
    size_t common = 0;
    for (uint32_t w : na)
      if (std::binary_search(nb.begin(), nb.end(), w)) ++common;
    if (common != inc.size()) return REMESH_LINK_CONDITION;
    if (inc.size() == 2 && is_boundary_vertex(a) && is_boundary_vertex(b)) return REMESH_LINK_CONDITION;
    const uint32_t keep = std::min(a, b), gone = std::max(a, b);
    V3 x;
    double wa, wb;
    collapse_target(a, b, &x, &wa, &wb);
    std::vector<uint32_t> star;
    for (uint32_t f : vfaces[a])
      if (std::find(inc.begin(), inc.end(), f) == inc.end()) star.push_back(f);
    for (uint32_t f : vfaces[b])
      if (std::find(inc.begin(), inc.end(), f) == inc.end()) star.push_back(f);
    std::sort(star.begin(), star.end());
    star.erase(std::unique(star.begin(), star.end()), star.end());
    auto moved_corners = [&](uint32_t f) {
      const Tri& t = faces[f];
      auto p = corners(f);
      for (int k = 0; k < 3; ++k)
        if (t[k] == a || t[k] == b) p[k] = x;
      return p;
    };
    if (foldover)
      for (uint32_t f : star)
        if (remesh_folds(remesh_tri_normal(corners(f)), remesh_tri_normal(moved_corners(f)))) return REMESH_FOLDOVER;
    for (uint32_t f : star)
      if (degenerate(moved_corners(f))) return REMESH_DEGENERATE_FACE;
    std::vector<Tri> new_star;
    for (uint32_t f : star) {
      Tri t = faces[f];
      for (auto& v : t)
        if (v == gone) v = keep;
      new_star.push_back(t);
    }
    for (size_t i = 0; i < new_star.size(); ++i) {
      const Tri si = sorted3(new_star[i]);
      for (size_t j = i + 1; j < new_star.size(); ++j)
        if (si == sorted3(new_star[j])) return REMESH_REPEATED_FACE;
    }
    if (guard && !guard_allow(guard, *this, a, b, keep, x, new_star)) return REMESH_GUARD;

    // apply, with a snapshot for the undo
    std::vector<uint32_t> touched = {a, b};
    for (uint32_t f : inc)
      for (uint32_t v : faces[f]) touched.push_back(v);
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> snap_v;
    for (uint32_t v : touched) snap_v.push_back({v, vfaces[v]});
    std::vector<std::pair<uint32_t, Tri>> snap_f;
    for (uint32_t f : star) snap_f.push_back({f, faces[f]});
    const V3 old_pos = pos[keep];
    const uint8_t old_pin = pinned[keep];
    const uint32_t old_origin = pin_origin[keep];
    const Src old_src = sources[keep];

    for (uint32_t f : inc) {
      face_alive[f] = 0;
      for (uint32_t v : faces[f]) remove_vface(v, f);
    }
    std::vector<uint32_t> moved;
    std::swap(moved, vfaces[gone]);
    for (uint32_t f : moved) {
      for (auto& v : faces[f])
        if (v == gone) v = keep;
      vfaces[keep].push_back(f);
    }
    v_alive[gone] = 0;
    pos[keep] = x;
    const bool pin = pinned[a] || pinned[b];
    const uint32_t origin = pinned[a] ? pin_origin[a] : (pinned[b] ? pin_origin[b] : kNone);
    pinned[keep] = pin;
    pin_origin[keep] = origin;
    sources[keep] = combine(sources[a], wa, sources[b], wb);

    auto check = neighbors(keep);
    check.push_back(keep);
    for (uint32_t v : touched)
      if (v != gone) check.push_back(v);
    std::sort(check.begin(), check.end());
    check.erase(std::unique(check.begin(), check.end()), check.end());
    for (uint32_t v : check)
      if (!vertex_ok(v)) {
        for (auto& [vv, l] : snap_v) vfaces[vv] = l;
        for (auto& [f, t] : snap_f) faces[f] = t;
        for (uint32_t f : inc) face_alive[f] = 1;
        v_alive[gone] = 1;
        pos[keep] = old_pos;
        pinned[keep] = old_pin;
        pin_origin[keep] = old_origin;
        sources[keep] = old_src;
        return REMESH_NON_MANIFOLD_RESULT;
      }
    *keep_out = keep;
    return REMESH_OK;
  }

  // mesh.rs:888 `flip`
  int flip(uint32_t a, uint32_t b, bool delaunay, bool foldover, bool* reordered) {
    const auto inc = edge_faces(a, b);
    if (inc.empty()) return REMESH_NO_SUCH_EDGE;
    if (inc.size() != 2) return REMESH_BOUNDARY_EDGE;
    const uint32_t f1 = inc[0], f2 = inc[1];
    const Tri t1 = faces[f1], t2 = faces[f2];
    uint32_t p, q, c, p2, q2, d;
    if (!orient(t1, a, b, &p, &q, &c) || !orient(t2, a, b, &p2, &q2, &d)) return REMESH_NO_SUCH_EDGE;
    *reordered = p2 == p;
    if (!edge_faces(c, d).empty() || c == d) return REMESH_EXISTING_EDGE;
    const V3 xp = pos[p], xq = pos[q], xc = pos[c], xd = pos[d];
    if (delaunay && !remesh_violates_delaunay(xp, xq, xc, xd)) return REMESH_NOT_DELAUNAY;
    if (foldover) {
      // mesh.rs:1058 `flip_folds`
      const V3 n1 = remesh_tri_normal(corners(f1));
      V3 n2 = remesh_tri_normal(corners(f2));
      if (*reordered) n2 = scale(n2, -1.0);
      if (remesh_folds(n1, remesh_tri_normal({xp, xd, xc})) || remesh_folds(n2, remesh_tri_normal({xd, xq, xc})))
        return REMESH_FOLDOVER;
    }
    if (degenerate({xp, xd, xc}) || degenerate({xd, xq, xc})) return REMESH_DEGENERATE_FACE;
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> snap;
    for (uint32_t v : {p, q, c, d}) snap.push_back({v, vfaces[v]});
    faces[f1] = {p, d, c};
    faces[f2] = {d, q, c};
    remove_vface(p, f2);
    remove_vface(q, f1);
    vfaces[c].push_back(f2);
    vfaces[d].push_back(f1);
    for (uint32_t v : {p, q, c, d})
      if (!vertex_ok(v)) {
        faces[f1] = t1;
        faces[f2] = t2;
        for (auto& [vv, l] : snap) vfaces[vv] = l;
        return REMESH_NON_MANIFOLD_RESULT;
      }
    return REMESH_OK;
  }
};

// ccd.rs:1448 `CcdCollapseGuard` (the grid of vertex cells, built on the first query)
struct CollapseGuard {
  double gap_abs = 0.0, cell = 0.0, max_edge = 0.0;
  bool built = false;
  std::map<std::array<int64_t, 3>, std::vector<uint32_t>> grid;
  bool has_last = false;
  uint32_t last_a = 0, last_b = 0;
  V3 last_x = {0, 0, 0};
  const std::vector<uint8_t>* band = nullptr;  // empty or null: no band
  uint64_t refused = 0, tested = 0, exempt = 0;

  std::array<int64_t, 3> cell_of(V3 p) const {
    return {(int64_t)std::floor(p[0] / cell), (int64_t)std::floor(p[1] / cell), (int64_t)std::floor(p[2] / cell)};
  }
  bool flag(uint32_t v) const { return band && (size_t)v < band->size() && (*band)[v]; }
  // ccd.rs:1514 `build`
  void build(const EditMesh& m) {
    double me = 0.0;
    for (uint32_t f = 0; f < (uint32_t)m.faces.size(); ++f)
      if (m.face_alive[f]) {
        const Tri& t = m.faces[f];
        for (int k = 0; k < 3; ++k) me = fmax_(me, norm(sub(m.pos[t[k]], m.pos[t[(k + 1) % 3]])));
      }
    max_edge = fmax_(me, gap_abs);
    cell = max_edge;
    grid.clear();
    for (uint32_t v = 0; v < (uint32_t)m.pos.size(); ++v)
      if (m.v_alive[v]) grid[cell_of(m.pos[v])].push_back(v);
    built = true;
  }
  // ccd.rs:1538 `allow`
  bool allow(const EditMesh& m, uint32_t ea, uint32_t eb, uint32_t survivor, V3 position,
             const std::vector<Tri>& new_star) {
    if (!built) build(m);
    if (has_last) {
      has_last = false;
      const uint32_t s = m.v_alive[last_a] ? last_a : last_b, r = m.v_alive[last_a] ? last_b : last_a;
      if (!m.v_alive[r] && m.v_alive[s]) grid[cell_of(last_x)].push_back(s);
    }
    has_last = true;
    last_a = ea, last_b = eb, last_x = position;
    auto at_new = [&](uint32_t v) { return v == survivor ? position : m.pos[v]; };
    double star_max = 0.0;
    for (auto& t : new_star)
      for (int k = 0; k < 3; ++k) star_max = fmax_(star_max, norm(sub(at_new(t[k]), at_new(t[(k + 1) % 3]))));
    max_edge = fmax_(max_edge, star_max);
    const double reach = gap_abs + max_edge;
    V3 lo = {INFINITY, INFINITY, INFINITY}, hi = {-INFINITY, -INFINITY, -INFINITY};
    for (auto& t : new_star)
      for (uint32_t v : t) {
        const V3 p = at_new(v);
        for (int k = 0; k < 3; ++k) {
          lo[k] = fmin_(lo[k], p[k]);
          hi[k] = fmax_(hi[k], p[k]);
        }
      }
    for (int k = 0; k < 3; ++k) {
      lo[k] = lo[k] - reach;
      hi[k] = hi[k] + reach;
    }
    const auto c0 = cell_of(lo), c1 = cell_of(hi);
    std::vector<uint32_t> cand;
    for (int64_t i = c0[0]; i <= c1[0]; ++i)
      for (int64_t j = c0[1]; j <= c1[1]; ++j)
        for (int64_t k = c0[2]; k <= c1[2]; ++k) {
          auto it = grid.find({i, j, k});
          if (it == grid.end()) continue;
          for (uint32_t w : it->second) {
            if (w == ea || w == eb || !m.v_alive[w]) continue;
            const V3 p = m.pos[w];
            if (!(p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] && p[2] <= hi[2]))
              continue;
            for (uint32_t f : m.vfaces[w]) cand.push_back(f);
          }
        }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    std::vector<bool> star_in(new_star.size());
    for (size_t i = 0; i < new_star.size(); ++i) {
      bool in = false;
      for (uint32_t v : new_star[i]) in = in || flag(v) || (v == survivor && (flag(ea) || flag(eb)));
      star_in[i] = in;
    }
    const bool has_band = band && !band->empty();
    bool refuse = false;
    for (uint32_t g : cand) {
      const Tri& gv = m.faces[g];
      if (m.contains(gv, ea) || m.contains(gv, eb)) continue;
      const std::array<V3, 3> gt = {m.pos[gv[0]], m.pos[gv[1]], m.pos[gv[2]]};
      V3 glo = {INFINITY, INFINITY, INFINITY}, ghi = {-INFINITY, -INFINITY, -INFINITY};
      for (auto& p : gt)
        for (int k = 0; k < 3; ++k) {
          glo[k] = fmin_(glo[k], p[k]);
          ghi[k] = fmax_(ghi[k], p[k]);
        }
      for (int k = 0; k < 3; ++k) {
        glo[k] = glo[k] - gap_abs;
        ghi[k] = ghi[k] + gap_abs;
      }
      const bool g_in = has_band && (flag(gv[0]) || flag(gv[1]) || flag(gv[2]));
      for (size_t ti = 0; ti < new_star.size(); ++ti) {
        const Tri& t = new_star[ti];
        if (m.contains(gv, t[0]) || m.contains(gv, t[1]) || m.contains(gv, t[2])) continue;
        if (g_in && star_in[ti]) {
          ++exempt;
          continue;
        }
        const std::array<V3, 3> tt = {at_new(t[0]), at_new(t[1]), at_new(t[2])};
        V3 tlo = {INFINITY, INFINITY, INFINITY}, thi = {-INFINITY, -INFINITY, -INFINITY};
        for (auto& p : tt)
          for (int k = 0; k < 3; ++k) {
            tlo[k] = fmin_(tlo[k], p[k]);
            thi[k] = fmax_(thi[k], p[k]);
          }
        bool ov = true;
        for (int k = 0; k < 3; ++k) ov = ov && glo[k] <= thi[k] && tlo[k] <= ghi[k];
        if (!ov) continue;
        ++tested;
        if (remesh_triangle_distance(tt, gt) < gap_abs) {
          refuse = true;
          break;
        }
      }
      if (refuse) break;
    }
    if (refuse) ++refused;
    return !refuse;
  }
};

bool guard_allow(CollapseGuard* g, const EditMesh& m, uint32_t ea, uint32_t eb, uint32_t survivor, V3 position,
                 const std::vector<std::array<uint32_t, 3>>& new_star) {
  return g->allow(m, ea, eb, survivor, position, new_star);
}

namespace {

// mod.rs:154 `RemeshOptions::validate`
bool validate(const RemeshOptions& o, Error* err) {
  auto bad = [&](const char* name, const std::string& req) {
    err->status = INVALID_OPTION;
    err->detail = std::string(name) + ": " + req;
    return false;
  };
  if (!(std::isfinite(o.split) && o.split > 0.0)) return bad("split", "must be finite and positive");
  if (!(std::isfinite(o.collapse) && o.collapse >= 0.0)) return bad("collapse", "must be finite and non-negative");
  if (o.split <= 2.0 * o.collapse)
    return bad("split", "must exceed 2 x collapse (the halves of a split edge must not be collapsible)");
  if (!(o.rho > 0.0 && o.rho <= 1.0)) return bad("rho", "must lie in (0, 1]");
  if (o.flip_cap_per_vertex == 0) return bad("flip_cap_per_vertex", "must be at least 1");
  if (std::isnan(o.max_growth) || o.max_growth < 1.0)
    return bad("max_growth", "must be at least 1 (f64::INFINITY disables it)");
  return true;
}

double edge_len(const EditMesh& m, uint32_t a, uint32_t b) { return norm(sub(m.pos[a], m.pos[b])); }

uint64_t f64_bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

using ShortKey = std::array<uint64_t, 3>;  // (length bits, min, max)

void push_short(std::priority_queue<ShortKey, std::vector<ShortKey>, std::greater<ShortKey>>& heap,
                const EditMesh& m, uint32_t a, uint32_t b, double thr) {
  const double l = edge_len(m, a, b);
  if (l < thr) heap.push({f64_bits(l), std::min(a, b), std::max(a, b)});
}

// mod.rs:338 `collapse_short`
void collapse_short(EditMesh& m, double thr, CollapseGuard* guard, RemeshReport* rep) {
  std::priority_queue<ShortKey, std::vector<ShortKey>, std::greater<ShortKey>> heap;
  for (auto& e : m.edges()) push_short(heap, m, e[0], e[1], thr);
  while (!heap.empty()) {
    const ShortKey top = heap.top();
    heap.pop();
    const uint32_t a = (uint32_t)top[1], b = (uint32_t)top[2];
    if (!m.v_alive[a] || !m.v_alive[b]) continue;
    if (m.edge_faces(a, b).empty()) continue;
    const double l = edge_len(m, a, b);
    if (f64_bits(l) != top[0] || l >= thr) continue;
    uint32_t keep;
    const int r = m.collapse(a, b, true, guard, &keep);
    if (r == REMESH_OK) {
      ++rep->collapses;
      for (uint32_t w : m.neighbors(keep)) push_short(heap, m, keep, w, thr);
    } else {
      ++rep->refused_collapses[r - 1];
    }
  }
}

uint64_t edge_key(uint32_t a, uint32_t b) { return ((uint64_t)std::min(a, b) << 32) | std::max(a, b); }

// mod.rs:380 `delaunay_flips`
void delaunay_flips(EditMesh& m, size_t cap_per_vertex, RemeshReport* rep) {
  size_t nv_live = 0;
  for (uint8_t a : m.v_alive) nv_live += a;
  const uint64_t cap = (uint64_t)cap_per_vertex * (uint64_t)nv_live;
  std::deque<std::array<uint32_t, 2>> queue;
  std::unordered_set<uint64_t> in_queue;
  for (auto& e : m.edges()) {
    queue.push_back(e);
    in_queue.insert(edge_key(e[0], e[1]));
  }
  uint64_t flips = 0;
  while (!queue.empty()) {
    const auto [a, b] = queue.front();
    queue.pop_front();
    in_queue.erase(edge_key(a, b));
    const auto inc = m.edge_faces(a, b);
    if (inc.size() != 2) continue;
    auto quad = [&](uint32_t f) {
      for (uint32_t v : m.faces[f])
        if (v != a && v != b) return v;
      return a;
    };
    const uint32_t c = quad(inc[0]), d = quad(inc[1]);
    if (!remesh_violates_delaunay(m.pos[a], m.pos[b], m.pos[c], m.pos[d])) continue;
    if (flips >= cap) {
      ++rep->flip_cap_hits;
      break;
    }
    bool reordered;
    const int r = m.flip(a, b, false, true, &reordered);
    if (r == REMESH_OK) {
      ++flips;
      const std::array<std::array<uint32_t, 2>, 4> quad_edges = {{{a, c}, {c, b}, {b, d}, {d, a}}};
      for (auto& xy : quad_edges)
        if (in_queue.insert(edge_key(xy[0], xy[1])).second)
          queue.push_back({std::min(xy[0], xy[1]), std::max(xy[0], xy[1])});
    } else {
      ++rep->refused_flips[r - 1];
    }
  }
  rep->flips += flips;
}

// mod.rs:424 `smooth`
void smooth(EditMesh& m, double rho, RemeshReport* rep) {
  const size_t nf = m.faces.size(), nv = m.pos.size();
  std::vector<uint8_t> boundary(nv, 0), face_on_boundary(nf, 0);
  for (auto& [e, n] : m.edges_with_counts())
    if (n == 1) {
      boundary[e[0]] = boundary[e[1]] = 1;
      for (uint32_t f : m.edge_faces(e[0], e[1])) face_on_boundary[f] = 1;
    }
  std::vector<V3> normal0(nf, V3{0, 0, 0}), center(nf, V3{0, 0, 0});
  std::vector<double> area(nf, 0.0);
  for (uint32_t f = 0; f < (uint32_t)nf; ++f) {
    if (!m.face_alive[f]) continue;
    const auto p = m.corners(f);
    const V3 n = remesh_tri_normal(p);
    normal0[f] = n;
    area[f] = 0.5 * norm(n);
    center[f] = face_on_boundary[f] ? remesh_barycenter(p) : remesh_circumcenter(p);
  }
  std::vector<std::pair<uint32_t, V3>> moves;
  for (uint32_t v = 0; v < (uint32_t)nv; ++v) {
    if (!m.v_alive[v] || m.pinned[v] || boundary[v]) continue;
    const V3 x = m.pos[v];
    Fan fan;
    if (!m.fan(v, &fan)) {
      ++rep->smoothing_skipped;
      continue;
    }
    V3 nrm = {0, 0, 0};
    for (size_t i = 0; i < fan.faces.size(); ++i) {
      const uint32_t f = fan.faces[i];
      const Tri& t = m.faces[f];
      int k = 0;
      for (int j = 0; j < 3; ++j)
        if (t[j] == v) {
          k = j;
          break;
        }
      const double ang = remesh_corner_angle(x, m.pos[t[(k + 1) % 3]], m.pos[t[(k + 2) % 3]]);
      const V3 n = normal0[f];
      const double len = norm(n);
      const double s = fan.flip[i] ? -ang / len : ang / len;
      nrm = add(nrm, scale(n, s));
    }
    const double len = norm(nrm);
    std::vector<std::pair<double, V3>> fc;
    for (uint32_t f : fan.faces) fc.push_back({area[f], center[f]});
    V3 u;
    const bool disp = remesh_smoothing_displacement(x, rho, fc, &u);
    if (!(disp && len > 0.0 && std::isfinite(len))) {
      ++rep->smoothing_skipped;
      continue;
    }
    const V3 n = scale(nrm, 1.0 / len);
    u = sub(u, scale(n, dot(n, u)));
    if (std::isfinite(u[0]) && std::isfinite(u[1]) && std::isfinite(u[2]))
      moves.push_back({v, u});
    else
      ++rep->smoothing_skipped;
  }
  for (auto& [v, u] : moves) {
    const V3 x = m.pos[v];
    V3 step = u;
    bool applied = false;
    for (int halvings = 0; halvings <= 3; ++halvings) {
      const V3 y = add(x, step);
      bool ok = true;
      for (uint32_t f : m.vfaces[v]) {
        const Tri& t = m.faces[f];
        auto p = m.corners(f);
        for (int k = 0; k < 3; ++k)
          if (t[k] == v) p[k] = y;
        if (remesh_folds(normal0[f], remesh_tri_normal(p))) {
          ok = false;
          break;
        }
      }
      if (ok) {
        m.pos[v] = y;
        ++rep->smoothing_moves;
        if (halvings > 0) ++rep->smoothing_halved;
        applied = true;
        break;
      }
      step = scale(step, 0.5);
    }
    if (!applied) ++rep->smoothing_refused;
  }
}

}  // namespace

// mesh.rs:283 `EditMesh::new`
bool edit_mesh_new(const Mesh& s, const Positions& p, const std::vector<uint32_t>& pins, EditMesh** out, Error* err) {
  const size_t nv = (size_t)s.nv;
  if (p.size() != nv) {
    err->status = INVALID_OPTION;
    err->detail = "positions: one per vertex";
    return false;
  }
  for (size_t v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(p[v][k])) {
        err->status = NON_FINITE_POSITION;
        err->detail = "vertex " + std::to_string(v);
        err->a = (int64_t)v;
        return false;
      }
  auto* m = new EditMesh();
  m->pinned.assign(nv, 0);
  m->pin_origin.assign(nv, kNone);
  for (uint32_t q : pins) {
    if ((size_t)q >= nv) {
      delete m;
      err->status = INVALID_OPTION;
      err->detail = "pinned: " + std::to_string(q) + " (the surface has " + std::to_string(nv) + " vertices)";
      return false;
    }
    m->pinned[q] = 1;
    m->pin_origin[q] = q;
  }
  m->faces = s.faces;
  m->vfaces.assign(nv, {});
  for (uint32_t f = 0; f < (uint32_t)m->faces.size(); ++f)
    for (uint32_t v : m->faces[f]) m->vfaces[v].push_back(f);
  m->pos = p;
  m->face_alive.assign(m->faces.size(), 1);
  m->parent.resize(m->faces.size());
  for (uint32_t f = 0; f < (uint32_t)m->faces.size(); ++f) m->parent[f] = f;
  m->v_alive.assign(nv, 1);
  m->sources.resize(nv);
  for (uint32_t v = 0; v < (uint32_t)nv; ++v) m->sources[v] = {{v, 1.0}};
  m->pins_in = pins;
  m->n_vertices_in = nv;
  *out = m;
  return true;
}

void edit_mesh_free(EditMesh* m) { delete m; }

int edit_mesh_op(EditMesh& m, int op, uint32_t a, uint32_t b, bool delaunay, bool foldover, bool* reordered) {
  *reordered = false;
  uint32_t r;
  switch (op) {
    case 0:
      return m.split(a, b, &r);
    case 1:
      return m.collapse(a, b, foldover, nullptr, &r);
    default:
      return m.flip(a, b, delaunay, foldover, reordered);
  }
}

// mesh.rs:991 `finish`: drops the dead slots keeping the survivors' order, revalidates
bool edit_mesh_finish(const EditMesh& m, Remeshed* out, Error* err) {
  std::vector<uint32_t> new_id(m.pos.size(), kNone);
  Remeshed r;
  r.map_offsets.push_back(0);
  for (size_t v = 0; v < m.pos.size(); ++v)
    if (m.v_alive[v]) {
      new_id[v] = (uint32_t)r.positions.size();
      r.positions.push_back(m.pos[v]);
      for (auto& [o, w] : m.sources[v]) {
        r.map_sources.push_back(o);
        r.map_weights.push_back(w);
      }
      r.map_offsets.push_back((uint32_t)r.map_sources.size());
    }
  for (size_t f = 0; f < m.faces.size(); ++f)
    if (m.face_alive[f]) {
      r.faces.push_back({new_id[m.faces[f][0]], new_id[m.faces[f][1]], new_id[m.faces[f][2]]});
      r.face_parents.push_back(m.parent[f]);
    }
  // the result must be a valid surface (Surface::new + validate_positions): every index a live
  // vertex, no degenerate face
  for (size_t f = 0; f < r.faces.size(); ++f) {
    for (uint32_t v : r.faces[f])
      if (v == kNone) {
        err->status = INVALID_OPTION;
        err->detail = "internal: face " + std::to_string(f) + " references a removed vertex";
        return false;
      }
    if (degenerate({r.positions[r.faces[f][0]], r.positions[r.faces[f][1]], r.positions[r.faces[f][2]]})) {
      err->status = DEGENERATE_FACE;
      err->detail = "internal: face " + std::to_string(f) + " is degenerate after remeshing";
      err->a = (int64_t)f;
      return false;
    }
  }
  std::vector<uint32_t> pin_new(m.n_vertices_in, kNone);
  for (size_t v = 0; v < m.pos.size(); ++v)
    if (m.v_alive[v] && m.pin_origin[v] != kNone) pin_new[m.pin_origin[v]] = new_id[v];
  for (uint32_t q : m.pins_in) r.pinned.push_back(pin_new[q]);
  r.n_old = m.n_vertices_in;
  *out = std::move(r);
  return true;
}

// mod.rs:235 `remesh`
bool remesh(const Mesh& s, const Positions& p, const std::vector<uint32_t>& pins, double l0, const RemeshOptions& o,
            double guard_gap, const std::vector<uint8_t>* band, Remeshed* out, RemeshReport* rep, Error* err) {
  *rep = RemeshReport{};
  if (!validate(o, err)) return false;
  if (!(std::isfinite(l0) && l0 > 0.0)) {
    err->status = INVALID_OPTION;
    err->detail = "l0: the target edge length must be finite and positive";
    return false;
  }
  EditMesh* mp = nullptr;
  if (!edit_mesh_new(s, p, pins, &mp, err)) return false;
  std::unique_ptr<EditMesh> hold(mp);
  EditMesh& m = *mp;
  if (o.en_split) {
    const double thr = o.split * l0;
    // the input's edges in lexicographic order, with their face counts
    const auto ewc = m.edges_with_counts();
    struct Long {
      size_t e;
      double l;
      uint64_t added;
    };
    std::vector<Long> lng;
    for (size_t e = 0; e < ewc.size(); ++e) {
      const double l = edge_len(m, ewc[e].first[0], ewc[e].first[1]);
      if (l > thr) lng.push_back({e, l, ewc[e].second == 1 ? 1u : 2u});
    }
    const double budget = (o.max_growth - 1.0) * (double)s.faces.size();
    uint64_t total = 0;
    for (auto& x : lng) total += x.added;
    if ((double)total > budget) {
      std::vector<size_t> order(lng.size());
      for (size_t i = 0; i < order.size(); ++i) order[i] = i;
      std::sort(order.begin(), order.end(), [&](size_t i, size_t j) {
        if (lng[i].l != lng[j].l) return lng[j].l < lng[i].l;  // longest first (total_cmp on finite lengths)
        return lng[i].e < lng[j].e;
      });
      std::vector<bool> keep(lng.size(), false);
      double used = 0.0;
      for (size_t i : order) {
        if (used + (double)lng[i].added > budget) continue;
        used += (double)lng[i].added;
        keep[i] = true;
      }
      std::vector<Long> kept;
      for (size_t i = 0; i < lng.size(); ++i)
        if (keep[i])
          kept.push_back(lng[i]);
        else
          ++rep->splits_deferred;
      lng = std::move(kept);
    }
    for (auto& x : lng) {
      uint32_t mid;
      const int r = m.split(ewc[x.e].first[0], ewc[x.e].first[1], &mid);
      if (r == REMESH_OK)
        ++rep->splits;
      else
        ++rep->refused_splits[r - 1];
    }
  }
  CollapseGuard guard;
  if (guard_gap > 0.0) {
    guard.gap_abs = guard_gap;
    guard.band = band && !band->empty() ? band : nullptr;
  }
  if (o.en_collapse) collapse_short(m, o.collapse * l0, guard_gap > 0.0 ? &guard : nullptr, rep);
  for (int r = 0; r < o.rounds; ++r) {
    if (o.order == 0) {
      if (o.en_flip) delaunay_flips(m, o.flip_cap_per_vertex, rep);
      if (o.en_smooth) smooth(m, o.rho, rep);
    } else {
      if (o.en_smooth) smooth(m, o.rho, rep);
      if (o.en_flip) delaunay_flips(m, o.flip_cap_per_vertex, rep);
    }
  }
  if (!edit_mesh_finish(m, out, err)) return false;
  rep->n_vertices = out->positions.size();
  rep->n_faces = out->faces.size();
  rep->guard_refused = guard.refused;
  rep->guard_tested = guard.tested;
  rep->guard_exempt = guard.exempt;
  return true;
}

}  // namespace orc
