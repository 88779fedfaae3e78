// Host construction of the tree structures (see trees.hpp).  Follows the reference's rules so
// that the structures are the reference's exactly:
//   RC/src/spatial/bvh.rs:156-205 (split rule), :221-254 (bottom-up node data, leaves in element
//   order), RS/src/spatial/mod.rs:113-145 (box distance, farthest corner, Barnes–Hut test),
//   RS/src/energy/barnes_hut.rs:743-827 (lists, transposes), RS/src/metric/hierarchical.rs:116
//   (separation), RS/src/spatial/block_tree.rs:129-372 (blocks, grouping, near offsets, chains).
// The BVH is built level by level (all nodes of one depth, then the next): the split of a node
// depends only on its own elements, so the result equals the reference's depth-first build; node
// indices are then assigned in pre-order from the subtree sizes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "host_parallel.hpp"
#include "trees.hpp"

#pragma clang fp contract(off)

namespace ddgrs {

namespace {

// f64::total_cmp as an integer key.
inline int64_t total_key(double x) {
  int64_t b;
  std::memcpy(&b, &x, 8);
  return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
}

struct Box {
  double lo[3], hi[3];
};
inline Box empty_box() { return {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}}; }
inline void unite(Box& a, const double* lo, const double* hi) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::fmin(a.lo[k], lo[k]);
    a.hi[k] = std::fmax(a.hi[k], hi[k]);
  }
}
inline double diam2(const Box& b) {
  if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) return 0.0;
  double s = 0.0;
  for (int k = 0; k < 3; ++k) {
    double w = b.hi[k] - b.lo[k];
    s += w * w;
  }
  return s;
}


inline double box_distance_sq(const double* alo, const double* ahi, const double* blo, const double* bhi) {
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double d = std::fmax(std::fmax(alo[k], blo[k]) - std::fmin(ahi[k], bhi[k]), 0.0);
    d2 += d * d;
  }
  return d2;
}

}  // namespace

Topology build_topology(size_t n, const double* pts, const double* lo, const double* hi, size_t leaf_size) {
  using Seg = TreeSegment;
  std::vector<uint32_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
  std::vector<Seg> segs{{0, (uint32_t)n, 0}};
  std::vector<size_t> level{0};
  while (!level.empty()) {
    std::vector<int64_t> split(level.size(), -1);
    parallel(level.size(), [&](size_t li) {
      const Seg& s = segs[level[li]];
      const size_t cnt = s.end - s.start;
      if (cnt <= leaf_size) return;
      const int axis = (int)(s.depth % 3);
      uint32_t* sl = order.data() + s.start;
      std::sort(sl, sl + cnt, [&](uint32_t a, uint32_t b) {
        int64_t ka = total_key(pts[3 * a + axis]), kb = total_key(pts[3 * b + axis]);
        return ka != kb ? ka < kb : a < b;
      });
      std::vector<Box> suffix(cnt);
      Box acc = empty_box();
      for (size_t k = cnt; k-- > 0;) {
        unite(acc, lo + 3 * sl[k], hi + 3 * sl[k]);
        suffix[k] = acc;
      }
      Box prefix = empty_box();
      double best = INFINITY;
      size_t best_k = cnt / 2;
      size_t low = (size_t)std::ceil(0.25 * (double)cnt);
      low = std::min(std::max(low, (size_t)1), cnt / 2);
      for (size_t k = 1; k <= cnt - low; ++k) {
        unite(prefix, lo + 3 * sl[k - 1], hi + 3 * sl[k - 1]);
        if (k < low) continue;
        double cost = diam2(prefix) + diam2(suffix[k]);
        if (cost < best) {
          best = cost;
          best_k = k;
        }
      }
      split[li] = (int64_t)(s.start + best_k);
    });
    std::vector<size_t> next;
    for (size_t li = 0; li < level.size(); ++li) {
      if (split[li] < 0) continue;
      size_t si = level[li];
      Seg l{segs[si].start, (uint32_t)split[li], segs[si].depth + 1};
      Seg r{(uint32_t)split[li], segs[si].end, segs[si].depth + 1};
      segs[si].left = (int64_t)segs.size();
      segs.push_back(l);
      segs[si].right = (int64_t)segs.size();
      segs.push_back(r);
      next.push_back((size_t)segs[si].left);
      next.push_back((size_t)segs[si].right);
    }
    level = std::move(next);
  }
  return finish_topology(segs, std::move(order));
}

Topology finish_topology(const std::vector<TreeSegment>& segs, std::vector<uint32_t> order) {
  // Subtree sizes (children were appended after their parents), then pre-order ids.
  std::vector<uint32_t> count(segs.size(), 1), id(segs.size(), 0);
  for (size_t i = segs.size(); i-- > 0;)
    if (segs[i].left >= 0) count[i] = 1 + count[segs[i].left] + count[segs[i].right];
  for (size_t i = 0; i < segs.size(); ++i)
    if (segs[i].left >= 0) {
      id[segs[i].left] = id[i] + 1;
      id[segs[i].right] = id[i] + 1 + count[segs[i].left];
    }
  Topology t;
  const size_t nn = segs.size();
  t.order = std::move(order);
  t.start.resize(nn);
  t.end.resize(nn);
  t.right.assign(nn, 0);
  t.parent.assign(nn, UINT32_MAX);
  t.depth.resize(nn);
  for (size_t i = 0; i < nn; ++i) {
    uint32_t a = id[i];
    t.start[a] = segs[i].start;
    t.end[a] = segs[i].end;
    t.depth[a] = segs[i].depth;
    t.height = std::max(t.height, segs[i].depth + 1);
    if (segs[i].left >= 0) {
      t.right[a] = id[segs[i].right];
      t.parent[id[segs[i].left]] = a;
      t.parent[id[segs[i].right]] = a;
    }
  }
  t.leaf_ordinal.assign(nn, UINT32_MAX);
  for (size_t a = 0; a < nn; ++a)
    if (t.right[a] == 0) {
      t.leaf_ordinal[a] = (uint32_t)t.leaves.size();
      t.leaves.push_back((uint32_t)a);
    }
  return t;
}

Clusters build_clusters(const Topology& t, const double* pts, const double* lo, const double* hi, const double* mass,
                        const double* normals) {
  const size_t nn = t.n_nodes(), nl = t.leaves.size();
  Clusters c;
  c.mass.assign(nn, 0.0);
  c.centroid.assign(3 * nn, 0.0);
  c.lo.assign(3 * nn, 0.0);
  c.hi.assign(3 * nn, 0.0);
  c.radius.assign(nn, 0.0);
  std::vector<double> mom(3 * nn, 0.0);
  // fold of (Σa, Σ a NxNx, Σ a NxNy, Σ a NxNz, Σ a NyNy, Σ a NyNz, Σ a NzNz), then P̄ = s / Σa
  std::vector<double> ps(normals ? 7 * nn : 0, 0.0);
  // leaves (independent: in parallel), each over its elements in order
  parallel(nl, [&](size_t k) {
    const uint32_t a = t.leaves[k];
    Box b = empty_box();
    double m = 0.0, mo[3] = {0.0, 0.0, 0.0};
    for (uint32_t q = t.start[a]; q < t.end[a]; ++q) {
      uint32_t e = t.order[q];
      unite(b, lo + 3 * e, hi + 3 * e);
      m += mass[e];
      for (int k2 = 0; k2 < 3; ++k2) mo[k2] += mass[e] * pts[3 * e + k2];
    }
    c.mass[a] = m;
    for (int k2 = 0; k2 < 3; ++k2) {
      mom[3 * a + k2] = mo[k2];
      c.lo[3 * a + k2] = b.lo[k2];
      c.hi[3 * a + k2] = b.hi[k2];
    }
    if (normals) {
      double* o = &ps[7 * (size_t)a];
      for (uint32_t q = t.start[a]; q < t.end[a]; ++q) {
        uint32_t e = t.order[q];
        double w = mass[e];
        const double* n = normals + 3 * e;
        o[0] += w;
        o[1] += w * n[0] * n[0];
        o[2] += w * n[0] * n[1];
        o[3] += w * n[0] * n[2];
        o[4] += w * n[1] * n[1];
        o[5] += w * n[1] * n[2];
        o[6] += w * n[2] * n[2];
      }
    }
  });
  // internal nodes bottom-up (children have larger pre-order ids): left + right
  for (size_t a = nn; a-- > 0;) {
    if (t.is_leaf(a)) continue;
    const size_t l = a + 1, r = t.right[a];
    Box b = empty_box();
    unite(b, &c.lo[3 * l], &c.hi[3 * l]);
    unite(b, &c.lo[3 * r], &c.hi[3 * r]);
    c.mass[a] = c.mass[l] + c.mass[r];
    for (int k = 0; k < 3; ++k) {
      mom[3 * a + k] = mom[3 * l + k] + mom[3 * r + k];
      c.lo[3 * a + k] = b.lo[k];
      c.hi[3 * a + k] = b.hi[k];
    }
    if (normals)
      for (int k = 0; k < 7; ++k) ps[7 * a + k] = ps[7 * l + k] + ps[7 * r + k];
  }
  if (normals) c.projector.resize(6 * nn);
  parallel(nn, [&](size_t a) {
    if (normals)
      for (int k = 0; k < 6; ++k) c.projector[6 * a + k] = ps[7 * a + k + 1] / ps[7 * a];
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      double x = mom[3 * a + k] / c.mass[a];
      c.centroid[3 * a + k] = x;
    }
    for (int k = 0; k < 3; ++k) {
      double x = c.centroid[3 * a + k];
      double d = std::fmax(c.hi[3 * a + k] - x, x - c.lo[3 * a + k]);
      r2 += d * d;
    }
    c.radius[a] = std::sqrt(r2);
  });
  return c;
}

BhLists build_bh_lists(const Topology& t, const Clusters& c, double theta) {
  const size_t nl = t.leaves.size(), nn = t.n_nodes();
  std::vector<std::vector<uint32_t>> lum(nl), exa(nl);
  const double th2 = theta * theta;
  parallel(nl, [&](size_t k) {
    const uint32_t l = t.leaves[k];
    std::vector<uint32_t> stack{0};
    while (!stack.empty()) {
      uint32_t a = stack.back();
      stack.pop_back();
      double h = std::fmax(c.radius[l], c.radius[a]);
      bool adm = h * h < th2 * box_distance_sq(&c.lo[3 * l], &c.hi[3 * l], &c.lo[3 * a], &c.hi[3 * a]);
      if (adm)
        lum[k].push_back(a);
      else if (t.is_leaf(a))
        exa[k].push_back(a);
      else {
        stack.push_back(t.right[a]);
        stack.push_back(a + 1);
      }
    }
  });
  BhLists L;
  L.lumped_off.push_back(0);
  L.exact_off.push_back(0);
  for (size_t k = 0; k < nl; ++k) {
    uint64_t ns = t.end[t.leaves[k]] - t.start[t.leaves[k]];
    for (uint32_t a : exa[k]) L.pairs_exact += a == t.leaves[k] ? ns * (ns - 1) : ns * (uint64_t)(t.end[a] - t.start[a]);
    L.lumped.insert(L.lumped.end(), lum[k].begin(), lum[k].end());
    L.exact.insert(L.exact.end(), exa[k].begin(), exa[k].end());
    L.lumped_off.push_back((uint32_t)L.lumped.size());
    L.exact_off.push_back((uint32_t)L.exact.size());
  }
  // transposes, entries increasing within a target
  L.lumped_t_off.assign(nn + 1, 0);
  for (uint32_t a : L.lumped) L.lumped_t_off[a + 1]++;
  for (size_t i = 0; i < nn; ++i) L.lumped_t_off[i + 1] += L.lumped_t_off[i];
  L.lumped_t.resize(L.lumped.size());
  {
    std::vector<uint32_t> fill(L.lumped_t_off.begin(), L.lumped_t_off.end() - 1);
    for (size_t e = 0; e < L.lumped.size(); ++e) L.lumped_t[fill[L.lumped[e]]++] = (uint32_t)e;
  }
  L.exact_t_off.assign(nl + 1, 0);
  for (uint32_t a : L.exact) L.exact_t_off[t.leaf_ordinal[a] + 1]++;
  for (size_t i = 0; i < nl; ++i) L.exact_t_off[i + 1] += L.exact_t_off[i];
  L.exact_t.resize(L.exact.size());
  {
    std::vector<uint32_t> fill(L.exact_t_off.begin(), L.exact_t_off.end() - 1);
    for (size_t k = 0; k < nl; ++k)
      for (uint32_t e = L.exact_off[k]; e < L.exact_off[k + 1]; ++e) L.exact_t[fill[t.leaf_ordinal[L.exact[e]]]++] = (uint32_t)k;
  }
  return L;
}

BlockStructure build_block_structure(const Topology& t, const Clusters& c, double chi, size_t leaf_pairs) {
  const size_t nn = t.n_nodes();
  auto len = [&](size_t a) { return (size_t)(t.end[a] - t.start[a]); };
  auto separated = [&](size_t a, size_t b) {
    if (!(chi > 0.0)) return false;
    double d2 = box_distance_sq(&c.lo[3 * a], &c.hi[3 * a], &c.lo[3 * b], &c.hi[3 * b]);
    double h = std::fmax(c.radius[a], c.radius[b]);
    return d2 > 0.0 && h * h <= chi * chi * d2;
  };
  std::vector<std::pair<uint32_t, uint32_t>> adm, nearb, level{{0, 0}};
  BlockStructure s;
  while (!level.empty()) {
    s.depth++;
    std::vector<uint8_t> cls(level.size());
    parallel(level.size(), [&](size_t i) {
      auto [a, b] = level[i];
      if (a != b && separated(a, b))
        cls[i] = 0;
      else if ((t.is_leaf(a) && t.is_leaf(b)) || len(a) * len(b) <= leaf_pairs)
        cls[i] = 1;
      else
        cls[i] = 2;
    });
    std::vector<std::pair<uint32_t, uint32_t>> next;
    for (size_t i = 0; i < level.size(); ++i) {
      auto [a, b] = level[i];
      if (cls[i] == 0)
        adm.push_back({a, b});
      else if (cls[i] == 1)
        nearb.push_back({a, b});
      else {
        uint32_t ka[2] = {a, a}, kb[2] = {b, b};
        int na = 1, nb = 1;
        if (!t.is_leaf(a)) { ka[0] = a + 1; ka[1] = t.right[a]; na = 2; }
        if (!t.is_leaf(b)) { kb[0] = b + 1; kb[1] = t.right[b]; nb = 2; }
        for (int x = 0; x < na; ++x)
          for (int y = 0; y < nb; ++y) next.push_back({ka[x], kb[y]});
      }
    }
    level = std::move(next);
  }
  auto group = [&](const std::vector<std::pair<uint32_t, uint32_t>>& pairs, std::vector<uint32_t>& ptr,
                   std::vector<uint32_t>& col) {
    ptr.assign(nn + 1, 0);
    for (auto& p : pairs) ptr[p.first + 1]++;
    for (size_t i = 0; i < nn; ++i) ptr[i + 1] += ptr[i];
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    col.resize(pairs.size());
    for (auto& p : pairs) col[fill[p.first]++] = p.second;
  };
  group(adm, s.far_ptr, s.far_col);
  group(nearb, s.near_ptr, s.near_col);
  s.far_needed.assign(nn, 0);
  for (uint32_t b : s.far_col) s.far_needed[b] = 1;
  for (auto [a, b] : adm) s.admissible_pairs += (uint64_t)len(a) * len(b);
  const size_t nb = s.near_col.size();
  s.near_row.resize(nb);
  for (size_t a = 0; a < nn; ++a)
    for (uint32_t q = s.near_ptr[a]; q < s.near_ptr[a + 1]; ++q) s.near_row[q] = (uint32_t)a;
  s.near_off.assign(nb, 0);
  uint64_t total = 0;
  for (size_t q = 0; q < nb; ++q)
    if (s.near_row[q] <= s.near_col[q]) {
      s.near_off[q] = total;
      total += len(s.near_row[q]) * len(s.near_col[q]);
    }
  for (size_t q = 0; q < nb; ++q)
    if (s.near_row[q] > s.near_col[q]) {
      s.near_off[q] = total;
      total += len(s.near_row[q]) * len(s.near_col[q]);
    }
  s.near_total = total;
  for (auto [a, b] : nearb) s.near_entries += a == b ? len(a) * (len(a) - 1) : len(a) * len(b);
  s.chain_ptr.push_back(0);
  for (uint32_t l : t.leaves) {
    size_t s0 = s.chain.size();
    for (uint32_t x = l;; x = t.parent[x]) {
      s.chain.push_back(x);
      if (x == 0) break;
    }
    std::reverse(s.chain.begin() + s0, s.chain.end());
    s.chain_ptr.push_back((uint32_t)s.chain.size());
  }
  return s;
}

}  // namespace ddgrs
