// Continuous collision detection on the device (see ccd.hpp).  Every FP64 expression is written
// in the reference's order and compiled without contraction, so distances, advances and boxes
// round as there:
//   crates/ddg-repulsive-surfaces/src/ccd.rs            (windows, box trees, advance, band)
//   crates/ddg-repulsive-curves/src/ccd.rs:236-320      (point-segment, segment-segment and
//                                                        point-triangle distances)
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "ccd.hpp"
#include "ccd_geom.hpp"

#pragma clang fp contract(off)

namespace ddgrs {

namespace {

constexpr double kMargin = 1.0 + 1e-12;  // ccd.rs:498 MARGIN
constexpr int kQueryThreads = 64;
constexpr uint32_t kIndexBits = 26;      // a stack entry: level << 26 | node

__device__ __forceinline__ bool overlaps(const CcdBox& a, const CcdBox& b) {
  for (int k = 0; k < 3; ++k)
    if (!(a.lo[k] <= b.hi[k] && b.lo[k] <= a.hi[k])) return false;
  return true;
}

__device__ __forceinline__ CcdBox unite(const CcdBox& a, const CcdBox& b) {
  CcdBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = fmin(a.lo[k], b.lo[k]);
    r.hi[k] = fmax(a.hi[k], b.hi[k]);
  }
  return r;
}

__device__ __forceinline__ CcdBox empty_box() {
  CcdBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = INFINITY;
    r.hi[k] = -INFINITY;
  }
  return r;
}

__device__ __forceinline__ CcdBox inflate(const CcdBox& b, double h) {
  CcdBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = b.lo[k] - h;
    r.hi[k] = b.hi[k] + h;
  }
  return r;
}

// ccd.rs:480
__device__ double lipschitz(const V* ua, int na, const V* ub, int nb) {
  const double n = (double)(na + nb);
  V mean{0.0, 0.0, 0.0};
  for (int i = 0; i < na; ++i) mean = vadd(mean, ua[i]);
  for (int i = 0; i < nb; ++i) mean = vadd(mean, ub[i]);
  mean = vscale(mean, 1.0 / n);
  double l0a = 0.0, l0b = 0.0, lma = 0.0, lmb = 0.0;
  const V zero{0.0, 0.0, 0.0};
  for (int i = 0; i < na; ++i) l0a = fmax(l0a, vnorm(vsub(ua[i], zero)));
  for (int i = 0; i < nb; ++i) l0b = fmax(l0b, vnorm(vsub(ub[i], zero)));
  for (int i = 0; i < na; ++i) lma = fmax(lma, vnorm(vsub(ua[i], mean)));
  for (int i = 0; i < nb; ++i) lmb = fmax(lmb, vnorm(vsub(ub[i], mean)));
  return fmin(l0a + l0b, lma + lmb) * (1.0 + 1e-12);
}

struct Tally {
  uint32_t iterations = 0, pairs = 0, close = 0, touching = 0, exempt = 0;
};

struct Best {
  bool has = false;
  double tau = 0.0;
  uint32_t kind = 0, a = 0, b = 0;
  // the reference's total order: τ, then the primitive (vertex-triangle before edge-edge, ids)
  __device__ void offer(double t, uint32_t k, uint32_t x, uint32_t y) {
    bool take = !has;
    if (has) {
      if (t != tau) take = t < tau;
      else if (k != kind) take = k < kind;
      else if (x != a) take = x < a;
      else take = y < b;
    }
    if (take) {
      has = true;
      tau = t;
      kind = k;
      a = x;
      b = y;
    }
  }
  __device__ double horizon(double hi) const { return has ? fmin(tau, hi) : hi; }
};

struct Motion {
  const double *start, *vel, *speed;
  double vscale;
  const uint32_t *faces, *edges;
  const uint8_t* band;
  double gap_abs, close_factor;
  uint32_t max_it;
  double lo, hi;
};

// ccd.rs:529 advance
template <class Dist, class LRef>
__device__ bool advance(const Dist& dist, double l_cheap, const LRef& l_refined, const Motion& M, double lo, double hi,
                        Tally& t, double* out) {
  const double d0 = dist(0.0);
  t.pairs += 1;
  t.iterations += 1;
  const double gap_abs = M.gap_abs;
  if (lo == 0.0) {
    if (d0 <= 1e-12 * gap_abs) {
      t.touching += 1;
      return false;
    }
    if (d0 < gap_abs) t.close += 1;
  } else if (d0 <= 1e-12 * gap_abs) {
    return false;
  }
  const double target = d0 > gap_abs ? gap_abs : M.close_factor * d0;
  const double slack = 1e-9 * d0;
  double l = l_cheap;
  bool refined = false;
  double tau = lo;
  double d = lo == 0.0 ? d0 : dist(lo);
  if (lo != 0.0) t.iterations += 1;
  for (uint32_t it = 0; it < M.max_it; ++it) {
    if (isnan(d) || d - target <= slack) {
      *out = tau;
      return true;
    }
    double next = tau + (d - target) / l;
    if (next >= hi && !refined) {
      l = fmin(l, l_refined());
      refined = true;
      next = tau + (d - target) / l;
    }
    if (next >= hi || isnan(next)) return false;
    tau = next;
    d = dist(tau);
    t.iterations += 1;
  }
  *out = tau;
  return true;
}

struct TreeView {
  const CcdBox *nodes, *sboxes;
  const uint32_t* order;
  uint32_t n, leaf;
  CcdLevels lv;
};

// ccd.rs:359 query: f(element) for every element whose box overlaps q, in the reference's order.
template <class F>
__device__ void walk(const TreeView& T, const CcdBox& q, F&& f) {
  uint32_t stack[kCcdMaxLevels + 4];
  stack[0] = (T.lv.nlev - 1) << kIndexBits;
  int sp = 1;
  const uint32_t mask = (1u << kIndexBits) - 1;
  while (sp > 0) {
    --sp;
    const uint32_t k = stack[sp] >> kIndexBits, i = stack[sp] & mask;
    if (!overlaps(T.nodes[T.lv.off[k] + i], q)) continue;
    if (k == 0) {
      const uint32_t a = i * T.leaf, b = min(a + T.leaf, T.n);
      for (uint32_t p = a; p < b; ++p)
        if (overlaps(T.sboxes[p], q)) f(T.order[p]);
    } else {
      stack[sp++] = ((k - 1) << kIndexBits) | (2 * i);
      if (2 * i + 1 < T.lv.cnt[k - 1]) stack[sp++] = ((k - 1) << kIndexBits) | (2 * i + 1);
    }
  }
}

__device__ __forceinline__ void store_query(CcdQuery* o, const Best& b, const Tally& t, uint32_t kind) {
  CcdQuery q;
  q.tau = b.has ? b.tau : 0.0;
  q.flags = (b.has ? 1u : 0u) | (kind << 1);
  q.a = b.a;
  q.b = b.b;
  q.iterations = t.iterations;
  q.pairs = t.pairs;
  q.close = t.close;
  q.touching = t.touching;
  q.exempt = t.exempt;
  *o = q;
}

// ---- kernels ---------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void ccd_init_kernel(unsigned long long* scal) {
  if (threadIdx.x < 2) scal[threadIdx.x] = 0ull;
}

// speed_v = |vscale·vel_v| and the largest speed (non-negative doubles order like their bits).
__global__ __launch_bounds__(256) void ccd_speed_kernel(uint32_t nv, const double* __restrict__ vel, double vscale,
                                                        double* __restrict__ speed, unsigned long long* scal) {
  __shared__ double sh[256];
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (v < nv) {
    s = vnorm(vel3(vel, vscale, v));
    speed[v] = s;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(&scal[0], (unsigned long long)__double_as_longlong(sh[0]));
}

// Swept boxes of the vertices on [lo, hi] (hull of the two end positions).
__global__ __launch_bounds__(256) void ccd_vbox_kernel(uint32_t nv, const double* __restrict__ start,
                                                       const double* __restrict__ vel, double vscale, double lo,
                                                       double hi, CcdBox* __restrict__ vbox) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const V p = load3(start, v), u = vel3(vel, vscale, v);
  const V a = at(p, u, lo), b = at(p, u, hi);
  CcdBox r;
  const double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = fmin(fmin(INFINITY, av[k]), bv[k]);
    r.hi[k] = fmax(fmax(-INFINITY, av[k]), bv[k]);
  }
  vbox[v] = r;
}

// Face (3 corners) or edge (2) boxes as unions of vertex boxes, and the same inflated by h.
__global__ __launch_bounds__(256) void ccd_union_kernel(uint32_t n, int arity, const uint32_t* __restrict__ idx,
                                                        const CcdBox* __restrict__ vbox, double h,
                                                        CcdBox* __restrict__ box, CcdBox* __restrict__ inflated) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  CcdBox r = vbox[idx[(size_t)arity * i]];
  for (int c = 1; c < arity; ++c) r = unite(r, vbox[idx[(size_t)arity * i + c]]);
  box[i] = r;
  inflated[i] = inflate(r, h);
}

// Boxes of the faces at fixed positions (hull of the corners).
__global__ __launch_bounds__(256) void ccd_face_box_kernel(uint32_t nf, const uint32_t* __restrict__ faces,
                                                           const double* __restrict__ pos, CcdBox* __restrict__ box) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf) return;
  CcdBox r = empty_box();
  for (int c = 0; c < 3; ++c) {
    const V p = load3(pos, faces[3 * (size_t)f + c]);
    const double pv[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) {
      r.lo[k] = fmin(r.lo[k], pv[k]);
      r.hi[k] = fmax(r.hi[k], pv[k]);
    }
  }
  box[f] = r;
}

// Per workgroup the bounds of the box centres: part[6 b + k] = min, part[6 b + 3 + k] = max.
__global__ __launch_bounds__(256) void ccd_bounds_kernel(uint32_t n, const CcdBox* __restrict__ box,
                                                         double* __restrict__ part) {
  __shared__ double sh[6][256];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  for (int k = 0; k < 3; ++k) {
    double c = 0.0;
    const bool in = i < n;
    if (in) c = 0.5 * (box[i].lo[k] + box[i].hi[k]);
    sh[k][threadIdx.x] = in ? c : INFINITY;
    sh[3 + k][threadIdx.x] = in ? c : -INFINITY;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) {
        sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        sh[3 + k][threadIdx.x] = fmax(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void ccd_bounds_final_kernel(uint32_t nb, const double* __restrict__ part,
                                                              double* __restrict__ bounds) {
  const int k = threadIdx.x;
  if (k >= 6) return;
  double r = k < 3 ? INFINITY : -INFINITY;
  for (uint32_t b = 0; b < nb; ++b) r = k < 3 ? fmin(r, part[6 * (size_t)b + k]) : fmax(r, part[6 * (size_t)b + k]);
  bounds[k] = r;
}

__device__ __forceinline__ unsigned long long split3(unsigned long long x) {
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x001f00000000ffffull;
  x = (x | x << 16) & 0x001f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// Morton keys of the box centres quantized to 21 bits per axis (ccd.rs:313), with the indices.
__global__ __launch_bounds__(256) void ccd_keys_kernel(uint32_t n, const CcdBox* __restrict__ box,
                                                       const double* __restrict__ bounds,
                                                       unsigned long long* __restrict__ key, uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  unsigned long long code = 0;
  for (int k = 0; k < 3; ++k) {
    const double c = 0.5 * (box[i].lo[k] + box[i].hi[k]);
    const double w = bounds[3 + k] - bounds[k];
    unsigned long long q = 0;
    if (w > 0.0) {
      const double x = ((c - bounds[k]) / w) * 2097151.0;
      q = x > 0.0 ? (unsigned long long)x : 0ull;  // Rust's saturating cast (x <= 2097151)
    }
    code |= split3(q) << k;
  }
  key[i] = code;
  idx[i] = i;
}

__global__ __launch_bounds__(256) void ccd_gather_kernel(uint32_t n, const uint32_t* __restrict__ order,
                                                         const CcdBox* __restrict__ box, CcdBox* __restrict__ sboxes) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) sboxes[i] = box[order[i]];
}

// Level 0: unions of blocks of `leaf` sorted boxes; above: unions of pairs (ccd.rs:330-347).
__global__ __launch_bounds__(256) void ccd_level_kernel(uint32_t count, uint32_t width, uint32_t n_below,
                                                        const CcdBox* __restrict__ below, CcdBox* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  CcdBox r = empty_box();
  const uint32_t a = i * width, b = min(a + width, n_below);
  for (uint32_t p = a; p < b; ++p) r = unite(r, below[p]);
  out[i] = r;
}

// Vertex against triangle (ccd.rs:930-975): one thread per vertex.
__global__ __launch_bounds__(kQueryThreads) void ccd_vertex_kernel(uint32_t nv, Motion M, TreeView T,
                                                                   const CcdBox* __restrict__ vbox, double h,
                                                                   CcdQuery* __restrict__ out) {
  const uint32_t v = blockIdx.x * kQueryThreads + threadIdx.x;
  if (v >= nv) return;
  Best best;
  Tally tally;
  const CcdBox q = inflate(vbox[v], h);
  const bool v_in = M.band && M.band[v];
  const V p = load3(M.start, v), u = vel3(M.vel, M.vscale, v);
  const double sv = M.speed[v];
  walk(T, q, [&](uint32_t f) {
    const uint32_t t0 = M.faces[3 * (size_t)f], t1 = M.faces[3 * (size_t)f + 1], t2 = M.faces[3 * (size_t)f + 2];
    if (t0 == v || t1 == v || t2 == v) return;
    if (v_in && (M.band[t0] || M.band[t1] || M.band[t2])) {
      tally.exempt += 1;
      return;
    }
    const V tp[3] = {at(load3(M.start, t0), vel3(M.vel, M.vscale, t0), 0.0),
                     at(load3(M.start, t1), vel3(M.vel, M.vscale, t1), 0.0),
                     at(load3(M.start, t2), vel3(M.vel, M.vscale, t2), 0.0)};
    const V tu[3] = {vel3(M.vel, M.vscale, t0), vel3(M.vel, M.vscale, t1), vel3(M.vel, M.vscale, t2)};
    const double ts = fmax(fmax(M.speed[t0], M.speed[t1]), M.speed[t2]);
    const double lc = (sv + ts) * kMargin;
    auto dist = [&](double t) {
      return point_triangle_distance(at(p, u, t), at(tp[0], tu[0], t), at(tp[1], tu[1], t), at(tp[2], tu[2], t));
    };
    auto lref = [&]() { return lipschitz(&u, 1, tu, 3); };
    double tau = 0.0;
    if (advance(dist, lc, lref, M, M.lo, best.horizon(M.hi), tally, &tau)) best.offer(tau, 0, v, f);
  });
  store_query(out + v, best, tally, 0);
}

// Edge against edge (ccd.rs:978-1036): one thread per edge, partners with a larger id.
__global__ __launch_bounds__(kQueryThreads) void ccd_edge_kernel(uint32_t ne, Motion M, TreeView T,
                                                                 const CcdBox* __restrict__ ebox, double h,
                                                                 CcdQuery* __restrict__ out) {
  const uint32_t e = blockIdx.x * kQueryThreads + threadIdx.x;
  if (e >= ne) return;
  Best best;
  Tally tally;
  const uint32_t a = M.edges[2 * (size_t)e], b = M.edges[2 * (size_t)e + 1];
  const CcdBox q = inflate(ebox[e], h);
  const bool ab_in = M.band && (M.band[a] || M.band[b]);
  const V pp[2] = {load3(M.start, a), load3(M.start, b)};
  const V uu[2] = {vel3(M.vel, M.vscale, a), vel3(M.vel, M.vscale, b)};
  const double sab = fmax(M.speed[a], M.speed[b]);
  walk(T, q, [&](uint32_t e2) {
    if (e2 <= e) return;
    const uint32_t c = M.edges[2 * (size_t)e2], d = M.edges[2 * (size_t)e2 + 1];
    if (c == a || c == b || d == a || d == b) return;
    if (ab_in && (M.band[c] || M.band[d])) {
      tally.exempt += 1;
      return;
    }
    const V qq[2] = {load3(M.start, c), load3(M.start, d)};
    const V ww[2] = {vel3(M.vel, M.vscale, c), vel3(M.vel, M.vscale, d)};
    const double lc = (sab + fmax(M.speed[c], M.speed[d])) * kMargin;
    auto dist = [&](double t) {
      return segment_segment_distance(at(pp[0], uu[0], t), at(pp[1], uu[1], t), at(qq[0], ww[0], t),
                                      at(qq[1], ww[1], t));
    };
    auto lref = [&]() { return lipschitz(uu, 2, ww, 2); };
    double tau = 0.0;
    if (advance(dist, lc, lref, M, M.lo, best.horizon(M.hi), tally, &tau)) best.offer(tau, 1, e, e2);
  });
  store_query(out + e, best, tally, 1);
}

// The earliest contact over all queries and the summed counts (one workgroup; the order of the
// combination does not matter: a minimum in a total order and integer sums).
__global__ __launch_bounds__(1024) void ccd_reduce_kernel(uint32_t n, const CcdQuery* __restrict__ q,
                                                          CcdSummary* __restrict__ out) {
  __shared__ double s_tau[1024];
  __shared__ uint32_t s_flags[1024], s_a[1024], s_b[1024];
  __shared__ unsigned long long s_cnt[5][1024];
  Best best;
  unsigned long long cnt[5] = {0, 0, 0, 0, 0};
  for (uint32_t i = threadIdx.x; i < n; i += 1024) {
    const CcdQuery x = q[i];
    if (x.flags & 1u) best.offer(x.tau, x.flags >> 1, x.a, x.b);
    cnt[0] += x.iterations;
    cnt[1] += x.pairs;
    cnt[2] += x.close;
    cnt[3] += x.touching;
    cnt[4] += x.exempt;
  }
  const int t = threadIdx.x;
  s_tau[t] = best.tau;
  s_flags[t] = (best.has ? 1u : 0u) | (best.kind << 1);
  s_a[t] = best.a;
  s_b[t] = best.b;
  for (int k = 0; k < 5; ++k) s_cnt[k][t] = cnt[k];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      Best x;
      x.has = s_flags[t] & 1u;
      x.tau = s_tau[t];
      x.kind = s_flags[t] >> 1;
      x.a = s_a[t];
      x.b = s_b[t];
      if (s_flags[t + w] & 1u) x.offer(s_tau[t + w], s_flags[t + w] >> 1, s_a[t + w], s_b[t + w]);
      s_tau[t] = x.tau;
      s_flags[t] = (x.has ? 1u : 0u) | (x.kind << 1);
      s_a[t] = x.a;
      s_b[t] = x.b;
      for (int k = 0; k < 5; ++k) s_cnt[k][t] += s_cnt[k][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    CcdSummary r;
    r.tau = s_tau[0];
    r.flags = s_flags[0];
    r.a = s_a[0];
    r.b = s_b[0];
    r.iterations = s_cnt[0][0];
    r.pairs = s_cnt[1][0];
    r.close = s_cnt[2][0];
    r.touching = s_cnt[3][0];
    r.exempt = s_cnt[4][0];
    *out = r;
  }
}

// intersecting_pairs (ccd.rs:1229): one thread per face f, partners g > f from the face tree.
__global__ __launch_bounds__(kQueryThreads) void ccd_pairs_kernel(uint32_t nf, const uint32_t* __restrict__ faces,
                                                                  const double* __restrict__ pos, TreeView T,
                                                                  const CcdBox* __restrict__ box,
                                                                  uint32_t* __restrict__ pairs, uint64_t cap,
                                                                  unsigned long long* count) {
  const uint32_t f = blockIdx.x * kQueryThreads + threadIdx.x;
  if (f >= nf) return;
  const uint32_t a[3] = {faces[3 * (size_t)f], faces[3 * (size_t)f + 1], faces[3 * (size_t)f + 2]};
  const V tf[3] = {load3(pos, a[0]), load3(pos, a[1]), load3(pos, a[2])};
  walk(T, box[f], [&](uint32_t g) {
    if (g <= f) return;
    const uint32_t b[3] = {faces[3 * (size_t)g], faces[3 * (size_t)g + 1], faces[3 * (size_t)g + 2]};
    int common[3], nc = 0;
    for (int i = 0; i < 3; ++i)
      if (b[0] == a[i] || b[1] == a[i] || b[2] == a[i]) common[nc++] = i;
    const V tg[3] = {load3(pos, b[0]), load3(pos, b[1]), load3(pos, b[2])};
    bool hit = false;
    if (nc == 0) {
      hit = triangles_intersect(tf, tg);
    } else if (nc == 1) {
      const int i = common[0];
      int j = 0;
      while (b[j] != a[i]) ++j;
      hit = segment_meets_triangle(tf[(i + 1) % 3], tf[(i + 2) % 3], tg[0], tg[1], tg[2]) ||
            segment_meets_triangle(tg[(j + 1) % 3], tg[(j + 2) % 3], tf[0], tf[1], tf[2]);
    }
    if (hit) {
      const unsigned long long k = atomicAdd(count, 1ull);
      if (k < cap) {
        pairs[2 * k] = f;
        pairs[2 * k + 1] = g;
      }
    }
  });
}

// The band: the vertices of the intersecting faces (core), widened by one ring along the edges.
__global__ __launch_bounds__(256) void ccd_core_kernel(uint32_t npairs, const uint32_t* __restrict__ pairs,
                                                       const uint32_t* __restrict__ faces, uint8_t* __restrict__ core) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 2 * npairs) return;
  const uint32_t f = pairs[i];
  for (int c = 0; c < 3; ++c) core[faces[3 * (size_t)f + c]] = 1;
}

__global__ __launch_bounds__(256) void ccd_widen_kernel(uint32_t ne, const uint32_t* __restrict__ edges,
                                                        const uint8_t* __restrict__ core, uint8_t* __restrict__ band) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const uint32_t a = edges[2 * (size_t)e], b = edges[2 * (size_t)e + 1];
  if (core[a] || core[b]) {
    band[a] = 1;
    band[b] = 1;
  }
}

__global__ __launch_bounds__(256) void ccd_fill_u8_kernel(uint32_t n, uint8_t* __restrict__ x, uint8_t v) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ __launch_bounds__(256) void ccd_copy_u8_kernel(uint32_t n, const uint8_t* __restrict__ a,
                                                          uint8_t* __restrict__ b) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

TreeView view(const CcdTreeDev& t) {
  return TreeView{t.nodes.ptr, t.sboxes.ptr, t.order.ptr, t.n, t.leaf, t.lv};
}

}  // namespace

void check_ccd_options(const CcdOpts& o) {
  if (!(std::isfinite(o.gap) && o.gap > 0.0)) invalid("ccd.gap", std::to_string(o.gap), "finite and positive");
  if (!std::isnan(o.reference_length) && !(std::isfinite(o.reference_length) && o.reference_length > 0.0))
    invalid("ccd.reference_length", std::to_string(o.reference_length), "finite and positive (NaN: none)");
  if (o.max_iterations == 0) invalid("ccd.max_iterations", "0", "at least 1");
  if (!(o.close_factor >= 0.0 && o.close_factor < 1.0))
    invalid("ccd.close_factor", std::to_string(o.close_factor), "in [0, 1)");
  if (o.leaf_size == 0) invalid("ccd.leaf_size", "0", "at least 1");
}

double mean_edge_length_host(const std::vector<uint32_t>& edges, const double* pos) {
  const size_t ne = edges.size() / 2;
  double s = 0.0;
  for (size_t e = 0; e < ne; ++e) {
    const double* a = pos + 3 * (size_t)edges[2 * e];
    const double* b = pos + 3 * (size_t)edges[2 * e + 1];
    const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    s += std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  }
  return s / (double)ne;
}

void CcdTreeDev::build(ddgrs_ctx* ctx, const CcdBox* boxes, uint32_t n_, uint32_t leaf_) {
  n = n_;
  leaf = leaf_;
  if (n == 0) throw Error(DDGRS_GPU, "ccd", "a box tree over no elements");
  // level sizes (ccd.rs:330-347)
  lv = CcdLevels{};
  uint32_t c = (uint32_t)((n + (uint64_t)leaf - 1) / leaf), off = 0;
  if (c >= (1u << kIndexBits)) throw Error(DDGRS_GPU, "ccd", "too many leaf blocks for the box tree");
  for (;;) {
    if (lv.nlev >= kCcdMaxLevels) throw Error(DDGRS_GPU, "ccd", "box tree too deep");
    lv.off[lv.nlev] = off;
    lv.cnt[lv.nlev] = c;
    ++lv.nlev;
    off += c;
    if (c == 1) break;
    c = (c + 1) / 2;
  }
  const unsigned nb = blocks_for(n, 256);
  part.reserve(6 * (size_t)nb);
  bounds.reserve(6);
  key_in.reserve(n);
  key_out.reserve(n);
  idx_in.reserve(n);
  order.reserve(n);
  sboxes.reserve(n);
  nodes.reserve(off);
  launch(ctx, "ccd_bounds", ccd_bounds_kernel, dim3(nb), dim3(256), 0, n, boxes, part.ptr);
  launch(ctx, "ccd_bounds_final", ccd_bounds_final_kernel, dim3(1), dim3(64), 0, (uint32_t)nb, (const double*)part.ptr,
         bounds.ptr);
  launch(ctx, "ccd_keys", ccd_keys_kernel, dim3(nb), dim3(256), 0, n, boxes, (const double*)bounds.ptr, key_in.ptr,
         idx_in.ptr);
  size_t tb = 0;
  HIPCHECK(rocprim::radix_sort_pairs(nullptr, tb, key_in.ptr, key_out.ptr, idx_in.ptr, order.ptr, n, 0, 63, ctx->stream));
  temp.reserve(std::max<size_t>(tb, 1));
  HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, key_in.ptr, key_out.ptr, idx_in.ptr, order.ptr, n, 0, 63, ctx->stream));
  launch(ctx, "ccd_gather", ccd_gather_kernel, dim3(nb), dim3(256), 0, n, (const uint32_t*)order.ptr, boxes, sboxes.ptr);
  launch(ctx, "ccd_level", ccd_level_kernel, dim3(blocks_for(lv.cnt[0], 256)), dim3(256), 0, lv.cnt[0], leaf, n,
         (const CcdBox*)sboxes.ptr, nodes.ptr);
  for (uint32_t k = 1; k < lv.nlev; ++k)
    launch(ctx, "ccd_level", ccd_level_kernel, dim3(blocks_for(lv.cnt[k], 256)), dim3(256), 0, lv.cnt[k], 2u,
           lv.cnt[k - 1], (const CcdBox*)(nodes.ptr + lv.off[k - 1]), nodes.ptr + lv.off[k]);
}

void CcdDev::set_mesh(ddgrs_ctx* ctx, const std::vector<uint32_t>& faces_h, uint32_t nv_) {
  nv = nv_;
  nf = (uint32_t)(faces_h.size() / 3);
  std::vector<uint64_t> keys;
  keys.reserve(faces_h.size());
  for (size_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) {
      const uint64_t a = faces_h[3 * f + c], b = faces_h[3 * f + (c + 1) % 3];
      keys.push_back(std::min(a, b) << 32 | std::max(a, b));
    }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  ne = (uint32_t)keys.size();
  edges_h.resize(2 * (size_t)ne);
  for (size_t e = 0; e < ne; ++e) {
    edges_h[2 * e] = (uint32_t)(keys[e] >> 32);
    edges_h[2 * e + 1] = (uint32_t)keys[e];
  }
  upload(ctx, edges, edges_h);
}

CcdResult CcdDev::first_contact(ddgrs_ctx* ctx, const uint32_t* faces, const double* start, const double* start_h,
                                const double* vel, double vscale, double tau_end, const CcdOpts& o,
                                const uint8_t* band_dev) {
  check_ccd_options(o);
  if (!(std::isfinite(tau_end) && tau_end >= 0.0))
    invalid("tau_end", std::to_string(tau_end), "finite and non-negative");
  const double mel = mean_edge_length_host(edges_h, start_h);
  const double ell = std::isnan(o.reference_length) ? mel : o.reference_length;
  CcdResult rep;
  rep.gap_abs = o.gap * ell;
  rep.tau_max = tau_end;
  speed.reserve(nv);
  scal.reserve(2);
  vbox.reserve(nv);
  fbox.reserve(nf);
  fq.reserve(nf);
  ebox.reserve(ne);
  eq.reserve(ne);
  out.reserve((size_t)nv + ne);
  sum.reserve(1);
  launch(ctx, "ccd_init", ccd_init_kernel, dim3(1), dim3(64), 0, scal.ptr);
  launch(ctx, "ccd_speed", ccd_speed_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, vel, vscale, speed.ptr,
         scal.ptr);
  unsigned long long vbits = 0;
  download(ctx, &vbits, (const unsigned long long*)scal.ptr, 1);
  ctx->sync();
  double vmax;
  std::memcpy(&vmax, &vbits, sizeof vmax);
  Motion M{start, vel, speed.ptr, vscale, faces, edges.ptr, band_dev, rep.gap_abs, o.close_factor, o.max_iterations,
           0.0, 0.0};
  const double t1 = vmax > 0.0 ? mel / vmax : INFINITY;
  const double h = 0.5 * rep.gap_abs * (1.0 + 1e-6);
  double lo = 0.0, hi = std::fmin(t1, tau_end);
  for (;;) {
    if (vmax == 0.0) hi = 0.0;
    rep.windows += 1;
    // ccd.rs:886 window
    launch(ctx, "ccd_vbox", ccd_vbox_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, start, vel, vscale, lo, hi,
           vbox.ptr);
    launch(ctx, "ccd_union", ccd_union_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, 3, faces,
           (const CcdBox*)vbox.ptr, h, fbox.ptr, fq.ptr);
    launch(ctx, "ccd_union", ccd_union_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, ne, 2,
           (const uint32_t*)edges.ptr, (const CcdBox*)vbox.ptr, h, ebox.ptr, eq.ptr);
    ftree.build(ctx, fq.ptr, nf, o.leaf_size);
    etree.build(ctx, eq.ptr, ne, o.leaf_size);
    M.lo = lo;
    M.hi = hi;
    launch(ctx, "ccd_vertex", ccd_vertex_kernel, dim3(blocks_for(nv, kQueryThreads)), dim3(kQueryThreads), 0, nv, M,
           view(ftree), (const CcdBox*)vbox.ptr, h, out.ptr);
    launch(ctx, "ccd_edge", ccd_edge_kernel, dim3(blocks_for(ne, kQueryThreads)), dim3(kQueryThreads), 0, ne, M,
           view(etree), (const CcdBox*)ebox.ptr, h, out.ptr + nv);
    launch(ctx, "ccd_reduce", ccd_reduce_kernel, dim3(1), dim3(1024), 0, nv + ne, (const CcdQuery*)out.ptr, sum.ptr);
    CcdSummary s;
    download(ctx, &s, (const CcdSummary*)sum.ptr, 1);
    ctx->sync();
    rep.iterations += s.iterations;
    rep.pairs_tested += s.pairs;
    if (lo == 0.0) {
      rep.close_pairs = s.close;
      rep.touching_pairs = s.touching;
      rep.exempt_pairs = s.exempt;
    }
    if (s.flags & 1u) {
      rep.tau_max = s.tau;
      rep.contact_tau = s.tau;
      rep.has_contact = true;
      rep.kind = (int)(s.flags >> 1);
      rep.a = s.a;
      rep.b = s.b;
      break;
    }
    if (hi >= tau_end || vmax == 0.0) break;
    lo = hi;
    hi = std::fmin(std::fmax(2.0 * hi, hi + t1), tau_end);
    if (hi <= lo) break;
  }
  return rep;
}

size_t CcdDev::find_intersecting_pairs(ddgrs_ctx* ctx, const uint32_t* faces, const double* pos) {
  fbox.reserve(nf);
  scal.reserve(2);
  launch(ctx, "ccd_face_box", ccd_face_box_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, faces, pos, fbox.ptr);
  ftree.build(ctx, fbox.ptr, nf, 8);
  if (pair_cap == 0) pair_cap = 1024;
  for (;;) {
    pairs.reserve(2 * pair_cap);
    launch(ctx, "ccd_init", ccd_init_kernel, dim3(1), dim3(64), 0, scal.ptr);
    launch(ctx, "ccd_pairs", ccd_pairs_kernel, dim3(blocks_for(nf, kQueryThreads)), dim3(kQueryThreads), 0, nf, faces,
           pos, view(ftree), (const CcdBox*)fbox.ptr, pairs.ptr, (uint64_t)pair_cap, scal.ptr + 1);
    unsigned long long cnt = 0;
    download(ctx, &cnt, (const unsigned long long*)scal.ptr + 1, 1);
    ctx->sync();
    if (cnt <= pair_cap) return (size_t)cnt;
    pair_cap = (size_t)cnt + cnt / 4;
  }
}

size_t CcdDev::band_flags(ddgrs_ctx* ctx, const uint32_t* faces, const double* pos) {
  const size_t np = find_intersecting_pairs(ctx, faces, pos);
  core.reserve(nv);
  band.reserve(nv);
  launch(ctx, "ccd_fill", ccd_fill_u8_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, core.ptr, (uint8_t)0);
  launch(ctx, "ccd_core", ccd_core_kernel, dim3(blocks_for(2 * (int64_t)np, 256)), dim3(256), 0, (uint32_t)np,
         (const uint32_t*)pairs.ptr, faces, core.ptr);
  launch(ctx, "ccd_copy", ccd_copy_u8_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint8_t*)core.ptr,
         band.ptr);
  if (np)
    launch(ctx, "ccd_widen", ccd_widen_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, ne, (const uint32_t*)edges.ptr,
           (const uint8_t*)core.ptr, band.ptr);
  return np;
}

}  // namespace ddgrs
