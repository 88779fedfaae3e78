// One remeshing pass on the device, following
//   crates/ddg-repulsive-surfaces/src/remesh/mod.rs   (remesh :235, collapse_short :338,
//                                                      delaunay_flips :380, smooth :424)
//   crates/ddg-repulsive-surfaces/src/remesh/mesh.rs  (EditMesh :224, split :592, collapse :679,
//                                                      flip :888, fan :475, finish :991)
//   crates/ddg-repulsive-surfaces/src/remesh/geom.rs  (folds, violates_delaunay, circumcenter)
//   crates/ddg-repulsive-surfaces/src/ccd.rs:1448     (CcdCollapseGuard)
//
// Layout in HBM: positions (FP64, 3 per vertex slot), faces (3 u32 per face slot; slots are never
// reused: a split appends, a collapse marks two dead), per face slot its parent and a live flag,
// per vertex slot a list of its live faces (at most kVF), live / pinned flags, the input pin it
// carries and its vertex-map entry (offset and length into a pool of (old vertex, weight)).
//
// Which edits happen, and in which order, is the reference's:
// * splits: the input's long edges in lexicographic order (the longest within the growth budget
//   when they exceed it), each applied by the edit wave;
// * collapses: the short edges sorted by (length bits, min, max) on the device, merged by the edit
//   wave with a heap of the edges that collapses requeue — exactly the reference's binary
//   heap order, stale entries skipped by its own tests; the collapse guard is evaluated by the
//   whole workgroup over every live face (see guard_scan);
// * flips, per round: the reference pops every edge in lexicographic order, then the edges its
//   flips requeue.  An edge's Delaunay status changes only when a flip replaces one of its faces,
//   and those edges are exactly the flip's quad edges, which the reference queues.  So the wide
//   kernel classifies every edge once, and the edit wave visits in lexicographic order only
//   the violating edges and the still-pending quad edges of earlier flips (each re-tested when
//   visited), then the requeued ones in FIFO order: the same flips in the same order;
// * smoothing: displacements for all vertices at once (Jacobi, as the reference); the guarded
//   moves, which the reference applies in vertex order, each seeing its lower neighbours' new
//   positions and its higher neighbours' old ones, are solved as the fixed point of that
//   triangular system: every vertex is re-evaluated from the previous iterate until nothing
//   changes (the fixed point is unique, so it is the sequential result), with the sequential
//   order itself as the fallback after kMaxSmoothIterations.
// FP64 throughout, every expression in the reference's order, compiled without contraction.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "ccd_geom.hpp"
#include "remesh.hpp"

#pragma clang fp contract(off)

namespace ddgrs {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kVF = 32;                    // live faces per vertex slot
constexpr int kNb = 2 * kVF;               // neighbour entries of a vertex (with repeats)
constexpr double kCosFoldover = 0.4161468365471424;  // geom.rs:32 COS_FOLDOVER_LIMIT
constexpr int kMaxSmoothIterations = 64;
constexpr int kGuardThreads = 256;

enum Refusal : int {
  R_OK = 0,
  R_NO_SUCH_EDGE = 1,
  R_BOTH_PINNED = 2,
  R_LINK = 3,
  R_FOLDOVER = 4,
  R_REPEATED = 5,
  R_NON_MANIFOLD = 6,
  R_BOUNDARY = 7,
  R_NOT_DELAUNAY = 8,
  R_EXISTING = 9,
  R_GUARD = 10,
  R_DEGENERATE = 11,
};

// capacity flags
enum : uint32_t { F_VALENCE = 1, F_POOL = 2, F_HEAP = 4, F_EDGE = 8, F_SLOTS = 16, F_FIFO = 32 };

struct Counters {
  uint32_t nv, nf, pool, flags;
  unsigned long long splits, collapses, flips, splits_deferred;
  unsigned long long refused_split[11], refused_collapse[11], refused_flip[11];
  unsigned long long smoothing_moves, smoothing_halved, smoothing_refused, smoothing_skipped;
  unsigned long long flip_cap_hits, changed;
};

// Device view of the mesh.
struct Mesh {
  double* pos;
  uint32_t* faces;
  uint8_t* falive;
  uint32_t* parent;
  uint32_t* vfn;
  uint32_t* vf;
  uint8_t* valive;
  uint8_t* pinned;
  uint32_t* pin_origin;
  uint32_t* src_off;
  uint32_t* src_len;
  uint32_t* pool_v;
  double* pool_w;
  uint32_t vcap, fcap, pcap;
  Counters* cnt;
};

// ---- geometry (geom.rs, geometry.rs:33) ------------------------------------------------------

__device__ __forceinline__ V P(const Mesh& m, uint32_t v) {
  return {m.pos[3 * (size_t)v], m.pos[3 * (size_t)v + 1], m.pos[3 * (size_t)v + 2]};
}
__device__ __forceinline__ void setP(const Mesh& m, uint32_t v, V x) {
  m.pos[3 * (size_t)v] = x.x;
  m.pos[3 * (size_t)v + 1] = x.y;
  m.pos[3 * (size_t)v + 2] = x.z;
}
__device__ __forceinline__ V tri_normal(V p0, V p1, V p2) { return vcross(vsub(p1, p0), vsub(p2, p0)); }
__device__ __forceinline__ bool folds(V before, V after) {
  const double a = vnorm(after);
  return !(a > 0.0 && isfinite(a)) || vdot(before, after) < kCosFoldover * vnorm(before) * a;
}
// mesh.rs:1051 `degenerate` with geometry.rs:33 `face_quantities`
__device__ bool degenerate(V p0, V p1, V p2) {
  const V n = vcross(vsub(p1, p0), vsub(p2, p0));
  const double len = vnorm(n);
  const double area = 0.5 * len;
  const double inv = 1.0 / len;
  const V u = vscale(n, inv);
  return !(area > 0.0 && isfinite(area) && isfinite(u.x) && isfinite(u.y) && isfinite(u.z));
}
// geom.rs:50
__device__ bool violates_delaunay(V p, V q, V c, V d) {
  const V uc = vsub(p, c), vc = vsub(q, c), ud = vsub(p, d), vd = vsub(q, d);
  const double dc = vdot(uc, vc), sc = vnorm(vcross(uc, vc));
  const double dd = vdot(ud, vd), sd = vnorm(vcross(ud, vd));
  return sc > 0.0 && sd > 0.0 && dc * sd + dd * sc < 0.0;
}
// geom.rs:64
__device__ V circumcenter(V p0, V p1, V p2) {
  const V e0 = vsub(p2, p1), e1 = vsub(p2, p0), e2 = vsub(p1, p0);
  const double a2 = vdot(e0, e0), b2 = vdot(e1, e1), c2 = vdot(e2, e2);
  const double w0 = a2 * (b2 + c2 - a2), w1 = b2 * (c2 + a2 - b2), w2 = c2 * (a2 + b2 - c2);
  const double s = w0 + w1 + w2;
  return vadd(vadd(vscale(p0, w0 / s), vscale(p1, w1 / s)), vscale(p2, w2 / s));
}
__device__ __forceinline__ V barycenter(V p0, V p1, V p2) { return vscale(vadd(vadd(p0, p1), p2), 1.0 / 3.0); }
__device__ __forceinline__ double angle_between(V u, V v) { return atan2(vnorm(vcross(u, v)), vdot(u, v)); }

// ---- mesh queries (mesh.rs) ----------------------------------------------------------------

struct Tri {
  uint32_t v[3];
};
__device__ __forceinline__ Tri F(const Mesh& m, uint32_t f) {
  return {{m.faces[3 * (size_t)f], m.faces[3 * (size_t)f + 1], m.faces[3 * (size_t)f + 2]}};
}
__device__ __forceinline__ void setF(const Mesh& m, uint32_t f, uint32_t a, uint32_t b, uint32_t c) {
  m.faces[3 * (size_t)f] = a;
  m.faces[3 * (size_t)f + 1] = b;
  m.faces[3 * (size_t)f + 2] = c;
}
__device__ __forceinline__ bool has(const Tri& t, uint32_t v) { return t.v[0] == v || t.v[1] == v || t.v[2] == v; }
__device__ __forceinline__ void flag(const Mesh& m, uint32_t f) { atomicOr(&m.cnt->flags, f); }

template <class T>
__device__ void isort(T* a, int n) {
  for (int i = 1; i < n; ++i) {
    const T x = a[i];
    int j = i - 1;
    while (j >= 0 && x < a[j]) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = x;
  }
}

__device__ __forceinline__ bool valid_vertex(const Mesh& m, uint32_t v) { return v < m.cnt->nv && m.valive[v]; }

// mesh.rs:381 `edge_faces`: live faces on {a, b}, ascending (at most 8 kept)
__device__ int edge_faces(const Mesh& m, uint32_t a, uint32_t b, uint32_t* out) {
  if (!valid_vertex(m, a) || !valid_vertex(m, b) || a == b) return 0;
  int n = 0;
  const uint32_t k = m.vfn[a];
  for (uint32_t i = 0; i < k; ++i) {
    const uint32_t f = m.vf[(size_t)a * kVF + i];
    if (has(F(m, f), b)) {
      if (n < 8)
        out[n++] = f;
      else
        flag(m, F_EDGE);
    }
  }
  isort(out, n);
  return n;
}

// mesh.rs:434 `neighbors`, ascending; returns the count
__device__ int neighbors(const Mesh& m, uint32_t v, uint32_t* out) {
  int n = 0;
  const uint32_t k = m.vfn[v];
  for (uint32_t i = 0; i < k; ++i) {
    const Tri t = F(m, m.vf[(size_t)v * kVF + i]);
    for (int c = 0; c < 3; ++c)
      if (t.v[c] != v) out[n++] = t.v[c];
  }
  isort(out, n);
  int u = 0;
  for (int i = 0; i < n; ++i)
    if (u == 0 || out[u - 1] != out[i]) out[u++] = out[i];
  return u;
}

// mesh.rs:449 `is_boundary_vertex`
__device__ bool is_boundary_vertex(const Mesh& m, uint32_t v) {
  uint32_t nb[kNb];
  int n = 0;
  const uint32_t k = m.vfn[v];
  for (uint32_t i = 0; i < k; ++i) {
    const Tri t = F(m, m.vf[(size_t)v * kVF + i]);
    for (int c = 0; c < 3; ++c)
      if (t.v[c] != v) nb[n++] = t.v[c];
  }
  isort(nb, n);
  int i = 0;
  while (i < n) {
    int j = i;
    while (j < n && nb[j] == nb[i]) ++j;
    if (j - i == 1) return true;
    i = j;
  }
  return false;
}

__device__ __forceinline__ void wedge(const Tri& t, uint32_t v, uint32_t* w0, uint32_t* w1) {
  int k = 0;
  for (int i = 0; i < 3; ++i)
    if (t.v[i] == v) {
      k = i;
      break;
    }
  *w0 = t.v[(k + 1) % 3];
  *w1 = t.v[(k + 2) % 3];
}
__device__ __forceinline__ bool traverses(const Tri& t, uint32_t a, uint32_t b) {
  for (int k = 0; k < 3; ++k)
    if (t.v[k] == a && t.v[(k + 1) % 3] == b) return true;
  return false;
}

// mesh.rs:475 `fan`: faces around v in rotational order with their relative orientation; false
// when the faces around v are not one fan.  The reference sorts (neighbour, face) pairs; here each
// neighbour's occurrences are counted over the (at most kVF) wedges directly, and the face after
// `cur` across `exit` is the lowest-numbered other face with `exit` in its wedge — the first of
// the reference's sorted pairs for that neighbour.
__device__ bool fan(const Mesh& m, uint32_t v, uint32_t* faces, bool* flips, int* count) {
  const uint32_t nfs = m.vfn[v];
  if (nfs == 0 || nfs > (uint32_t)kVF) return false;
  uint32_t fl[kVF], wa[kVF], wb[kVF];
  uint32_t fmin = kNone, imin = 0;
  for (uint32_t i = 0; i < nfs; ++i) {
    const uint32_t f = m.vf[(size_t)v * kVF + i];
    fl[i] = f;
    if (f < fmin) fmin = f, imin = i;
    wedge(F(m, f), v, &wa[i], &wb[i]);
  }
  // neighbours seen once are the ends of a boundary fan; more than twice: not one fan
  uint32_t nends = 0, end_f = kNone, end_w = kNone;
  for (uint32_t i = 0; i < nfs; ++i)
    for (int side = 0; side < 2; ++side) {
      const uint32_t w = side ? wb[i] : wa[i];
      uint32_t cnt = 0;
      for (uint32_t j = 0; j < nfs; ++j) cnt += (wa[j] == w ? 1u : 0u) + (wb[j] == w ? 1u : 0u);
      if (cnt > 2) return false;
      if (cnt == 1) {
        ++nends;
        if (fl[i] < end_f || (fl[i] == end_f && w < end_w)) end_f = fl[i], end_w = w;
      }
    }
  if (nends != 0 && nends != 2) return false;
  uint32_t cur, entry;
  if (nends == 2) {
    cur = end_f;
    entry = end_w;
  } else {
    cur = fmin;
    entry = wa[imin];
  }
  const uint32_t start = cur;
  int c = 0;
  faces[c] = start;
  flips[c] = false;
  ++c;
  bool cur_flip = false;
  for (;;) {
    uint32_t ci = 0;
    for (uint32_t i = 0; i < nfs; ++i)
      if (fl[i] == cur) ci = i;
    const uint32_t exit = wa[ci] == entry ? wb[ci] : wa[ci];
    uint32_t next = kNone;
    for (uint32_t j = 0; j < nfs; ++j)
      if (fl[j] != cur && (wa[j] == exit || wb[j] == exit) && fl[j] < next) next = fl[j];
    if (next == kNone || next == start) break;
    if ((uint32_t)c > nfs || c >= kVF) return false;
    const bool same = traverses(F(m, cur), v, exit) == traverses(F(m, next), v, exit);
    cur_flip ^= same;
    faces[c] = next;
    flips[c] = cur_flip;
    ++c;
    cur = next;
    entry = exit;
  }
  if ((uint32_t)c != nfs) return false;
  *count = c;
  return true;
}

__device__ __forceinline__ void sorted3(Tri t, uint32_t* s) {
  s[0] = t.v[0], s[1] = t.v[1], s[2] = t.v[2];
  isort(s, 3);
}

// mesh.rs:558 `vertex_ok`
__device__ bool vertex_ok(const Mesh& m, uint32_t v) {
  if (!m.valive[v]) return true;
  uint32_t fs[kVF];
  bool fl[kVF];
  int c;
  if (!fan(m, v, fs, fl, &c)) return false;
  const uint32_t k = m.vfn[v];
  for (uint32_t i = 0; i < k; ++i) {
    uint32_t si[3];
    sorted3(F(m, m.vf[(size_t)v * kVF + i]), si);
    for (uint32_t j = i + 1; j < k; ++j) {
      uint32_t sj[3];
      sorted3(F(m, m.vf[(size_t)v * kVF + j]), sj);
      if (si[0] == sj[0] && si[1] == sj[1] && si[2] == sj[2]) return false;
    }
  }
  return true;
}

// ---- wave-cooperative checks of the edit wave -------------------------------------------------
// The edit kernels run one wave whose 64 lanes all execute the same edits with the same values
// (every lane writes what the others write); the star checks below spread a vertex's faces over the
// lanes (lane i: face i of the list), so a check costs a few wave instructions per face instead of a
// lane walking arrays in private memory.

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
  for (int o = 32; o >= 1; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o));
  return x;
}
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long y = __shfl_xor(x, o);
    x = y < x ? y : x;
  }
  return x;
}
__device__ __forceinline__ int first_lane(bool p) {
  const unsigned long long b = __ballot(p);
  return b ? __ffsll((long long)b) - 1 : -1;
}

// `fan(v).is_some()` (mesh.rs:475) by the whole wave (the faces' relative orientations do not
// decide it, so they are not tracked).
__device__ bool fan_ok_wave(const Mesh& m, uint32_t v) {
  const int lane = threadIdx.x & 63;
  const uint32_t nfs = m.vfn[v];
  if (nfs == 0 || nfs > (uint32_t)kVF) return false;
  const bool act = (uint32_t)lane < nfs;
  const uint32_t f = act ? m.vf[(size_t)v * kVF + lane] : kNone;
  uint32_t wa = kNone, wb = kNone;
  if (act) wedge(F(m, f), v, &wa, &wb);
  uint32_t ca = 0, cb = 0;
  for (uint32_t j = 0; j < nfs; ++j) {
    const uint32_t ja = (uint32_t)__shfl((int)wa, (int)j), jb = (uint32_t)__shfl((int)wb, (int)j);
    ca += (ja == wa ? 1u : 0u) + (jb == wa ? 1u : 0u);
    cb += (ja == wb ? 1u : 0u) + (jb == wb ? 1u : 0u);
  }
  if (__ballot(act && (ca > 2 || cb > 2))) return false;
  const bool ea = act && ca == 1, eb = act && cb == 1;
  const int nends = __popcll(__ballot(ea)) + __popcll(__ballot(eb));
  if (nends != 0 && nends != 2) return false;
  uint32_t cur, entry;
  if (nends == 2) {
    // the smallest (face, neighbour) of the ends
    unsigned long long k = ~0ull;
    if (ea) k = ((unsigned long long)f << 32) | wa;
    if (eb) {
      const unsigned long long kb = ((unsigned long long)f << 32) | wb;
      k = kb < k ? kb : k;
    }
    k = wave_min_u64(k);
    cur = (uint32_t)(k >> 32);
    entry = (uint32_t)(k & 0xffffffffull);
  } else {
    cur = wave_min_u32(f);
    entry = (uint32_t)__shfl((int)wa, first_lane(f == cur));
  }
  const uint32_t start = cur;
  uint32_t c = 1;
  for (;;) {
    const int ci = first_lane(f == cur);
    const uint32_t cwa = (uint32_t)__shfl((int)wa, ci), cwb = (uint32_t)__shfl((int)wb, ci);
    const uint32_t exit = cwa == entry ? cwb : cwa;
    const uint32_t next = wave_min_u32(act && f != cur && (wa == exit || wb == exit) ? f : kNone);
    if (next == kNone || next == start) break;
    if (c > nfs) return false;
    ++c;
    cur = next;
    entry = exit;
  }
  return c == nfs;
}

// mesh.rs:558 `vertex_ok` by the whole wave.
__device__ bool vertex_ok_wave(const Mesh& m, uint32_t v) {
  if (!m.valive[v]) return true;
  if (!fan_ok_wave(m, v)) return false;
  const int lane = threadIdx.x & 63;
  const uint32_t k = m.vfn[v];
  const bool act = (uint32_t)lane < k;
  uint32_t s3[3] = {kNone, kNone, kNone};
  if (act) {
    const Tri t = F(m, m.vf[(size_t)v * kVF + lane]);
    const uint32_t lo = min(t.v[0], t.v[1]), hi = max(t.v[0], t.v[1]);
    s3[0] = min(lo, t.v[2]);
    s3[2] = max(hi, t.v[2]);
    s3[1] = max(lo, min(hi, t.v[2]));
  }
  bool dup = false;
  for (uint32_t j = 0; j < k; ++j) {
    const uint32_t a = (uint32_t)__shfl((int)s3[0], (int)j), b = (uint32_t)__shfl((int)s3[1], (int)j),
                   c = (uint32_t)__shfl((int)s3[2], (int)j);
    dup = dup || (act && (uint32_t)lane != j && a == s3[0] && b == s3[1] && c == s3[2]);
  }
  return __ballot(dup) == 0;
}

// `edge_faces` (mesh.rs:381) by the whole wave: lane i tests face i of a's list; the matches are
// taken in ascending order by repeated wave minima.
__device__ int edge_faces_wave(const Mesh& m, uint32_t a, uint32_t b, uint32_t* out) {
  if (!valid_vertex(m, a) || !valid_vertex(m, b) || a == b) return 0;
  const int lane = threadIdx.x & 63;
  const uint32_t k = m.vfn[a];
  uint32_t f = kNone;
  if ((uint32_t)lane < k) {
    const uint32_t g = m.vf[(size_t)a * kVF + lane];
    if (has(F(m, g), b)) f = g;
  }
  const int n = __popcll(__ballot(f != kNone));
  if (n > 8) flag(m, F_EDGE);
  uint32_t last = 0;
  int c = 0;
  for (; c < n && c < 8; ++c) {
    const uint32_t x = wave_min_u32((f != kNone && (c == 0 || f > last)) ? f : kNone);
    out[c] = x;
    last = x;
  }
  return c;
}

// mesh.rs:577 `remove_vface` (swap-remove), by the whole wave: lane i compares entry i.
__device__ void remove_vface(const Mesh& m, uint32_t v, uint32_t f) {
  const int lane = threadIdx.x & 63;
  const uint32_t k = m.vfn[v];
  const int i = first_lane((uint32_t)lane < k && m.vf[(size_t)v * kVF + lane] == f);
  if (i < 0) return;
  const uint32_t last = m.vf[(size_t)v * kVF + k - 1];
  m.vf[(size_t)v * kVF + i] = last;
  m.vfn[v] = k - 1;
}
__device__ void push_vface(const Mesh& m, uint32_t v, uint32_t f) {
  const uint32_t k = m.vfn[v];
  if (k >= (uint32_t)kVF) {
    flag(m, F_VALENCE);
    return;
  }
  m.vf[(size_t)v * kVF + k] = f;
  m.vfn[v] = k + 1;
}

// mesh.rs:243 `orient`
__device__ bool orient(const Tri& f, uint32_t a, uint32_t b, uint32_t* p, uint32_t* q, uint32_t* c) {
  for (int k = 0; k < 3; ++k) {
    const uint32_t pp = f.v[k], qq = f.v[(k + 1) % 3], cc = f.v[(k + 2) % 3];
    if ((pp == a && qq == b) || (pp == b && qq == a)) {
      *p = pp, *q = qq, *c = cc;
      return true;
    }
  }
  return false;
}

// mesh.rs:1038 `combine`: the new entry of a vertex from the entries of a and b, appended to the
// pool.  Returns false (and flags) when the pool is full.
__device__ bool combine(const Mesh& m, uint32_t a, double wa, uint32_t b, double wb, uint32_t* off, uint32_t* len) {
  const uint32_t oa = m.src_off[a], la = m.src_len[a], ob = m.src_off[b], lb = m.src_len[b];
  const uint32_t o = m.cnt->pool;
  if (o + la + lb > m.pcap) {
    flag(m, F_POOL);
    return false;
  }
  uint32_t n = 0;
  for (uint32_t i = 0; i < la; ++i) {
    m.pool_v[o + n] = m.pool_v[oa + i];
    m.pool_w[o + n] = m.pool_w[oa + i] * wa;
    ++n;
  }
  for (uint32_t i = 0; i < lb; ++i) {
    const uint32_t s = m.pool_v[ob + i];
    const double w = m.pool_w[ob + i];
    bool found = false;
    for (uint32_t j = 0; j < n; ++j)
      if (m.pool_v[o + j] == s) {
        m.pool_w[o + j] += w * wb;
        found = true;
        break;
      }
    if (!found) {
      m.pool_v[o + n] = s;
      m.pool_w[o + n] = w * wb;
      ++n;
    }
  }
  m.cnt->pool = o + n;
  *off = o;
  *len = n;
  return true;
}

// ---- split (mesh.rs:592) --------------------------------------------------------------------

__device__ int split_edge(const Mesh& m, uint32_t a, uint32_t b) {
  uint32_t inc[8];
  const int ni = edge_faces_wave(m, a, b, inc);
  if (ni == 0) return R_NO_SUCH_EDGE;
  const uint32_t mv = m.cnt->nv;
  const V mid = vscale(vadd(P(m, a), P(m, b)), 0.5);
  for (int i = 0; i < ni; ++i) {
    uint32_t p, q, c;
    if (orient(F(m, inc[i]), a, b, &p, &q, &c))
      if (degenerate(P(m, p), mid, P(m, c)) || degenerate(mid, P(m, q), P(m, c))) return R_DEGENERATE;
  }
  if (mv >= m.vcap || m.cnt->nf + (uint32_t)ni > m.fcap) {
    flag(m, F_SLOTS);
    return R_DEGENERATE;
  }
  uint32_t so, sl;
  if (!combine(m, a, 0.5, b, 0.5, &so, &sl)) return R_DEGENERATE;
  setP(m, mv, mid);
  m.valive[mv] = 1;
  m.pinned[mv] = 0;
  m.pin_origin[mv] = kNone;
  m.src_off[mv] = so;
  m.src_len[mv] = sl;
  m.vfn[mv] = 0;
  m.cnt->nv = mv + 1;
  for (int i = 0; i < ni; ++i) {
    const uint32_t f = inc[i];
    uint32_t p, q, c;
    if (!orient(F(m, f), a, b, &p, &q, &c)) continue;
    const uint32_t nf = m.cnt->nf;
    setF(m, f, p, mv, c);
    setF(m, nf, mv, q, c);
    m.falive[nf] = 1;
    m.parent[nf] = m.parent[f];
    m.cnt->nf = nf + 1;
    remove_vface(m, q, f);
    push_vface(m, q, nf);
    push_vface(m, c, nf);
    push_vface(m, mv, f);
    push_vface(m, mv, nf);
  }
  return R_OK;
}

// ---- flip (mesh.rs:888) ---------------------------------------------------------------------

__device__ int flip_edge(const Mesh& m, uint32_t a, uint32_t b, bool delaunay, bool foldover, bool* reordered) {
  uint32_t inc[8];
  const int ni = edge_faces_wave(m, a, b, inc);
  *reordered = false;
  if (ni == 0) return R_NO_SUCH_EDGE;
  if (ni != 2) return R_BOUNDARY;
  const uint32_t f1 = inc[0], f2 = inc[1];
  const Tri t1 = F(m, f1), t2 = F(m, f2);
  uint32_t p, q, c, p2, q2, d;
  if (!orient(t1, a, b, &p, &q, &c) || !orient(t2, a, b, &p2, &q2, &d)) return R_NO_SUCH_EDGE;
  *reordered = p2 == p;
  uint32_t tmp[8];
  if (edge_faces_wave(m, c, d, tmp) != 0 || c == d) return R_EXISTING;
  const V xp = P(m, p), xq = P(m, q), xc = P(m, c), xd = P(m, d);
  if (delaunay && !violates_delaunay(xp, xq, xc, xd)) return R_NOT_DELAUNAY;
  if (foldover) {
    // mesh.rs:1058 `flip_folds`
    const V n1 = tri_normal(P(m, t1.v[0]), P(m, t1.v[1]), P(m, t1.v[2]));
    V n2 = tri_normal(P(m, t2.v[0]), P(m, t2.v[1]), P(m, t2.v[2]));
    if (*reordered) n2 = vscale(n2, -1.0);
    if (folds(n1, tri_normal(xp, xd, xc)) || folds(n2, tri_normal(xd, xq, xc))) return R_FOLDOVER;
  }
  if (degenerate(xp, xd, xc) || degenerate(xd, xq, xc)) return R_DEGENERATE;
  setF(m, f1, p, d, c);
  setF(m, f2, d, q, c);
  remove_vface(m, p, f2);
  remove_vface(m, q, f1);
  push_vface(m, c, f2);
  push_vface(m, d, f1);
  const uint32_t vs[4] = {p, q, c, d};
  for (int i = 0; i < 4; ++i)
    if (!vertex_ok_wave(m, vs[i])) {
      // undone by the inverse edits (the order of a face list never changes a result)
      setF(m, f1, t1.v[0], t1.v[1], t1.v[2]);
      setF(m, f2, t2.v[0], t2.v[1], t2.v[2]);
      remove_vface(m, c, f2);
      remove_vface(m, d, f1);
      push_vface(m, p, f2);
      push_vface(m, q, f1);
      return R_NON_MANIFOLD;
    }
  return R_OK;
}

// ---- collapse (mesh.rs:679) -----------------------------------------------------------------

// The state of one collapse between its checks and its application (the guard runs between).
struct Collapse {
  uint32_t a, b, keep, gone;
  V x;
  double wa, wb;
  uint32_t inc[8];
  int ni;
  uint32_t star[kNb];
  int nstar;
  uint32_t newstar[kNb][3];
};

// The checks up to the guard: returns R_OK when the collapse may proceed.
__device__ int collapse_check(const Mesh& m, uint32_t a, uint32_t b, bool foldover, Collapse& C) {
  C.a = a, C.b = b;
  C.ni = edge_faces_wave(m, a, b, C.inc);
  if (C.ni == 0) return R_NO_SUCH_EDGE;
  if (m.pinned[a] && m.pinned[b]) return R_BOTH_PINNED;
  {
    uint32_t na[kNb], nb[kNb];
    const int ka = neighbors(m, a, na), kb = neighbors(m, b, nb);
    int common = 0, j = 0;
    for (int i = 0; i < ka; ++i) {
      while (j < kb && nb[j] < na[i]) ++j;
      if (j < kb && nb[j] == na[i]) ++common;
    }
    if (common != C.ni) return R_LINK;
  }
  const bool ba = is_boundary_vertex(m, a), bb = is_boundary_vertex(m, b);
  if (C.ni == 2 && ba && bb) return R_LINK;
  C.keep = min(a, b);
  C.gone = max(a, b);
  // mesh.rs:652 `collapse_target`
  const bool pa = m.pinned[a], pb = m.pinned[b];
  if (pa) {
    C.x = P(m, a), C.wa = 1.0, C.wb = 0.0;
  } else if (pb) {
    C.x = P(m, b), C.wa = 0.0, C.wb = 1.0;
  } else if (ba && !bb) {
    C.x = P(m, a), C.wa = 1.0, C.wb = 0.0;
  } else if (bb && !ba) {
    C.x = P(m, b), C.wa = 0.0, C.wb = 1.0;
  } else {
    C.x = vscale(vadd(P(m, a), P(m, b)), 0.5), C.wa = 0.5, C.wb = 0.5;
  }
  // the faces touching exactly one end, ascending
  int ns = 0;
  for (int e = 0; e < 2; ++e) {
    const uint32_t v = e == 0 ? a : b;
    const uint32_t k = m.vfn[v];
    for (uint32_t i = 0; i < k; ++i) {
      const uint32_t f = m.vf[(size_t)v * kVF + i];
      bool on = false;
      for (int j = 0; j < C.ni; ++j) on = on || C.inc[j] == f;
      if (!on) C.star[ns++] = f;
    }
  }
  isort(C.star, ns);
  int u = 0;
  for (int i = 0; i < ns; ++i)
    if (u == 0 || C.star[u - 1] != C.star[i]) C.star[u++] = C.star[i];
  C.nstar = u;
  if (foldover)
    for (int i = 0; i < C.nstar; ++i) {
      const Tri t = F(m, C.star[i]);
      V p[3], q[3];
      for (int k = 0; k < 3; ++k) {
        p[k] = P(m, t.v[k]);
        q[k] = (t.v[k] == a || t.v[k] == b) ? C.x : p[k];
      }
      if (folds(tri_normal(p[0], p[1], p[2]), tri_normal(q[0], q[1], q[2]))) return R_FOLDOVER;
    }
  for (int i = 0; i < C.nstar; ++i) {
    const Tri t = F(m, C.star[i]);
    V q[3];
    for (int k = 0; k < 3; ++k) q[k] = (t.v[k] == a || t.v[k] == b) ? C.x : P(m, t.v[k]);
    if (degenerate(q[0], q[1], q[2])) return R_DEGENERATE;
  }
  for (int i = 0; i < C.nstar; ++i) {
    const Tri t = F(m, C.star[i]);
    for (int k = 0; k < 3; ++k) C.newstar[i][k] = t.v[k] == C.gone ? C.keep : t.v[k];
  }
  for (int i = 0; i < C.nstar; ++i) {
    uint32_t si[3] = {C.newstar[i][0], C.newstar[i][1], C.newstar[i][2]};
    isort(si, 3);
    for (int j = i + 1; j < C.nstar; ++j) {
      uint32_t sj[3] = {C.newstar[j][0], C.newstar[j][1], C.newstar[j][2]};
      isort(sj, 3);
      if (si[0] == sj[0] && si[1] == sj[1] && si[2] == sj[2]) return R_REPEATED;
    }
  }
  return R_OK;
}

// The application, with the stars checked afterwards and undone if one fails.
__device__ int collapse_apply(const Mesh& m, const Collapse& C) {
  const uint32_t a = C.a, b = C.b, keep = C.keep, gone = C.gone;
  uint32_t touched[2 + 3 * 8];
  int nt = 0;
  touched[nt++] = a;
  touched[nt++] = b;
  for (int i = 0; i < C.ni; ++i) {
    const Tri t = F(m, C.inc[i]);
    for (int k = 0; k < 3; ++k) touched[nt++] = t.v[k];
  }
  isort(touched, nt);
  int u = 0;
  for (int i = 0; i < nt; ++i)
    if (u == 0 || touched[u - 1] != touched[i]) touched[u++] = touched[i];
  nt = u;
  uint32_t snap[2 + 3 * 8][kVF], snapn[2 + 3 * 8];
  for (int i = 0; i < nt; ++i) {
    snapn[i] = m.vfn[touched[i]];
    for (uint32_t j = 0; j < snapn[i]; ++j) snap[i][j] = m.vf[(size_t)touched[i] * kVF + j];
  }
  const V old_pos = P(m, keep);
  const uint8_t old_pin = m.pinned[keep];
  const uint32_t old_origin = m.pin_origin[keep], old_off = m.src_off[keep], old_len = m.src_len[keep];
  uint32_t so, sl;
  if (!combine(m, a, C.wa, b, C.wb, &so, &sl)) return R_NON_MANIFOLD;
  // star faces are snapshotted by value (their vertex lists change below)
  uint32_t starv[kNb][3];
  for (int i = 0; i < C.nstar; ++i) {
    const Tri t = F(m, C.star[i]);
    for (int k = 0; k < 3; ++k) starv[i][k] = t.v[k];
  }
  for (int i = 0; i < C.ni; ++i) {
    const uint32_t f = C.inc[i];
    m.falive[f] = 0;
    const Tri t = F(m, f);
    for (int k = 0; k < 3; ++k) remove_vface(m, t.v[k], f);
  }
  const uint32_t kg = m.vfn[gone];
  for (uint32_t i = 0; i < kg; ++i) {
    const uint32_t f = m.vf[(size_t)gone * kVF + i];
    Tri t = F(m, f);
    for (int k = 0; k < 3; ++k)
      if (t.v[k] == gone) t.v[k] = keep;
    setF(m, f, t.v[0], t.v[1], t.v[2]);
    push_vface(m, keep, f);
  }
  m.vfn[gone] = 0;
  m.valive[gone] = 0;
  setP(m, keep, C.x);
  const bool pa = m.pinned[a], pb = m.pinned[b];
  const uint32_t origin = pa ? m.pin_origin[a] : (pb ? m.pin_origin[b] : kNone);
  m.pinned[keep] = (pa || pb) ? 1 : 0;
  m.pin_origin[keep] = origin;
  m.src_off[keep] = so;
  m.src_len[keep] = sl;
  // the stars of keep's neighbours, keep and the touched vertices
  uint32_t check[kNb + 3 + 3 * 8];
  int nc = neighbors(m, keep, check);
  check[nc++] = keep;
  for (int i = 0; i < nt; ++i)
    if (touched[i] != gone) check[nc++] = touched[i];
  isort(check, nc);
  bool ok = true;
  for (int i = 0; i < nc && ok; ++i)
    if (i == 0 || check[i] != check[i - 1]) ok = vertex_ok_wave(m, check[i]);
  if (!ok) {
    for (int i = 0; i < nt; ++i) {
      m.vfn[touched[i]] = snapn[i];
      for (uint32_t j = 0; j < snapn[i]; ++j) m.vf[(size_t)touched[i] * kVF + j] = snap[i][j];
    }
    for (int i = 0; i < C.nstar; ++i) setF(m, C.star[i], starv[i][0], starv[i][1], starv[i][2]);
    for (int i = 0; i < C.ni; ++i) m.falive[C.inc[i]] = 1;
    m.valive[gone] = 1;
    setP(m, keep, old_pos);
    m.pinned[keep] = old_pin;
    m.pin_origin[keep] = old_origin;
    m.src_off[keep] = old_off;
    m.src_len[keep] = old_len;
    return R_NON_MANIFOLD;
  }
  return R_OK;
}

// ---- kernels --------------------------------------------------------------------------------

struct Inputs {
  const uint32_t* faces;
  const double* pos;
  uint32_t nv, nf;
};

// Slots of the input: positions, faces, parents, flags and the identity vertex map.
__global__ __launch_bounds__(256) void rm_init_vertices(Mesh m, Inputs in) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= in.nv) return;
  for (int k = 0; k < 3; ++k) m.pos[3 * (size_t)v + k] = in.pos[3 * (size_t)v + k];
  m.valive[v] = 1;
  m.pinned[v] = 0;
  m.pin_origin[v] = kNone;
  m.src_off[v] = v;
  m.src_len[v] = 1;
  m.pool_v[v] = v;
  m.pool_w[v] = 1.0;
  m.vfn[v] = 0;
}
__global__ __launch_bounds__(256) void rm_init_faces(Mesh m, Inputs in) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= in.nf) return;
  for (int k = 0; k < 3; ++k) {
    const uint32_t v = in.faces[3 * (size_t)f + k];
    m.faces[3 * (size_t)f + k] = v;
    const uint32_t slot = atomicAdd(&m.vfn[v], 1u);
    if (slot < (uint32_t)kVF)
      m.vf[(size_t)v * kVF + slot] = f;
    else
      flag(m, F_VALENCE);
  }
  m.falive[f] = 1;
  m.parent[f] = f;
}
// The face lists in ascending order (their order never changes a result; sorted for run-to-run
// identical memory contents).
__global__ __launch_bounds__(256) void rm_sort_lists(Mesh m, uint32_t nv) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const uint32_t k = min(m.vfn[v], (uint32_t)kVF);
  m.vfn[v] = k;
  uint32_t l[kVF];
  for (uint32_t i = 0; i < k; ++i) l[i] = m.vf[(size_t)v * kVF + i];
  isort(l, (int)k);
  for (uint32_t i = 0; i < k; ++i) m.vf[(size_t)v * kVF + i] = l[i];
}
__global__ __launch_bounds__(256) void rm_pins(Mesh m, const uint32_t* pins, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  m.pinned[pins[i]] = 1;
  m.pin_origin[pins[i]] = pins[i];
}

// mesh.rs:404 `edges_with_counts`: per live vertex v its neighbours w > v with their face counts.
// Pass 0 counts, pass 1 writes at the scanned offsets (lexicographic order).
__global__ __launch_bounds__(256) void rm_edges(Mesh m, uint32_t nv, int pass, uint32_t* count,
                                                const uint32_t* off, uint32_t* ea, uint32_t* eb, uint8_t* ec) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  if (!m.valive[v]) {
    if (pass == 0) count[v] = 0;
    return;
  }
  uint32_t buf[kNb];
  int n = 0;
  const uint32_t k = m.vfn[v];
  for (uint32_t i = 0; i < k; ++i) {
    const Tri t = F(m, m.vf[(size_t)v * kVF + i]);
    for (int c = 0; c < 3; ++c)
      if (t.v[c] > v) buf[n++] = t.v[c];
  }
  isort(buf, n);
  uint32_t o = pass == 1 ? off[v] : 0, u = 0;
  int i = 0;
  while (i < n) {
    int j = i + 1;
    while (j < n && buf[j] == buf[i]) ++j;
    if (pass == 1) {
      ea[o + u] = v;
      eb[o + u] = buf[i];
      ec[o + u] = (uint8_t)min(j - i, 255);
    }
    ++u;
    i = j;
  }
  if (pass == 0) count[v] = u;
}

// Split candidates: edges longer than thr, with the faces a split adds (1 on the boundary).
__global__ __launch_bounds__(256) void rm_long(Mesh m, uint32_t ne, const uint32_t* ea, const uint32_t* eb,
                                               const uint8_t* ec, double thr, uint32_t* is_long, double* len) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const double l = vnorm(vsub(P(m, ea[e]), P(m, eb[e])));
  len[e] = l;
  is_long[e] = l > thr ? (ec[e] == 1 ? 1u : 2u) : 0u;
}

// Short edges: (length bits, lexicographic index) keys of the edges shorter than thr.
__global__ __launch_bounds__(256) void rm_short(Mesh m, uint32_t ne, const uint32_t* ea, const uint32_t* eb,
                                                double thr, uint32_t* is_short, unsigned long long* bits) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const double l = vnorm(vsub(P(m, ea[e]), P(m, eb[e])));
  is_short[e] = l < thr ? 1u : 0u;
  bits[e] = (unsigned long long)__double_as_longlong(l);
}

// Compacts the indices e with flag[e] != 0 through the exclusive scan `off`.
__global__ __launch_bounds__(256) void rm_compact(uint32_t n, const uint32_t* flagv, const uint32_t* off, uint32_t* out) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  if (flagv[e]) out[off[e]] = e;
}
__global__ __launch_bounds__(256) void rm_nonzero(uint32_t n, const uint32_t* in, uint32_t* out) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  out[e] = in[e] != 0 ? 1u : 0u;
}
__global__ __launch_bounds__(256) void rm_gather_keys(uint32_t n, const uint32_t* idx,
                                                      const unsigned long long* bits, unsigned long long* keys) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  keys[i] = bits[idx[i]];
}

// The long edges' added faces and lengths, in the order of their list.
__global__ __launch_bounds__(256) void rm_gather_long(uint32_t n, const uint32_t* list, const uint32_t* added_all,
                                                      const double* len_all, uint32_t* added, double* len) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  added[i] = added_all[list[i]];
  len[i] = len_all[list[i]];
}

// The growth budget (mod.rs:260-283): longest first (ties by edge index), each taken while it fits.
// Keys are the long edges sorted by (−length, index); one thread walks them (rare: only when the
// long edges would more than multiply the face count by max_growth).
__global__ void rm_budget(uint32_t nl, const uint32_t* order, const uint32_t* added, double budget, uint32_t* keep,
                          Counters* cnt) {
  if (threadIdx.x != 0) return;
  double used = 0.0;
  unsigned long long deferred = 0;
  for (uint32_t i = 0; i < nl; ++i) {
    const uint32_t j = order[i];
    const double ad = (double)added[j];
    if (used + ad > budget) {
      keep[j] = 0;
      ++deferred;
      continue;
    }
    used += ad;
    keep[j] = 1;
  }
  cnt->splits_deferred = deferred;
}


// Splits of the kept long edges in lexicographic order (mod.rs:290-297), by one wave.
__global__ void rm_split_engine(Mesh m, const uint32_t* ea, const uint32_t* eb, const uint32_t* list,
                                const uint32_t* keep, uint32_t nl) {
  if (threadIdx.x >= 64) return;
  for (uint32_t i = 0; i < nl; ++i) {
    if (keep && !keep[i]) continue;
    const uint32_t e = list[i];
    const int r = split_edge(m, ea[e], eb[e]);
    if (r == R_OK)
      ++m.cnt->splits;
    else
      ++m.cnt->refused_split[r - 1];
    if (m.cnt->flags) return;
  }
}

// ---- collapses (mod.rs:338 collapse_short) ---------------------------------------------------

struct HeapEntry {
  unsigned long long bits;
  uint32_t a, b;
};
__device__ __forceinline__ bool heap_less(const HeapEntry& x, const HeapEntry& y) {
  if (x.bits != y.bits) return x.bits < y.bits;
  if (x.a != y.a) return x.a < y.a;
  return x.b < y.b;
}
__device__ void heap_push(HeapEntry* h, uint32_t& n, uint32_t cap, HeapEntry e, const Mesh& m) {
  if (n >= cap) {
    flag(m, F_HEAP);
    return;
  }
  uint32_t i = n++;
  while (i > 0) {
    const uint32_t p = (i - 1) / 2;
    if (!heap_less(e, h[p])) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = e;
}
__device__ HeapEntry heap_pop(HeapEntry* h, uint32_t& n) {
  const HeapEntry top = h[0];
  const HeapEntry last = h[--n];
  uint32_t i = 0;
  for (;;) {
    const uint32_t l = 2 * i + 1, r = l + 1;
    if (l >= n) break;
    const uint32_t c = (r < n && heap_less(h[r], h[l])) ? r : l;
    if (!heap_less(h[c], last)) break;
    h[i] = h[c];
    i = c;
  }
  if (n > 0) h[i] = last;
  return top;
}

struct CollapseArgs {
  const unsigned long long* sbits;  // short edges sorted by (length bits, lexicographic index)
  const uint32_t* sidx;
  uint32_t ns;
  const uint32_t* ea;
  const uint32_t* eb;
  HeapEntry* heap;
  uint32_t heap_cap;
  double thr;
  double gap;            // > 0: the collapse guard
  const uint8_t* band;   // null: no band
  uint32_t nband;
};

__device__ __forceinline__ bool band_flag(const CollapseArgs& A, uint32_t v) {
  return A.band && v < A.nband && A.band[v];
}

// Collapses shortest first: the sorted short edges merged with the heap of requeued ones, the
// reference's checks, the guard evaluated by the whole workgroup (ccd.rs:1538 `allow`: a
// collapse is refused when a new star triangle comes within the gap of a live face that shares
// no vertex with it and does not touch the edge, pairs inside the band exempt).  Every face whose
// box within the gap overlaps a star triangle's box has its vertices inside the reference's query
// box (gap plus the longest edge around the star), so scanning every live face finds the pairs the
// reference's grid finds.
__global__ __launch_bounds__(kGuardThreads) void rm_collapse_engine(Mesh m, CollapseArgs A) {
  __shared__ uint32_t s_cmd;
  __shared__ int s_refuse;
  __shared__ uint32_t s_tri[kNb][3];
  __shared__ double s_pos[kNb][9];
  __shared__ double s_box[kNb][6];
  __shared__ uint8_t s_in[kNb];
  __shared__ uint32_t s_ea, s_eb, s_n;
  const int tid = threadIdx.x;
  Collapse C;
  uint32_t si = 0, hn = 0;
  for (;;) {
    if (tid < 64) {
      s_cmd = 0;
      while (!m.cnt->flags) {
        HeapEntry e;
        const bool from_sorted = si < A.ns;
        if (!from_sorted && hn == 0) break;
        if (from_sorted) {
          const uint32_t x = A.sidx[si];
          e = {A.sbits[si], A.ea[x], A.eb[x]};
        }
        if (hn > 0 && (!from_sorted || heap_less(A.heap[0], e))) {
          e = heap_pop(A.heap, hn);
        } else {
          ++si;
        }
        const uint32_t a = e.a, b = e.b;
        if (!m.valive[a] || !m.valive[b]) continue;
        uint32_t tmp[8];
        if (edge_faces_wave(m, a, b, tmp) == 0) continue;
        const double l = vnorm(vsub(P(m, a), P(m, b)));
        if ((unsigned long long)__double_as_longlong(l) != e.bits || l >= A.thr) continue;
        const int r = collapse_check(m, a, b, true, C);
        if (r != R_OK) {
          ++m.cnt->refused_collapse[r - 1];
          continue;
        }
        if (A.gap > 0.0) {
          s_ea = a;
          s_eb = b;
          s_n = (uint32_t)C.nstar;
          const bool fe = band_flag(A, a) || band_flag(A, b);
          for (int i = 0; i < C.nstar; ++i) {
            bool in = false;
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int k = 0; k < 3; ++k) {
              const uint32_t v = C.newstar[i][k];
              s_tri[i][k] = v;
              in = in || band_flag(A, v) || (v == C.keep && fe);
              const V p = v == C.keep ? C.x : P(m, v);
              s_pos[i][3 * k] = p.x, s_pos[i][3 * k + 1] = p.y, s_pos[i][3 * k + 2] = p.z;
              const double c[3] = {p.x, p.y, p.z};
              for (int q = 0; q < 3; ++q) {
                lo[q] = fmin(lo[q], c[q]);
                hi[q] = fmax(hi[q], c[q]);
              }
            }
            s_in[i] = in ? 1 : 0;
            for (int q = 0; q < 3; ++q) {
              s_box[i][q] = lo[q];
              s_box[i][3 + q] = hi[q];
            }
          }
          s_cmd = 1;
          break;
        }
        if (collapse_apply(m, C) == R_OK) {
          ++m.cnt->collapses;
          uint32_t nb[kNb];
          const int k = neighbors(m, C.keep, nb);
          for (int i = 0; i < k; ++i) {
            const double lw = vnorm(vsub(P(m, C.keep), P(m, nb[i])));
            if (lw < A.thr)
              heap_push(A.heap, hn, A.heap_cap,
                        {(unsigned long long)__double_as_longlong(lw), min(C.keep, nb[i]), max(C.keep, nb[i])}, m);
          }
        } else {
          ++m.cnt->refused_collapse[R_NON_MANIFOLD - 1];
        }
      }
    }
    __syncthreads();
    if (s_cmd == 0) break;
    if (tid == 0) s_refuse = 0;
    __syncthreads();
    const uint32_t ea = s_ea, eb = s_eb, ns = s_n, nf = m.cnt->nf;
    const bool has_band = A.band != nullptr;
    for (uint32_t g = tid; g < nf; g += kGuardThreads) {
      if (s_refuse) break;
      if (!m.falive[g]) continue;
      const Tri gv = F(m, g);
      if (has(gv, ea) || has(gv, eb)) continue;
      V gt[3];
      double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int k = 0; k < 3; ++k) {
        gt[k] = P(m, gv.v[k]);
        const double c[3] = {gt[k].x, gt[k].y, gt[k].z};
        for (int q = 0; q < 3; ++q) {
          glo[q] = fmin(glo[q], c[q]);
          ghi[q] = fmax(ghi[q], c[q]);
        }
      }
      for (int q = 0; q < 3; ++q) {
        glo[q] = glo[q] - A.gap;
        ghi[q] = ghi[q] + A.gap;
      }
      const bool g_in = has_band && (band_flag(A, gv.v[0]) || band_flag(A, gv.v[1]) || band_flag(A, gv.v[2]));
      for (uint32_t t = 0; t < ns; ++t) {
        if (has(gv, s_tri[t][0]) || has(gv, s_tri[t][1]) || has(gv, s_tri[t][2])) continue;
        if (g_in && s_in[t]) continue;
        bool ov = true;
        for (int q = 0; q < 3; ++q) ov = ov && glo[q] <= s_box[t][3 + q] && s_box[t][q] <= ghi[q];
        if (!ov) continue;
        V tt[3];
        for (int k = 0; k < 3; ++k) tt[k] = {s_pos[t][3 * k], s_pos[t][3 * k + 1], s_pos[t][3 * k + 2]};
        if (triangle_distance(tt, gt) < A.gap) {
          atomicOr(&s_refuse, 1);
          break;
        }
      }
    }
    __syncthreads();
    if (tid < 64) {
      if (s_refuse) {
        ++m.cnt->refused_collapse[R_GUARD - 1];
      } else if (collapse_apply(m, C) == R_OK) {
        ++m.cnt->collapses;
        uint32_t nb[kNb];
        const int k = neighbors(m, C.keep, nb);
        for (int i = 0; i < k; ++i) {
          const double lw = vnorm(vsub(P(m, C.keep), P(m, nb[i])));
          if (lw < A.thr)
            heap_push(A.heap, hn, A.heap_cap,
                      {(unsigned long long)__double_as_longlong(lw), min(C.keep, nb[i]), max(C.keep, nb[i])}, m);
        }
      } else {
        ++m.cnt->refused_collapse[R_NON_MANIFOLD - 1];
      }
    }
    __syncthreads();
  }
}

// ---- flips (mod.rs:380 delaunay_flips) -------------------------------------------------------

// Per edge of the round's lexicographic list: 1 when it is interior and violates the Delaunay test.
__global__ __launch_bounds__(256) void rm_violating(Mesh m, uint32_t ne, const uint32_t* ea, const uint32_t* eb,
                                                    uint32_t* viol) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const uint32_t a = ea[e], b = eb[e];
  uint32_t inc[8];
  uint32_t r = 0;
  if (edge_faces(m, a, b, inc) == 2) {
    uint32_t cd[2];
    for (int i = 0; i < 2; ++i) {
      const Tri t = F(m, inc[i]);
      cd[i] = a;
      for (int k = 0; k < 3; ++k)
        if (t.v[k] != a && t.v[k] != b) {
          cd[i] = t.v[k];
          break;
        }
    }
    r = violates_delaunay(P(m, a), P(m, b), P(m, cd[0]), P(m, cd[1])) ? 1u : 0u;
  }
  viol[e] = r;
}

struct FlipArgs {
  const uint32_t* ea;
  const uint32_t* eb;
  const uint32_t* eoff;   // per vertex: its edges' range in the lexicographic list
  uint32_t ne;
  const uint32_t* v0;     // violating edges at the round's start, ascending
  uint32_t nv0;
  uint32_t* revisit;      // heap of pending lexicographic positions to re-test
  uint8_t* rflag;         // per position: pushed to `revisit`
  unsigned long long* fifo;
  uint32_t fifo_cap;
  uint32_t* qn;           // per vertex: requeued edges (v, w > v) in the FIFO
  uint32_t* q;
  unsigned long long cap;  // flips allowed this round
};

__device__ uint32_t lex_position(const FlipArgs& A, uint32_t lo, uint32_t hi) {
  uint32_t l = A.eoff[lo], r = A.eoff[lo + 1];
  while (l < r) {
    const uint32_t mid = (l + r) / 2;
    if (A.eb[mid] < hi)
      l = mid + 1;
    else
      r = mid;
  }
  return (l < A.eoff[lo + 1] && A.eb[l] == hi) ? l : kNone;
}
// The requeued edges (lo, hi) are kept per lower end; lane i compares entry i (at most kNb).
__device__ int queued_lane(const FlipArgs& A, uint32_t lo, uint32_t hi) {
  const int lane = threadIdx.x & 63;
  const uint32_t n = A.qn[lo];
  return first_lane((uint32_t)lane < n && A.q[(size_t)lo * kNb + lane] == hi);
}
__device__ bool queued(const FlipArgs& A, uint32_t lo, uint32_t hi) { return queued_lane(A, lo, hi) >= 0; }
__device__ void unqueue(const FlipArgs& A, uint32_t lo, uint32_t hi) {
  const int i = queued_lane(A, lo, hi);
  if (i < 0) return;
  const uint32_t n = A.qn[lo];
  A.q[(size_t)lo * kNb + i] = A.q[(size_t)lo * kNb + n - 1];
  A.qn[lo] = n - 1;
}
__device__ void rheap_push(uint32_t* h, uint32_t& n, uint32_t x) {
  uint32_t i = n++;
  while (i > 0) {
    const uint32_t p = (i - 1) / 2;
    if (h[p] <= x) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}
__device__ uint32_t rheap_pop(uint32_t* h, uint32_t& n) {
  const uint32_t top = h[0], last = h[--n];
  uint32_t i = 0;
  for (;;) {
    const uint32_t l = 2 * i + 1, r = l + 1;
    if (l >= n) break;
    const uint32_t c = (r < n && h[r] < h[l]) ? r : l;
    if (h[c] >= last) break;
    h[i] = h[c];
    i = c;
  }
  if (n > 0) h[i] = last;
  return top;
}

// One round of flips by one wave: the reference's queue visited through the violating edges and
// the pending quad edges (pass over the lexicographic list), then its FIFO of requeued edges.
__global__ void rm_flip_engine(Mesh m, FlipArgs A) {
  if (threadIdx.x >= 64) return;
  unsigned long long flips = 0, head = 0, tail = 0;
  uint32_t i0 = 0, rn = 0, cursor = 0, last = kNone;
  bool abort = false;
  auto process = [&](uint32_t a, uint32_t b) {
    uint32_t inc[8];
    if (edge_faces_wave(m, a, b, inc) != 2) return;
    uint32_t cd[2];
    for (int i = 0; i < 2; ++i) {
      const Tri t = F(m, inc[i]);
      cd[i] = a;
      for (int k = 0; k < 3; ++k)
        if (t.v[k] != a && t.v[k] != b) {
          cd[i] = t.v[k];
          break;
        }
    }
    const uint32_t c = cd[0], d = cd[1];
    if (!violates_delaunay(P(m, a), P(m, b), P(m, c), P(m, d))) return;
    if (flips >= A.cap) {
      ++m.cnt->flip_cap_hits;
      abort = true;
      return;
    }
    bool reo;
    const int r = flip_edge(m, a, b, false, true, &reo);
    if (r != R_OK) {
      ++m.cnt->refused_flip[r - 1];
      return;
    }
    ++flips;
    const uint32_t qe[4][2] = {{a, c}, {c, b}, {b, d}, {d, a}};
    for (int j = 0; j < 4; ++j) {
      const uint32_t lo = min(qe[j][0], qe[j][1]), hi = max(qe[j][0], qe[j][1]);
      const uint32_t lp = lex_position(A, lo, hi);
      if (lp != kNone && lp > cursor && cursor < A.ne) {
        // still pending in the pass over the lexicographic list: re-tested when the pass gets there
        if (!A.rflag[lp]) {
          A.rflag[lp] = 1;
          rheap_push(A.revisit, rn, lp);
        }
      } else if (!queued(A, lo, hi)) {
        if (tail - head >= A.fifo_cap || A.qn[lo] >= (uint32_t)kNb) {
          flag(m, F_FIFO);
          abort = true;
          return;
        }
        A.fifo[tail % A.fifo_cap] = ((unsigned long long)lo << 32) | hi;
        ++tail;
        A.q[(size_t)lo * kNb + A.qn[lo]] = hi;
        A.qn[lo] += 1;
      }
    }
  };
  // pass over the lexicographic list
  while (!abort && !m.cnt->flags) {
    uint32_t next;
    const bool has0 = i0 < A.nv0;
    if (!has0 && rn == 0) break;
    if (rn > 0 && (!has0 || A.revisit[0] < A.v0[i0]))
      next = rheap_pop(A.revisit, rn);
    else
      next = A.v0[i0++];
    if (next == last) continue;
    last = next;
    cursor = next;
    process(A.ea[next], A.eb[next]);
  }
  cursor = A.ne;
  while (!abort && !m.cnt->flags && head < tail) {
    const unsigned long long k = A.fifo[head % A.fifo_cap];
    ++head;
    const uint32_t lo = (uint32_t)(k >> 32), hi = (uint32_t)(k & 0xffffffffull);
    unqueue(A, lo, hi);
    process(lo, hi);
  }
  m.cnt->flips += flips;
}

// ---- smoothing (mod.rs:424 smooth) -----------------------------------------------------------

// Boundary vertices and the faces on a boundary edge (an edge with one face).
__global__ __launch_bounds__(256) void rm_boundary(Mesh m, uint32_t ne, const uint32_t* ea, const uint32_t* eb,
                                                   const uint8_t* ec, uint8_t* bnd, uint8_t* fob) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne || ec[e] != 1) return;
  const uint32_t a = ea[e], b = eb[e];
  bnd[a] = 1;
  bnd[b] = 1;
  uint32_t inc[8];
  const int n = edge_faces(m, a, b, inc);
  for (int i = 0; i < n; ++i) fob[inc[i]] = 1;
}
__global__ __launch_bounds__(256) void rm_face_geom(Mesh m, uint32_t nf, const uint8_t* fob, double* normal0,
                                                    double* area, double* center) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf || !m.falive[f]) return;
  const Tri t = F(m, f);
  const V p0 = P(m, t.v[0]), p1 = P(m, t.v[1]), p2 = P(m, t.v[2]);
  const V n = tri_normal(p0, p1, p2);
  const V c = fob[f] ? barycenter(p0, p1, p2) : circumcenter(p0, p1, p2);
  normal0[3 * (size_t)f] = n.x, normal0[3 * (size_t)f + 1] = n.y, normal0[3 * (size_t)f + 2] = n.z;
  area[f] = 0.5 * vnorm(n);
  center[3 * (size_t)f] = c.x, center[3 * (size_t)f + 1] = c.y, center[3 * (size_t)f + 2] = c.z;
}
__device__ __forceinline__ V load3d(const double* a, uint32_t i) {
  return {a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]};
}

// The move of each vertex: angle-weighted normal of its fan (faces oriented consistently), the
// area-weighted circumcenter displacement, projected onto the tangent plane.  mover[v] = 1 with
// u[v] set, or 0 (skipped ones counted).
__global__ __launch_bounds__(256) void rm_smooth_moves(Mesh m, uint32_t nv, double rho, const uint8_t* bnd,
                                                       const double* normal0, const double* area, const double* center,
                                                       uint8_t* mover, double* u_out) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  mover[v] = 0;
  if (!m.valive[v] || m.pinned[v] || bnd[v]) return;
  const V x = P(m, v);
  uint32_t fs[kVF];
  bool fl[kVF];
  int c;
  if (!fan(m, v, fs, fl, &c)) {
    atomicAdd(&m.cnt->smoothing_skipped, 1ull);
    return;
  }
  V nrm = {0.0, 0.0, 0.0};
  for (int i = 0; i < c; ++i) {
    const Tri t = F(m, fs[i]);
    int k = 0;
    for (int j = 0; j < 3; ++j)
      if (t.v[j] == v) {
        k = j;
        break;
      }
    const double ang = angle_between(vsub(P(m, t.v[(k + 1) % 3]), x), vsub(P(m, t.v[(k + 2) % 3]), x));
    const V n = load3d(normal0, fs[i]);
    const double len = vnorm(n);
    const double s = fl[i] ? -ang / len : ang / len;
    nrm = vadd(nrm, vscale(n, s));
  }
  const double len = vnorm(nrm);
  V sum = {0.0, 0.0, 0.0};
  double asum = 0.0;
  for (int i = 0; i < c; ++i) {
    const double a = area[fs[i]];
    sum = vadd(sum, vscale(vsub(load3d(center, fs[i]), x), a));
    asum += a;
  }
  const bool disp = asum > 0.0 && isfinite(asum);
  if (!(disp && len > 0.0 && isfinite(len))) {
    atomicAdd(&m.cnt->smoothing_skipped, 1ull);
    return;
  }
  V u = vscale(sum, rho / asum);
  const V n = vscale(nrm, 1.0 / len);
  u = vsub(u, vscale(n, vdot(n, u)));
  if (!(isfinite(u.x) && isfinite(u.y) && isfinite(u.z))) {
    atomicAdd(&m.cnt->smoothing_skipped, 1ull);
    return;
  }
  mover[v] = 1;
  u_out[3 * (size_t)v] = u.x, u_out[3 * (size_t)v + 1] = u.y, u_out[3 * (size_t)v + 2] = u.z;
}

// The guarded move of v (mod.rs:485-513) given its neighbours' positions: lower ids from `low`,
// higher ids from the start positions.  Returns the halvings taken (4: refused) and the position.
__device__ int guarded_move(const Mesh& m, uint32_t v, const double* start, const double* low, const double* normal0,
                            const double* u, V* out) {
  const V x = load3d(start, v);
  V step = load3d(u, v);
  const uint32_t k = m.vfn[v];
  for (int h = 0; h <= 3; ++h) {
    const V y = vadd(x, step);
    bool ok = true;
    for (uint32_t i = 0; i < k && ok; ++i) {
      const uint32_t f = m.vf[(size_t)v * kVF + i];
      const Tri t = F(m, f);
      V p[3];
      for (int j = 0; j < 3; ++j) {
        const uint32_t w = t.v[j];
        p[j] = w == v ? y : (w < v ? load3d(low, w) : load3d(start, w));
      }
      ok = !folds(load3d(normal0, f), tri_normal(p[0], p[1], p[2]));
    }
    if (ok) {
      *out = y;
      return h;
    }
    step = vscale(step, 0.5);
  }
  *out = x;
  return 4;
}

// One Jacobi pass of the guarded moves: next[v] from cur (lower neighbours) and start (higher).
__global__ __launch_bounds__(256) void rm_smooth_iterate(Mesh m, uint32_t nv, const uint8_t* mover,
                                                         const double* start, const double* cur, const double* normal0,
                                                         const double* u, double* next, uint8_t* halv) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv || !mover[v]) return;
  V y;
  const int h = guarded_move(m, v, start, cur, normal0, u, &y);
  halv[v] = (uint8_t)h;
  const V c = load3d(cur, v);
  next[3 * (size_t)v] = y.x, next[3 * (size_t)v + 1] = y.y, next[3 * (size_t)v + 2] = y.z;
  if (__double_as_longlong(c.x) != __double_as_longlong(y.x) || __double_as_longlong(c.y) != __double_as_longlong(y.y) ||
      __double_as_longlong(c.z) != __double_as_longlong(y.z))
    atomicAdd(&m.cnt->changed, 1ull);
}
// The sequential order itself (fallback): vertices in ascending order, each from the final
// positions of the lower ones.
__global__ void rm_smooth_sequential(Mesh m, uint32_t nv, const uint8_t* mover, const double* start,
                                     const double* normal0, const double* u, double* cur, uint8_t* halv) {
  if (threadIdx.x != 0) return;
  for (uint32_t v = 0; v < nv; ++v) {
    if (!mover[v]) continue;
    V y;
    halv[v] = (uint8_t)guarded_move(m, v, start, cur, normal0, u, &y);
    cur[3 * (size_t)v] = y.x, cur[3 * (size_t)v + 1] = y.y, cur[3 * (size_t)v + 2] = y.z;
  }
}
__global__ __launch_bounds__(256) void rm_smooth_apply(Mesh m, uint32_t nv, const uint8_t* mover, const double* cur,
                                                       const uint8_t* halv) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv || !mover[v]) return;
  const int h = halv[v];
  if (h <= 3) {
    for (int k = 0; k < 3; ++k) m.pos[3 * (size_t)v + k] = cur[3 * (size_t)v + k];
    atomicAdd(&m.cnt->smoothing_moves, 1ull);
    if (h > 0) atomicAdd(&m.cnt->smoothing_halved, 1ull);
  } else {
    atomicAdd(&m.cnt->smoothing_refused, 1ull);
  }
}
// The optimistic start of the iteration: every mover at its full step.
__global__ __launch_bounds__(256) void rm_smooth_start(uint32_t nv, const uint8_t* mover, const double* start,
                                                       const double* u, double* cur) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  for (int k = 0; k < 3; ++k)
    cur[3 * (size_t)v + k] = mover[v] ? start[3 * (size_t)v + k] + u[3 * (size_t)v + k] : start[3 * (size_t)v + k];
}

// ---- finish (mesh.rs:991) --------------------------------------------------------------------

__global__ __launch_bounds__(256) void rm_flags_u32(uint32_t n, const uint8_t* a, uint32_t* out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = a[i] ? 1u : 0u;
}
__global__ __launch_bounds__(256) void rm_map_len(Mesh m, uint32_t nv, uint32_t* len) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v < nv) len[v] = m.valive[v] ? m.src_len[v] : 0u;
}
__global__ __launch_bounds__(256) void rm_out_vertices(Mesh m, uint32_t nv, const uint32_t* vid, const uint32_t* moff,
                                                       double* pos, uint32_t* map_off, uint32_t* map_src,
                                                       double* map_w, uint32_t* pin_new) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv || !m.valive[v]) return;
  const uint32_t id = vid[v], o = moff[v];
  for (int k = 0; k < 3; ++k) pos[3 * (size_t)id + k] = m.pos[3 * (size_t)v + k];
  map_off[id] = o;
  const uint32_t so = m.src_off[v], sl = m.src_len[v];
  for (uint32_t i = 0; i < sl; ++i) {
    map_src[o + i] = m.pool_v[so + i];
    map_w[o + i] = m.pool_w[so + i];
  }
  if (m.pin_origin[v] != kNone) pin_new[m.pin_origin[v]] = id;
}
__global__ __launch_bounds__(256) void rm_out_faces(Mesh m, uint32_t nf, const uint32_t* vid, const uint32_t* fid,
                                                    uint32_t* faces, uint32_t* parents, uint32_t* bad) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf || !m.falive[f]) return;
  const Tri t = F(m, f);
  const uint32_t id = fid[f];
  bool dead = false;
  for (int k = 0; k < 3; ++k) {
    dead = dead || !m.valive[t.v[k]];
    faces[3 * (size_t)id + k] = vid[t.v[k]];
  }
  parents[id] = m.parent[f];
  if (dead || degenerate(P(m, t.v[0]), P(m, t.v[1]), P(m, t.v[2]))) atomicMin(bad, id);
}

}  // namespace

// ---- host side -------------------------------------------------------------------------------

namespace {

struct MeshBufs {
  DevBuf<double> pos, pool_w;
  DevBuf<uint32_t> faces, parent, vfn, vf, pin_origin, src_off, src_len, pool_v;
  DevBuf<uint8_t> falive, valive, pinned;
  DevBuf<Counters> cnt;
  uint32_t vcap = 0, fcap = 0, pcap = 0;
  void alloc(uint32_t vc, uint32_t fc, uint32_t pc) {
    vcap = vc, fcap = fc, pcap = pc;
    pos.alloc(3 * (size_t)vc);
    faces.alloc(3 * (size_t)fc);
    falive.alloc(fc);
    parent.alloc(fc);
    vfn.alloc(vc);
    vf.alloc((size_t)vc * kVF);
    valive.alloc(vc);
    pinned.alloc(vc);
    pin_origin.alloc(vc);
    src_off.alloc(vc);
    src_len.alloc(vc);
    pool_v.alloc(pc);
    pool_w.alloc(pc);
    cnt.alloc(1);
  }
  Mesh view() {
    return Mesh{pos.ptr,    faces.ptr,      falive.ptr,  parent.ptr,  vfn.ptr,    vf.ptr, valive.ptr,
                pinned.ptr, pin_origin.ptr, src_off.ptr, src_len.ptr, pool_v.ptr, pool_w.ptr, vcap,
                fcap,       pcap,           cnt.ptr};
  }
};

template <class T>
void scan_ex(ddgrs_ctx* ctx, DevBuf<uint8_t>& temp, const T* in, T* out, size_t n) {
  size_t tb = 0;
  HIPCHECK(rocprim::exclusive_scan(nullptr, tb, in, out, T(0), n, rocprim::plus<T>(), ctx->stream));
  temp.reserve(tb);
  tb = temp.n;
  HIPCHECK(rocprim::exclusive_scan(temp.ptr, tb, in, out, T(0), n, rocprim::plus<T>(), ctx->stream));
}

Counters read_counters(ddgrs_ctx* ctx, MeshBufs& M) {
  auto* h = (Counters*)ctx->staging(sizeof(Counters));
  download(ctx, h, M.cnt.ptr, 1);
  ctx->sync();
  return *h;
}

void check_flags(const Counters& c) {
  if (!c.flags) return;
  std::string what;
  if (c.flags & F_VALENCE) what += " a vertex would have more than 32 faces;";
  if (c.flags & F_POOL) what += " the vertex-map pool is full;";
  if (c.flags & F_HEAP) what += " the collapse queue is full;";
  if (c.flags & F_EDGE) what += " an edge has more than 8 faces;";
  if (c.flags & F_SLOTS) what += " the vertex or face slots are full;";
  if (c.flags & F_FIFO) what += " the flip queue is full;";
  throw Error(DDGRS_GPU, "remeshing", "device mesh capacity:" + what);
}

// The current mesh's edges in lexicographic order with their face counts; off[v] (nv + 1 entries)
// the range of the edges whose lower end is v.
struct EdgeTable {
  DevBuf<uint32_t> count, off, ea, eb;
  DevBuf<uint8_t> ec;
  uint32_t ne = 0;
};

uint32_t build_edges(ddgrs_ctx* ctx, Mesh m, uint32_t nv, EdgeTable& E, DevBuf<uint8_t>& temp) {
  E.count.reserve((size_t)nv + 1);
  E.off.reserve((size_t)nv + 1);
  HIPCHECK(hipMemsetAsync(E.count.ptr + nv, 0, sizeof(uint32_t), ctx->stream));
  launch(ctx, "rm_edges", rm_edges, dim3(blocks_for(nv, 256)), dim3(256), 0, m, nv, 0, E.count.ptr,
         (const uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint8_t*)nullptr);
  scan_ex(ctx, temp, (const uint32_t*)E.count.ptr, E.off.ptr, (size_t)nv + 1);
  auto* h = (uint32_t*)ctx->staging(sizeof(uint32_t));
  download(ctx, h, E.off.ptr + nv, 1);
  ctx->sync();
  E.ne = *h;
  E.ea.reserve(std::max<uint32_t>(E.ne, 1));
  E.eb.reserve(std::max<uint32_t>(E.ne, 1));
  E.ec.reserve(std::max<uint32_t>(E.ne, 1));
  launch(ctx, "rm_edges", rm_edges, dim3(blocks_for(nv, 256)), dim3(256), 0, m, nv, 1, E.count.ptr,
         (const uint32_t*)E.off.ptr, E.ea.ptr, E.eb.ptr, E.ec.ptr);
  return E.ne;
}

// Indices e < n with flags[e] != 0, ascending, into out (returns their count).
uint32_t compact(ddgrs_ctx* ctx, const uint32_t* flags, uint32_t n, DevBuf<uint32_t>& nz, DevBuf<uint32_t>& off,
                 DevBuf<uint32_t>& out, DevBuf<uint8_t>& temp) {
  nz.reserve((size_t)n + 1);
  off.reserve((size_t)n + 1);
  HIPCHECK(hipMemsetAsync(nz.ptr + n, 0, sizeof(uint32_t), ctx->stream));
  launch(ctx, "rm_nonzero", rm_nonzero, dim3(blocks_for(n, 256)), dim3(256), 0, n, flags, nz.ptr);
  scan_ex(ctx, temp, (const uint32_t*)nz.ptr, off.ptr, (size_t)n + 1);
  auto* h = (uint32_t*)ctx->staging(sizeof(uint32_t));
  download(ctx, h, off.ptr + n, 1);
  ctx->sync();
  const uint32_t k = *h;
  out.reserve(std::max<uint32_t>(k, 1));
  launch(ctx, "rm_compact", rm_compact, dim3(blocks_for(n, 256)), dim3(256), 0, n, (const uint32_t*)nz.ptr,
         (const uint32_t*)off.ptr, out.ptr);
  return k;
}

void validate_options(const ddgrs_remesh_options& o) {
  auto num = [](double x) { return std::to_string(x); };
  if (!(std::isfinite(o.split) && o.split > 0.0)) invalid("split", num(o.split), "must be finite and positive");
  if (!(std::isfinite(o.collapse) && o.collapse >= 0.0))
    invalid("collapse", num(o.collapse), "must be finite and non-negative");
  if (o.split <= 2.0 * o.collapse)
    invalid("split", num(o.split), "must exceed 2 x collapse (the halves of a split edge must not be collapsible)");
  if (!(o.rho > 0.0 && o.rho <= 1.0)) invalid("rho", num(o.rho), "must lie in (0, 1]");
  if (o.flip_cap_per_vertex < 1) invalid("flip_cap_per_vertex", std::to_string(o.flip_cap_per_vertex), "must be at least 1");
  if (std::isnan(o.max_growth) || o.max_growth < 1.0)
    invalid("max_growth", num(o.max_growth), "must be at least 1 (f64::INFINITY disables it)");
  if (o.rounds < 0) invalid("rounds", std::to_string(o.rounds), "must be non-negative");
  if (o.order != 0 && o.order != 1) invalid("order", std::to_string(o.order), "0 (flip then smooth) or 1");
  if (!(o.guard_gap >= 0.0 && std::isfinite(o.guard_gap))) invalid("guard_gap", num(o.guard_gap), "finite, >= 0");
}

// Uploads and checks the input; builds the device mesh with room for `extra_v` vertices and
// `extra_f` faces.  Checks the positions as Surface::validate_positions (finite, no degenerate face).
void load_mesh(ddgrs_ctx* ctx, const RemeshInput& in, uint32_t extra_v, uint32_t extra_f, uint32_t pool_extra,
               MeshBufs& M, DevBuf<uint32_t>& faces_in, DevBuf<double>& pos_in) {
  const uint32_t nv = (uint32_t)in.nv, nf = (uint32_t)in.nf;
  M.alloc(nv + extra_v, nf + extra_f, nv + pool_extra);
  upload(ctx, faces_in, in.faces, 3 * (size_t)nf);
  upload(ctx, pos_in, in.positions, 3 * (size_t)nv);
  HIPCHECK(hipMemsetAsync(M.cnt.ptr, 0, sizeof(Counters), ctx->stream));
  Mesh m = M.view();
  const Inputs ins{faces_in.ptr, pos_in.ptr, nv, nf};
  launch(ctx, "rm_init_vertices", rm_init_vertices, dim3(blocks_for(nv, 256)), dim3(256), 0, m, ins);
  launch(ctx, "rm_init_faces", rm_init_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, m, ins);
  launch(ctx, "rm_sort_lists", rm_sort_lists, dim3(blocks_for(nv, 256)), dim3(256), 0, m, nv);
  if (!in.pinned.empty()) {
    DevBuf<uint32_t> pins;
    upload(ctx, pins, in.pinned);
    launch(ctx, "rm_pins", rm_pins, dim3(blocks_for((int64_t)in.pinned.size(), 256)), dim3(256), 0, m,
           (const uint32_t*)pins.ptr, (uint32_t)in.pinned.size());
    ctx->sync();
  }
  Counters c{};
  c.nv = nv;
  c.nf = nf;
  c.pool = nv;
  auto* h = (Counters*)ctx->staging(sizeof(Counters));
  *h = c;
  HIPCHECK(hipMemcpyAsync(M.cnt.ptr, h, sizeof(Counters), hipMemcpyHostToDevice, ctx->stream));
  ctx->sync();
}

void check_input(const RemeshInput& in) {
  if (in.nv <= 0 || in.nf <= 0) invalid("surface", std::to_string(in.nv) + " vertices, " + std::to_string(in.nf) + " faces", "a non-empty mesh");
  if (in.nv >= (int64_t)(1u << 30) || in.nf >= (int64_t)(1u << 29)) invalid("surface", std::to_string(in.nf), "fewer than 2^29 faces");
  for (int64_t i = 0; i < 3 * in.nf; ++i)
    if ((int64_t)in.faces[i] >= in.nv) invalid("faces", std::to_string(in.faces[i]), "vertex ids below nv");
  for (int64_t v = 0; v < in.nv; ++v)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(in.positions[3 * v + k])) {
        Error e(DDGRS_NON_FINITE_POSITION, "remeshing", "vertex " + std::to_string(v) + " has a non-finite coordinate");
        e.a = v;
        throw e;
      }
  for (uint32_t p : in.pinned)
    if ((int64_t)p >= in.nv)
      invalid("pinned", std::to_string(p) + " (the surface has " + std::to_string(in.nv) + " vertices)",
              "every pinned vertex must be a vertex of the surface");
}

// EditMesh::finish: survivors in slot order, the vertex map and the pins; the faces revalidated.
void finish(ddgrs_ctx* ctx, MeshBufs& M, const RemeshInput& in, DevBuf<uint8_t>& temp, ddgrs_remeshed* out) {
  const Counters c = read_counters(ctx, M);
  check_flags(c);
  const uint32_t nvs = c.nv, nfs = c.nf;
  Mesh m = M.view();
  DevBuf<uint32_t> vflag(nvs + 1), vid(nvs + 1), fflag(nfs + 1), fid(nfs + 1), mlen(nvs + 1), moff(nvs + 1);
  HIPCHECK(hipMemsetAsync(vflag.ptr + nvs, 0, 4, ctx->stream));
  HIPCHECK(hipMemsetAsync(fflag.ptr + nfs, 0, 4, ctx->stream));
  HIPCHECK(hipMemsetAsync(mlen.ptr + nvs, 0, 4, ctx->stream));
  launch(ctx, "rm_flags_u32", rm_flags_u32, dim3(blocks_for(nvs, 256)), dim3(256), 0, nvs,
         (const uint8_t*)M.valive.ptr, vflag.ptr);
  launch(ctx, "rm_flags_u32", rm_flags_u32, dim3(blocks_for(nfs, 256)), dim3(256), 0, nfs,
         (const uint8_t*)M.falive.ptr, fflag.ptr);
  launch(ctx, "rm_map_len", rm_map_len, dim3(blocks_for(nvs, 256)), dim3(256), 0, m, nvs, mlen.ptr);
  scan_ex(ctx, temp, (const uint32_t*)vflag.ptr, vid.ptr, (size_t)nvs + 1);
  scan_ex(ctx, temp, (const uint32_t*)fflag.ptr, fid.ptr, (size_t)nfs + 1);
  scan_ex(ctx, temp, (const uint32_t*)mlen.ptr, moff.ptr, (size_t)nvs + 1);
  auto* h = (uint32_t*)ctx->staging(3 * sizeof(uint32_t));
  download(ctx, h, vid.ptr + nvs, 1);
  download(ctx, h + 1, fid.ptr + nfs, 1);
  download(ctx, h + 2, moff.ptr + nvs, 1);
  ctx->sync();
  const uint32_t nv = h[0], nf = h[1], nm = h[2];
  DevBuf<double> pos(3 * (size_t)nv + 1), mw((size_t)nm + 1);
  DevBuf<uint32_t> faces(3 * (size_t)nf + 1), parents((size_t)nf + 1), moffo((size_t)nv + 1), msrc((size_t)nm + 1),
      pin_new((size_t)in.nv + 1), bad(1);
  HIPCHECK(hipMemsetAsync(pin_new.ptr, 0xff, ((size_t)in.nv + 1) * 4, ctx->stream));
  HIPCHECK(hipMemsetAsync(bad.ptr, 0xff, 4, ctx->stream));
  launch(ctx, "rm_out_vertices", rm_out_vertices, dim3(blocks_for(nvs, 256)), dim3(256), 0, m, nvs,
         (const uint32_t*)vid.ptr, (const uint32_t*)moff.ptr, pos.ptr, moffo.ptr, msrc.ptr, mw.ptr, pin_new.ptr);
  launch(ctx, "rm_out_faces", rm_out_faces, dim3(blocks_for(nfs, 256)), dim3(256), 0, m, nfs, (const uint32_t*)vid.ptr,
         (const uint32_t*)fid.ptr, faces.ptr, parents.ptr, bad.ptr);
  out->positions.resize(3 * (size_t)nv);
  out->faces.resize(3 * (size_t)nf);
  out->parents.resize(nf);
  out->map_offsets.resize((size_t)nv + 1);
  out->map_sources.resize(nm);
  out->map_weights.resize(nm);
  std::vector<uint32_t> pn((size_t)in.nv);
  uint32_t badh = 0;
  download(ctx, out->positions.data(), (const double*)pos.ptr, 3 * (size_t)nv);
  download(ctx, out->faces.data(), (const uint32_t*)faces.ptr, 3 * (size_t)nf);
  download(ctx, out->parents.data(), (const uint32_t*)parents.ptr, nf);
  download(ctx, out->map_offsets.data(), (const uint32_t*)moffo.ptr, nv);
  download(ctx, out->map_sources.data(), (const uint32_t*)msrc.ptr, nm);
  download(ctx, out->map_weights.data(), (const double*)mw.ptr, nm);
  download(ctx, pn.data(), (const uint32_t*)pin_new.ptr, (size_t)in.nv);
  download(ctx, &badh, (const uint32_t*)bad.ptr, 1);
  ctx->sync();
  out->map_offsets[nv] = nm;
  if (badh != kNone) {
    Error e(DDGRS_DEGENERATE_FACE, "remeshing",
            "internal: face " + std::to_string(badh) + " of the remeshed surface is degenerate or dangling");
    e.a = badh;
    throw e;
  }
  out->pinned.clear();
  for (uint32_t p : in.pinned) out->pinned.push_back(pn[p]);
}

// Degenerate input faces (Surface::validate_positions): the lowest index, or kNone.
__global__ __launch_bounds__(256) void rm_check_faces(Mesh m, uint32_t nf, uint32_t* bad) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf) return;
  const Tri t = F(m, f);
  if (degenerate(P(m, t.v[0]), P(m, t.v[1]), P(m, t.v[2]))) atomicMin(bad, f);
}
void check_degenerate(ddgrs_ctx* ctx, MeshBufs& M, uint32_t nf) {
  DevBuf<uint32_t> bad(1);
  HIPCHECK(hipMemsetAsync(bad.ptr, 0xff, 4, ctx->stream));
  launch(ctx, "rm_check_faces", rm_check_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, M.view(), nf, bad.ptr);
  auto* h = (uint32_t*)ctx->staging(sizeof(uint32_t));
  download(ctx, h, (const uint32_t*)bad.ptr, 1);
  ctx->sync();
  if (*h != kNone) {
    Error e(DDGRS_DEGENERATE_FACE, "remeshing", "face " + std::to_string(*h) + " has zero or non-finite area");
    e.a = *h;
    throw e;
  }
}

}  // namespace

void remesh_pass(ddgrs_ctx* ctx, const RemeshInput& in, double l0, const ddgrs_remesh_options& o,
                 ddgrs_remeshed* out, ddgrs_remesh_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  validate_options(o);
  if (!(std::isfinite(l0) && l0 > 0.0))
    invalid("l0", std::to_string(l0), "the target edge length must be finite and positive");
  check_input(in);
  *rep = ddgrs_remesh_report{};
  DevBuf<uint8_t> temp;
  DevBuf<uint32_t> faces_in;
  DevBuf<double> pos_in;
  const uint32_t nv0 = (uint32_t)in.nv, nf0 = (uint32_t)in.nf;

  // The input's edges and its long ones size the device mesh: every split adds one vertex and at
  // most two faces.  A first mesh without room for splits gives the edge table.
  MeshBufs M;
  EdgeTable E;
  DevBuf<uint32_t> is_long, nz, off, list, keep;
  DevBuf<double> len;
  uint32_t nlong = 0;
  {
    MeshBufs M0;
    load_mesh(ctx, in, 0, 0, 0, M0, faces_in, pos_in);
    check_degenerate(ctx, M0, nf0);
    check_flags(read_counters(ctx, M0));
    if (o.split_on) {
      const uint32_t ne = build_edges(ctx, M0.view(), nv0, E, temp);
      is_long.reserve(std::max<uint32_t>(ne, 1));
      len.reserve(std::max<uint32_t>(ne, 1));
      launch(ctx, "rm_long", rm_long, dim3(blocks_for(ne, 256)), dim3(256), 0, M0.view(), ne, (const uint32_t*)E.ea.ptr,
             (const uint32_t*)E.eb.ptr, (const uint8_t*)E.ec.ptr, o.split * l0, is_long.ptr, len.ptr);
      nlong = compact(ctx, is_long.ptr, ne, nz, off, list, temp);
    }
  }
  // pool: the input's entries, two per split, and room for the collapses' combined entries
  const uint32_t pool_extra = 2 * nlong + 4 * nv0 + 65536;
  load_mesh(ctx, in, nlong, 2 * nlong, pool_extra, M, faces_in, pos_in);
  Mesh m = M.view();

  if (o.split_on && nlong > 0) {
    // the growth budget (mod.rs:260): only when the long edges would exceed it
    DevBuf<uint32_t> added(nlong), order_in(nlong), order(nlong);
    DevBuf<unsigned long long> key_in(nlong), key_out(nlong);
    std::vector<uint32_t> addh(nlong);
    std::vector<double> lenh(nlong);
    {
      DevBuf<double> lg(nlong);
      launch(ctx, "rm_gather_long", rm_gather_long, dim3(blocks_for(nlong, 256)), dim3(256), 0, nlong,
             (const uint32_t*)list.ptr, (const uint32_t*)is_long.ptr, (const double*)len.ptr, added.ptr, lg.ptr);
      download(ctx, addh.data(), (const uint32_t*)added.ptr, nlong);
      download(ctx, lenh.data(), (const double*)lg.ptr, nlong);
      ctx->sync();
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < nlong; ++i) total += addh[i];
    const double budget = (o.max_growth - 1.0) * (double)nf0;
    bool cut = (double)total > budget;
    if (cut) {
      // longest first (ties by edge order): radix sort of the complemented length bits, stable
      std::vector<unsigned long long> kh(nlong);
      std::vector<uint32_t> oh(nlong);
      for (uint32_t i = 0; i < nlong; ++i) {
        uint64_t b;
        std::memcpy(&b, &lenh[i], 8);
        kh[i] = ~b;
        oh[i] = i;
      }
      upload(ctx, key_in, kh);
      upload(ctx, order_in, oh);
      size_t tb = 0;
      HIPCHECK(rocprim::radix_sort_pairs(nullptr, tb, key_in.ptr, key_out.ptr, order_in.ptr, order.ptr, nlong, 0, 64,
                                         ctx->stream));
      temp.reserve(tb);
      tb = temp.n;
      HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, key_in.ptr, key_out.ptr, order_in.ptr, order.ptr, nlong, 0, 64,
                                         ctx->stream));
      keep.reserve(nlong);
      launch(ctx, "rm_budget", rm_budget, dim3(1), dim3(64), 0, nlong, (const uint32_t*)order.ptr,
             (const uint32_t*)added.ptr, budget, keep.ptr, M.cnt.ptr);
    }
    launch(ctx, "rm_split_engine", rm_split_engine, dim3(1), dim3(64), 0, m, (const uint32_t*)E.ea.ptr,
           (const uint32_t*)E.eb.ptr, (const uint32_t*)list.ptr, cut ? (const uint32_t*)keep.ptr : nullptr, nlong);
    check_flags(read_counters(ctx, M));
  }

  if (o.collapse_on) {
    Counters c = read_counters(ctx, M);
    const uint32_t ne = build_edges(ctx, m, c.nv, E, temp);
    DevBuf<uint32_t> is_short(std::max<uint32_t>(ne, 1)), sidx_out;
    DevBuf<unsigned long long> bits(std::max<uint32_t>(ne, 1));
    const double thr = o.collapse * l0;
    launch(ctx, "rm_short", rm_short, dim3(blocks_for(ne, 256)), dim3(256), 0, m, ne, (const uint32_t*)E.ea.ptr,
           (const uint32_t*)E.eb.ptr, thr, is_short.ptr, bits.ptr);
    const uint32_t ns = compact(ctx, is_short.ptr, ne, nz, off, list, temp);
    DevBuf<unsigned long long> skeys(std::max<uint32_t>(ns, 1)), skeys_out(std::max<uint32_t>(ns, 1));
    sidx_out.reserve(std::max<uint32_t>(ns, 1));
    if (ns > 0) {
      launch(ctx, "rm_gather_keys", rm_gather_keys, dim3(blocks_for(ns, 256)), dim3(256), 0, ns,
             (const uint32_t*)list.ptr, (const unsigned long long*)bits.ptr, skeys.ptr);
      size_t tb = 0;
      HIPCHECK(rocprim::radix_sort_pairs(nullptr, tb, skeys.ptr, skeys_out.ptr, list.ptr, sidx_out.ptr, ns, 0, 64,
                                         ctx->stream));
      temp.reserve(tb);
      tb = temp.n;
      HIPCHECK(rocprim::radix_sort_pairs(temp.ptr, tb, skeys.ptr, skeys_out.ptr, list.ptr, sidx_out.ptr, ns, 0, 64,
                                         ctx->stream));
    }
    const uint32_t heap_cap = ns + 8 * c.nv + 4096;
    DevBuf<HeapEntry> heap(heap_cap);
    CollapseArgs A{skeys_out.ptr, sidx_out.ptr, ns, E.ea.ptr, E.eb.ptr, heap.ptr, heap_cap, thr, o.guard_gap,
                   nullptr, 0};
    DevBuf<uint8_t> band;
    if (o.guard_gap > 0.0 && in.band) {
      upload(ctx, band, in.band, (size_t)in.nv);
      A.band = band.ptr;
      A.nband = (uint32_t)in.nv;
    }
    if (ns > 0)
      launch(ctx, "rm_collapse_engine", rm_collapse_engine, dim3(1), dim3(kGuardThreads), 0, m, A);
    check_flags(read_counters(ctx, M));
  }

  DevBuf<uint32_t> viol, rev, qn, q;
  DevBuf<uint8_t> rflag, bnd, fob, mover, halv;
  DevBuf<unsigned long long> fifo;
  DevBuf<double> normal0, area, center, u, cur, nxt;
  auto flip_round = [&]() {
    const Counters c = read_counters(ctx, M);
    const uint32_t ne = build_edges(ctx, m, c.nv, E, temp);
    viol.reserve(std::max<uint32_t>(ne, 1));
    launch(ctx, "rm_violating", rm_violating, dim3(blocks_for(ne, 256)), dim3(256), 0, m, ne,
           (const uint32_t*)E.ea.ptr, (const uint32_t*)E.eb.ptr, viol.ptr);
    const uint32_t n0 = compact(ctx, viol.ptr, ne, nz, off, list, temp);
    if (n0 == 0) return;
    rev.reserve(ne + 1);
    rflag.reserve(ne + 1);
    fifo.reserve(ne + 64);
    qn.reserve(c.nv);
    q.reserve((size_t)c.nv * kNb);
    HIPCHECK(hipMemsetAsync(rflag.ptr, 0, ne + 1, ctx->stream));
    HIPCHECK(hipMemsetAsync(qn.ptr, 0, (size_t)c.nv * 4, ctx->stream));
    const uint64_t live = (uint64_t)nv0 + c.splits - c.collapses;
    FlipArgs A{E.ea.ptr, E.eb.ptr, E.off.ptr, ne, list.ptr, n0, rev.ptr, rflag.ptr, fifo.ptr, ne + 64, qn.ptr, q.ptr,
               (unsigned long long)o.flip_cap_per_vertex * live};
    launch(ctx, "rm_flip_engine", rm_flip_engine, dim3(1), dim3(64), 0, m, A);
    check_flags(read_counters(ctx, M));
  };
  auto smooth_round = [&]() {
    const Counters c = read_counters(ctx, M);
    const uint32_t nv = c.nv, nf = c.nf;
    const uint32_t ne = build_edges(ctx, m, nv, E, temp);
    bnd.reserve(nv);
    fob.reserve(nf);
    HIPCHECK(hipMemsetAsync(bnd.ptr, 0, nv, ctx->stream));
    HIPCHECK(hipMemsetAsync(fob.ptr, 0, nf, ctx->stream));
    launch(ctx, "rm_boundary", rm_boundary, dim3(blocks_for(ne, 256)), dim3(256), 0, m, ne, (const uint32_t*)E.ea.ptr,
           (const uint32_t*)E.eb.ptr, (const uint8_t*)E.ec.ptr, bnd.ptr, fob.ptr);
    normal0.reserve(3 * (size_t)nf);
    area.reserve(nf);
    center.reserve(3 * (size_t)nf);
    launch(ctx, "rm_face_geom", rm_face_geom, dim3(blocks_for(nf, 256)), dim3(256), 0, m, nf, (const uint8_t*)fob.ptr,
           normal0.ptr, area.ptr, center.ptr);
    mover.reserve(nv);
    halv.reserve(nv);
    u.reserve(3 * (size_t)nv);
    cur.reserve(3 * (size_t)nv);
    nxt.reserve(3 * (size_t)nv);
    launch(ctx, "rm_smooth_moves", rm_smooth_moves, dim3(blocks_for(nv, 256)), dim3(256), 0, m, nv, o.rho,
           (const uint8_t*)bnd.ptr, (const double*)normal0.ptr, (const double*)area.ptr, (const double*)center.ptr,
           mover.ptr, u.ptr);
    launch(ctx, "rm_smooth_start", rm_smooth_start, dim3(blocks_for(nv, 256)), dim3(256), 0, nv,
           (const uint8_t*)mover.ptr, (const double*)M.pos.ptr, (const double*)u.ptr, cur.ptr);
    HIPCHECK(hipMemcpyAsync(nxt.ptr, cur.ptr, 3 * (size_t)nv * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    bool converged = false;
    for (int it = 0; it < kMaxSmoothIterations; ++it) {
      HIPCHECK(hipMemsetAsync(&M.cnt.ptr->changed, 0, sizeof(unsigned long long), ctx->stream));
      launch(ctx, "rm_smooth_iterate", rm_smooth_iterate, dim3(blocks_for(nv, 256)), dim3(256), 0, m, nv,
             (const uint8_t*)mover.ptr, (const double*)M.pos.ptr, (const double*)cur.ptr, (const double*)normal0.ptr,
             (const double*)u.ptr, nxt.ptr, halv.ptr);
      std::swap(cur, nxt);
      ++rep->smoothing_iterations;
      if (read_counters(ctx, M).changed == 0) {
        converged = true;
        break;
      }
    }
    if (!converged)
      launch(ctx, "rm_smooth_sequential", rm_smooth_sequential, dim3(1), dim3(64), 0, m, nv, (const uint8_t*)mover.ptr,
             (const double*)M.pos.ptr, (const double*)normal0.ptr, (const double*)u.ptr, cur.ptr, halv.ptr);
    launch(ctx, "rm_smooth_apply", rm_smooth_apply, dim3(blocks_for(nv, 256)), dim3(256), 0, m, nv,
           (const uint8_t*)mover.ptr, (const double*)cur.ptr, (const uint8_t*)halv.ptr);
  };
  for (int r = 0; r < o.rounds; ++r) {
    if (o.order == 0) {
      if (o.flip_on) flip_round();
      if (o.smooth_on) smooth_round();
    } else {
      if (o.smooth_on) smooth_round();
      if (o.flip_on) flip_round();
    }
  }
  const Counters c = read_counters(ctx, M);
  finish(ctx, M, in, temp, out);
  rep->splits = (int64_t)c.splits;
  rep->collapses = (int64_t)c.collapses;
  rep->flips = (int64_t)c.flips;
  rep->splits_deferred = (int64_t)c.splits_deferred;
  for (int k = 0; k < 11; ++k) {
    rep->refused_splits[k] = (int64_t)c.refused_split[k];
    rep->refused_collapses[k] = (int64_t)c.refused_collapse[k];
    rep->refused_flips[k] = (int64_t)c.refused_flip[k];
  }
  rep->smoothing_moves = (int64_t)c.smoothing_moves;
  rep->smoothing_halved = (int64_t)c.smoothing_halved;
  rep->smoothing_refused = (int64_t)c.smoothing_refused;
  rep->smoothing_skipped = (int64_t)c.smoothing_skipped;
  rep->flip_cap_hits = (int64_t)c.flip_cap_hits;
  rep->n_vertices = (int64_t)out->positions.size() / 3;
  rep->n_faces = (int64_t)out->faces.size() / 3;
  rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

namespace {
__global__ void rm_single_op(Mesh m, int op, uint32_t a, uint32_t b, int delaunay, int foldover, int* result) {
  if (threadIdx.x >= 64) return;
  int r = R_OK;
  bool reo = false;
  if (op == 0) {
    r = split_edge(m, a, b);
  } else if (op == 1) {
    Collapse C;
    r = collapse_check(m, a, b, foldover != 0, C);
    if (r == R_OK) r = collapse_apply(m, C);
  } else {
    r = flip_edge(m, a, b, delaunay != 0, foldover != 0, &reo);
  }
  result[0] = r;
  result[1] = reo ? 1 : 0;
}
}  // namespace

int remesh_single_op(ddgrs_ctx* ctx, const RemeshInput& in, int op, uint32_t a, uint32_t b, bool delaunay,
                     bool foldover, bool* reordered, ddgrs_remeshed* out) {
  check_input(in);
  if (op < 0 || op > 2) invalid("op", std::to_string(op), "0 split, 1 collapse, 2 flip");
  DevBuf<uint32_t> faces_in;
  DevBuf<double> pos_in;
  MeshBufs M;
  load_mesh(ctx, in, 1, 2, 16, M, faces_in, pos_in);
  check_degenerate(ctx, M, (uint32_t)in.nf);
  DevBuf<int> res(2);
  launch(ctx, "rm_single_op", rm_single_op, dim3(1), dim3(64), 0, M.view(), op, a, b, delaunay ? 1 : 0,
         foldover ? 1 : 0, res.ptr);
  int h[2];
  download(ctx, h, (const int*)res.ptr, 2);
  ctx->sync();
  DevBuf<uint8_t> temp;
  finish(ctx, M, in, temp, out);
  *reordered = h[1] != 0;
  return h[0];
}

}  // namespace ddgrs
