// FP64 geometry shared by the collision detection and the remesher's collapse guard: vectors,
// the curves crate's distances (crates/ddg-repulsive-curves/src/ccd.rs:236-320) and the static
// intersection test (crates/ddg-repulsive-surfaces/src/ccd.rs:1153-1193), every expression in the
// reference's order.  Included by translation units compiled without contraction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#pragma clang fp contract(off)

namespace ddgrs {
namespace {

struct V {
  double x, y, z;
};
__device__ __forceinline__ V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V vscale(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V vcross(V a, V b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double vnorm(V a) { return sqrt(vdot(a, a)); }
__device__ __forceinline__ double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
__device__ __forceinline__ V at(V p, V u, double t) { return {p.x + t * u.x, p.y + t * u.y, p.z + t * u.z}; }
__device__ __forceinline__ V load3(const double* a, uint32_t i) { return {a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]}; }
__device__ __forceinline__ V vel3(const double* a, double s, uint32_t i) {
  return {s * a[3 * (size_t)i], s * a[3 * (size_t)i + 1], s * a[3 * (size_t)i + 2]};
}

// ddg-repulsive-curves/src/ccd.rs:236
__device__ double point_segment_distance(V p, V a, V b) {
  const V ab = vsub(b, a), ap = vsub(p, a);
  const double e = vdot(ab, ab);
  const double t = e > 0.0 ? clamp01(vdot(ap, ab) / e) : 0.0;
  return vnorm(vsub(ap, vscale(ab, t)));
}

// ddg-repulsive-curves/src/ccd.rs:266
__device__ double segment_segment_distance(V p1, V q1, V p2, V q2) {
  double best = fmin(fmin(fmin(point_segment_distance(p1, p2, q2), point_segment_distance(q1, p2, q2)),
                          point_segment_distance(p2, p1, q1)),
                     point_segment_distance(q2, p1, q1));
  const V d1 = vsub(q1, p1), d2 = vsub(q2, p2);
  const double a = vdot(d1, d1), e = vdot(d2, d2);
  const V n = vcross(d1, d2);
  const double nn = vdot(n, n);
  if (a > 0.0 && e > 0.0 && nn > 0.0) {
    const V r = vsub(p1, p2);
    const double b = vdot(d1, d2), c = vdot(d1, r), f = vdot(d2, r);
    double s = clamp01(vdot(vcross(vsub(p2, p1), d2), n) / nn);
    const double t = clamp01((b * s + f) / e);
    s = clamp01((b * t - c) / a);
    const V c1 = vadd(p1, vscale(d1, s)), c2 = vadd(p2, vscale(d2, t));
    best = fmin(best, vnorm(vsub(c1, c2)));
  }
  return best;
}

// ddg-repulsive-curves/src/ccd.rs:298
__device__ double point_triangle_distance(V p, V a, V b, V c) {
  const double edges =
      fmin(fmin(point_segment_distance(p, a, b), point_segment_distance(p, b, c)), point_segment_distance(p, c, a));
  const V ab = vsub(b, a), ac = vsub(c, a);
  const V n = vcross(ab, ac);
  const double nn = vdot(n, n);
  if (nn <= 0.0) return edges;
  const V ap = vsub(p, a);
  const double v = vdot(vcross(ap, ac), n) / nn;
  const double w = vdot(vcross(ab, ap), n) / nn;
  if (v >= 0.0 && w >= 0.0 && v + w <= 1.0) {
    const double h = fabs(vdot(ap, n)) / sqrt(nn);
    return fmin(edges, h);
  }
  return edges;
}

// ccd.rs:1153-1186
__device__ __forceinline__ double orient(V a, V b, V c, V d) { return vdot(vcross(vsub(b, a), vsub(c, a)), vsub(d, a)); }

__device__ bool segment_meets_triangle(V p, V q, V a, V b, V c) {
  const double d1 = orient(a, b, c, p), d2 = orient(a, b, c, q);
  if ((d1 > 0.0 && d2 > 0.0) || (d1 < 0.0 && d2 < 0.0)) return false;
  const double n = vnorm(vcross(vsub(b, a), vsub(c, a)));
  const double len = fmax(fmax(fmax(vnorm(vsub(q, p)), vnorm(vsub(b, a))), vnorm(vsub(c, a))), vnorm(vsub(c, b)));
  const double flat = 1e-12 * n * len;
  if (fabs(d1) <= flat && fabs(d2) <= flat) {
    const double eps = 1e-14 * (vnorm(vsub(b, a)) + vnorm(vsub(c, a)));
    return point_triangle_distance(p, a, b, c) <= eps || point_triangle_distance(q, a, b, c) <= eps ||
           segment_segment_distance(p, q, a, b) <= eps || segment_segment_distance(p, q, b, c) <= eps ||
           segment_segment_distance(p, q, c, a) <= eps;
  }
  const double s1 = orient(p, q, a, b), s2 = orient(p, q, b, c), s3 = orient(p, q, c, a);
  return (s1 >= 0.0 && s2 >= 0.0 && s3 >= 0.0) || (s1 <= 0.0 && s2 <= 0.0 && s3 <= 0.0);
}

// ccd.rs:1188
__device__ bool triangles_intersect(const V* t1, const V* t2) {
  for (int k = 0; k < 3; ++k)
    if (segment_meets_triangle(t1[k], t1[(k + 1) % 3], t2[0], t2[1], t2[2])) return true;
  for (int k = 0; k < 3; ++k)
    if (segment_meets_triangle(t2[k], t2[(k + 1) % 3], t1[0], t1[1], t1[2])) return true;
  return false;
}

// ccd.rs:1195 `triangle_distance`: 0 for intersecting triangles, else the smallest of the
// vertex–triangle and edge–edge distances.
__device__ double triangle_distance(const V* t1, const V* t2) {
  if (triangles_intersect(t1, t2)) return 0.0;
  double d = INFINITY;
  for (int k = 0; k < 3; ++k) {
    d = fmin(d, point_triangle_distance(t1[k], t2[0], t2[1], t2[2]));
    d = fmin(d, point_triangle_distance(t2[k], t1[0], t1[1], t1[2]));
    for (int j = 0; j < 3; ++j) d = fmin(d, segment_segment_distance(t1[k], t1[(k + 1) % 3], t2[j], t2[(j + 1) % 3]));
  }
  return d;
}

}  // namespace
}  // namespace ddgrs
