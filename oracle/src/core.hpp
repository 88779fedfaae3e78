// FP64 CPU oracle of the Hs-descent path of ddg-repulsive-surfaces — TEST INFRASTRUCTURE ONLY.
//
// This file restates, operation for operation, the reference's kernel, exponent and per-face
// geometry code so that the oracle's roundings are the reference's:
//   crates/ddg-repulsive-surfaces/src/kernel.rs        (Exponents :26-195, kernel_with_gradient_of_v :238,
//                                                      pair_terms :256)
//   crates/ddg-repulsive-curves/src/kernel.rs:12-52    (v3 helpers)
//   crates/ddg-repulsive-surfaces/src/geometry.rs      (face_quantities :33, face_pull_back :87, pull_back :172,
//                                                      vertex_areas :194, cotan_stiffness :221)
// It is compiled with -ffp-contract=off (the reference's Rust code never contracts a*b+c into an
// FMA), so sums and products round exactly as there.  Only the tests, the smoke check and the CPU
// baseline leg of bench.py load it; the GPU library never does.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace orc {

using V3 = std::array<double, 3>;

inline V3 add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 scale(V3 a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
// a + s b
inline V3 axpy(V3 a, double s, V3 b) { return {a[0] + s * b[0], a[1] + s * b[1], a[2] + s * b[2]}; }
inline double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(V3 a, V3 b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// Rust's f64::min / f64::max: a NaN argument yields the other one (IEEE minNum/maxNum).
inline double fmin_(double a, double b) { return std::fmin(a, b); }
inline double fmax_(double a, double b) { return std::fmax(a, b); }

// Error kinds, numbered as the C-ABI of the product library numbers them.
enum Status : int {
  OK = 0,
  INVALID_OPTION = 1,
  INVALID_EXPONENTS = 2,
  NON_FINITE_ENERGY = 3,
  SOLVER_DIVERGED = 4,
  SINGULAR_SADDLE = 5,
  DEGENERATE_FACE = 6,
  NON_FINITE_POSITION = 7,
};

struct Error {
  int status = OK;
  std::string detail;
  int64_t a = -1, b = -1;  // e.g. the face pair of NonFiniteEnergy
};

// kernel.rs:37-39 `as_small_int`
inline int as_small_int(double x, double max) {
  return (x == std::trunc(x) && x > 0.0 && x <= max) ? (int)(uint32_t)x : -1;
}

// kernel.rs:183 `ipow`: binary powering in the reference's exact order.
inline double ipow(double x, uint32_t n) {
  double acc = 1.0;
  while (n > 0) {
    if (n & 1u) acc *= x;
    n >>= 1;
    if (n > 0) x *= x;
  }
  return acc;
}

// kernel.rs:26 `Exponents`.
struct Exponents {
  double alpha = 6.0, beta = 12.0;
  int alpha_int = 6, beta_int = 12;

  // kernel.rs:49 `Exponents::new`: returns false (and the requirement) for invalid exponents.
  static bool make(double alpha, double beta, Exponents* out, std::string* why) {
    if (!std::isfinite(alpha) || !std::isfinite(beta)) {
      *why = "both exponents must be finite";
      return false;
    }
    if (alpha <= 1.0) {
      *why = "alpha must be greater than 1";
      return false;
    }
    if (beta <= alpha + 2.0) {
      *why = "beta must be greater than alpha + 2 (s = (beta - 2)/alpha > 1)";
      return false;
    }
    if (beta >= 2.0 * alpha + 2.0) {
      *why = "beta must be smaller than 2 alpha + 2 (s = (beta - 2)/alpha < 2)";
      return false;
    }
    out->alpha = alpha;
    out->beta = beta;
    out->alpha_int = as_small_int(alpha, 64.0);
    out->beta_int = as_small_int(beta, 128.0);
    return true;
  }
  double s() const { return (beta - 2.0) / alpha; }

  // kernel.rs:149 `nu_powers`: (|nu|^alpha, sign(nu)|nu|^(alpha-1)).
  inline void nu_powers(double nu, double* pa, double* sp) const {
    double a = std::fabs(nu);
    double p1 = alpha_int >= 0 ? ipow(a, (uint32_t)(alpha_int - 1)) : std::pow(a, alpha - 1.0);
    *pa = p1 * a;
    *sp = std::copysign(p1, nu);
  }
  // kernel.rs:160 `r_power`: r^-beta from r^2.
  inline double r_power(double r2) const {
    if (beta_int >= 0) {
      uint32_t k = (uint32_t)beta_int;
      if (k % 2 == 0) return 1.0 / ipow(r2, k / 2);
      return 1.0 / (ipow(r2, k / 2) * std::sqrt(r2));
    }
    return std::pow(r2, -0.5 * beta);
  }
};

// kernel.rs:216 `kernel_of_v`.
inline double kernel_of_v(const Exponents& e, V3 v, V3 n) {
  double pa, sp;
  e.nu_powers(dot(n, v), &pa, &sp);
  return pa * e.r_power(dot(v, v));
}

struct KernelEval {
  double value;
  V3 d_v, d_n;
};

// kernel.rs:238 `kernel_with_gradient_of_v`.
inline KernelEval kernel_with_gradient_of_v(const Exponents& e, V3 v, V3 n) {
  double r2 = dot(v, v);
  double rb = e.r_power(r2);
  double pa, sp1;
  e.nu_powers(dot(n, v), &pa, &sp1);
  double value = pa * rb;
  double c_n = e.alpha * sp1 * rb;
  V3 d_v = axpy(scale(n, c_n), -e.beta * value / r2, v);
  return {value, d_v, scale(v, c_n)};
}

// kernel.rs:256 `pair_terms`.
inline void pair_terms(const Exponents& e, V3 v, V3 n_s, V3 n_t, double* k_st, double* k_ts,
                       V3* dv_st, V3* dv_ts, V3* dn_st) {
  double r2 = dot(v, v);
  double rb = e.r_power(r2);
  double pa_s, sp_s, pa_t, sp_t;
  e.nu_powers(dot(n_s, v), &pa_s, &sp_s);
  e.nu_powers(-dot(n_t, v), &pa_t, &sp_t);
  *k_st = pa_s * rb;
  *k_ts = pa_t * rb;
  double c_s = e.alpha * sp_s * rb;
  double c_t = e.alpha * sp_t * rb;
  double inv_r2 = 1.0 / r2;
  *dv_st = axpy(scale(n_s, c_s), -e.beta * *k_st * inv_r2, v);
  *dv_ts = axpy(scale(n_t, c_t), e.beta * *k_ts * inv_r2, v);
  *dn_st = scale(v, c_s);
}

// ---------------------------------------------------------------------------------------------
// Geometry (geometry.rs)

struct Mesh {
  int64_t nv = 0;
  std::vector<std::array<uint32_t, 3>> faces;
};

using Positions = std::vector<V3>;

inline std::array<V3, 3> corners(const Positions& p, const std::array<uint32_t, 3>& t) {
  return {p[t[0]], p[t[1]], p[t[2]]};
}

// geometry.rs:33 `face_quantities`: (area, unit normal, barycenter).
inline void face_quantities(const std::array<V3, 3>& p, double* a, V3* n_out, V3* x) {
  V3 n = cross(sub(p[1], p[0]), sub(p[2], p[0]));
  double len = norm(n);
  *x = scale(add(add(p[0], p[1]), p[2]), 1.0 / 3.0);
  *a = 0.5 * len;
  *n_out = scale(n, 1.0 / len);
}

// geometry.rs:47 `opposite_edges`.
inline std::array<V3, 3> opposite_edges(const std::array<V3, 3>& p) {
  return {sub(p[2], p[1]), sub(p[0], p[2]), sub(p[1], p[0])};
}

// geometry.rs:87 `face_pull_back`.
inline std::array<V3, 3> face_pull_back(const std::array<V3, 3>& p, double d_area, V3 d_bary,
                                        V3 d_normal) {
  V3 n = cross(sub(p[1], p[0]), sub(p[2], p[0]));
  double len = norm(n);
  V3 nrm = scale(n, 1.0 / len);
  auto e = opposite_edges(p);
  V3 xb = scale(d_bary, 1.0 / 3.0);
  double inv_2a = 1.0 / len;
  std::array<V3, 3> out;
  for (int c = 0; c < 3; ++c) {
    V3 ga = scale(cross(nrm, e[c]), 0.5 * d_area);
    V3 gn = scale(nrm, dot(cross(e[c], nrm), d_normal) * inv_2a);
    out[c] = add(add(ga, xb), gn);
  }
  return out;
}

struct FaceGeometry {
  std::vector<double> areas;
  std::vector<V3> normals, barycenters;
};

inline FaceGeometry face_geometry(const Mesh& m, const Positions& p) {
  FaceGeometry g;
  size_t nf = m.faces.size();
  g.areas.resize(nf);
  g.normals.resize(nf);
  g.barycenters.resize(nf);
  for (size_t f = 0; f < nf; ++f)
    face_quantities(corners(p, m.faces[f]), &g.areas[f], &g.normals[f], &g.barycenters[f]);
  return g;
}

// geometry.rs:172 `pull_back` (faces in face order).
inline std::vector<V3> pull_back(const Mesh& m, const Positions& p, const std::vector<double>& da,
                                 const std::vector<V3>& dx, const std::vector<V3>& dn) {
  std::vector<V3> out(m.nv, V3{0.0, 0.0, 0.0});
  for (size_t f = 0; f < m.faces.size(); ++f) {
    auto g = face_pull_back(corners(p, m.faces[f]), da[f], dx[f], dn[f]);
    for (int c = 0; c < 3; ++c) {
      uint32_t v = m.faces[f][c];
      out[v] = add(out[v], g[c]);
    }
  }
  return out;
}

// geometry.rs:194 `vertex_areas`.
inline std::vector<double> vertex_areas(const Mesh& m, const std::vector<double>& fa) {
  std::vector<double> out(m.nv, 0.0);
  for (size_t f = 0; f < m.faces.size(); ++f) {
    double a = fa[f] / 3.0;
    for (int c = 0; c < 3; ++c) out[m.faces[f][c]] += a;
  }
  return out;
}

// Validation of positions (surface.rs `validate_positions`): counts, finiteness, zero-area faces.
bool validate_positions(const Mesh& m, const Positions& p, Error* err);

// Parallel loop over [0, n) on `threads` threads (0: all cores); every index writes its own
// output, so results do not depend on the thread count.
void parallel_for(int64_t n, int threads, const std::function<void(int64_t)>& f);

}  // namespace orc
