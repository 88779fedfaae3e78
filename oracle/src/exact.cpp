// FP64 oracle — TEST INFRASTRUCTURE ONLY: the exact all-pairs energy and its differential.
//
// Restates crates/ddg-repulsive-surfaces/src/energy/exact.rs:
//   face_partials :87 (the pair loop :100-116, rows combined in face order :126-131),
//   energy_value :61, energy_and_differential :146, culprit_pair :165,
// and crates/ddg-repulsive-surfaces/src/surface.rs:513 `validate_positions`.
#include <algorithm>
#include <thread>

#include "oracle.hpp"

namespace orc {

void parallel_for(int64_t n, int threads, const std::function<void(int64_t)>& f) {
  int t = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if (t <= 1 || n < 2) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  t = (int)std::min<int64_t>(t, n);
  std::vector<std::thread> pool;
  std::atomic<int64_t> next{0};
  const int64_t grain = std::max<int64_t>(1, n / (16 * t));
  for (int k = 0; k < t; ++k)
    pool.emplace_back([&] {
      for (;;) {
        int64_t i0 = next.fetch_add(grain);
        if (i0 >= n) break;
        int64_t i1 = std::min(n, i0 + grain);
        for (int64_t i = i0; i < i1; ++i) f(i);
      }
    });
  for (auto& th : pool) th.join();
}

bool validate_positions(const Mesh& m, const Positions& p, Error* err) {
  if ((int64_t)p.size() != m.nv) {
    err->status = INVALID_OPTION;
    err->detail = "position count mismatch";
    return false;
  }
  for (size_t v = 0; v < p.size(); ++v)
    if (!(std::isfinite(p[v][0]) && std::isfinite(p[v][1]) && std::isfinite(p[v][2]))) {
      err->status = NON_FINITE_POSITION;
      err->a = (int64_t)v;
      err->detail = "vertex " + std::to_string(v) + " has a non-finite position";
      return false;
    }
  for (size_t f = 0; f < m.faces.size(); ++f) {
    double a;
    V3 n, x;
    face_quantities(corners(p, m.faces[f]), &a, &n, &x);
    if (!(a > 0.0 && std::isfinite(a) && std::isfinite(n[0]) && std::isfinite(n[1]) &&
          std::isfinite(n[2]))) {
      err->status = DEGENERATE_FACE;
      err->a = (int64_t)f;
      err->detail = "face " + std::to_string(f) + " has zero area";
      return false;
    }
  }
  return true;
}

FacePartials exact_face_partials(const FaceGeometry& g, const Exponents& e, int threads) {
  const int64_t nf = (int64_t)g.areas.size();
  std::vector<double> e_rows(nf);
  FacePartials out;
  out.d_area.resize(nf);
  out.d_bary.resize(nf);
  out.d_normal.resize(nf);
  parallel_for(nf, threads, [&](int64_t s) {
    V3 xs = g.barycenters[s], ns = g.normals[s];
    double a_s = g.areas[s];
    double e_row = 0.0, d_a = 0.0;
    V3 d_x{0.0, 0.0, 0.0}, d_n{0.0, 0.0, 0.0};
    for (int64_t t = 0; t < nf; ++t) {
      if (t == s) continue;
      double a_t = g.areas[t];
      V3 v = sub(xs, g.barycenters[t]);
      double k_st, k_ts;
      V3 dv_st, dv_ts, dn_st;
      pair_terms(e, v, ns, g.normals[t], &k_st, &k_ts, &dv_st, &dv_ts, &dn_st);
      e_row += a_t * k_st;
      d_a += a_t * (k_st + k_ts);
      d_x = axpy(d_x, a_t, sub(dv_st, dv_ts));
      d_n = axpy(d_n, a_t, dn_st);
    }
    e_rows[s] = a_s * e_row;
    out.d_area[s] = d_a;
    out.d_bary[s] = scale(d_x, a_s);
    out.d_normal[s] = scale(d_n, a_s);
  });
  out.energy = 0.0;
  for (int64_t s = 0; s < nf; ++s) out.energy += e_rows[s];
  return out;
}

// exact.rs:61 `energy_value`: rows a_S Σ_{T≠S} a_T K(X_S − X_T, N_S), summed in face order.
double exact_energy_value(const FaceGeometry& g, const Exponents& e, int threads) {
  const int64_t nf = (int64_t)g.areas.size();
  std::vector<double> rows((size_t)nf);
  parallel_for(nf, threads, [&](int64_t s) {
    double acc = 0.0;
    for (int64_t t = 0; t < nf; ++t)
      if (t != s) acc += g.areas[t] * kernel_of_v(e, sub(g.barycenters[s], g.barycenters[t]), g.normals[s]);
    rows[s] = g.areas[s] * acc;
  });
  double sum = 0.0;
  for (int64_t s = 0; s < nf; ++s) sum += rows[s];
  return sum;
}

// exact.rs:165 `culprit_pair`.
static void culprit_pair(const FaceGeometry& g, const Exponents& e, int64_t* ps, int64_t* pt) {
  const int64_t nf = (int64_t)g.areas.size();
  int64_t bs = 0, bt = 0;
  double best = -INFINITY;
  for (int64_t s = 0; s < nf; ++s)
    for (int64_t t = 0; t < nf; ++t) {
      if (s == t) continue;
      KernelEval k = kernel_with_gradient_of_v(e, sub(g.barycenters[s], g.barycenters[t]), g.normals[s]);
      double term = k.value * g.areas[s] * g.areas[t];
      bool finite = std::isfinite(term);
      for (int c = 0; c < 3; ++c) finite = finite && std::isfinite(k.d_v[c]) && std::isfinite(k.d_n[c]);
      if (!finite) {
        *ps = s;
        *pt = t;
        return;
      }
      if (term > best) {
        best = term;
        bs = s;
        bt = t;
      }
    }
  *ps = bs;
  *pt = bt;
}

bool exact_energy_and_differential(const Mesh& m, const Positions& p, const Exponents& e, int threads,
                                   double* energy, std::vector<V3>* diff, FacePartials* partials,
                                   Error* err) {
  if (!validate_positions(m, p, err)) return false;
  FaceGeometry g = face_geometry(m, p);
  FacePartials fp = exact_face_partials(g, e, threads);
  std::vector<V3> d = pull_back(m, p, fp.d_area, fp.d_bary, fp.d_normal);
  bool finite = std::isfinite(fp.energy);
  for (auto& x : d)
    for (int c = 0; c < 3; ++c) finite = finite && std::isfinite(x[c]);
  if (!finite) {
    err->status = NON_FINITE_ENERGY;
    culprit_pair(g, e, &err->a, &err->b);
    err->detail = "faces (" + std::to_string(err->a) + ", " + std::to_string(err->b) +
                  ") give a non-finite energy";
    return false;
  }
  *energy = fp.energy;
  *diff = std::move(d);
  if (partials) *partials = std::move(fp);
  return true;
}

}  // namespace orc
