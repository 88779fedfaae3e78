// Nonlinear constraint rows and their Schur complement: the oracle's restatement (FP64) of
//   crates/ddg-repulsive-surfaces/src/constraints/mod.rs   area_of :462, volume_of :474,
//       values :519, area_gradient :606, volume_gradient :627, tolerance :435
//       (NONLINEAR_TOLERANCE :147), project_positional :657 (one barycenter group, no pins)
//   crates/ddg-repulsive-surfaces/src/geometry.rs           face_area_gradient :50
//   crates/ddg-repulsive-surfaces/src/solve/schur.rs        from_solutions :142 (S = −C_N Z),
//       factor :185 (equilibrated, pivoted Cholesky of N = −½(S + Sᵀ), rows dropped at pivots
//       ≤ SCHUR_DROP_TOLERANCE :33), solve_n :251, combine :281, direction_from :318,
//       correction :341, project :495 (Newton rounds with the safeguard)
// Test infrastructure only.
//
// Parity pins: the values and Jacobians against the reference's constraint fixtures
// (rs-icosa-cons-area-volume, rs-mobius-cons-area, rs-patch-cons-open) and central differences;
// the whole constrained step against rs-icosa-step, rs-mobius-step and rs-torus-step-ref.
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace orc {

double total_area(const Mesh& m, const Positions& p) {
  double s = 0.0;
  for (size_t f = 0; f < m.faces.size(); ++f) {
    double a;
    V3 n, x;
    face_quantities(corners(p, m.faces[f]), &a, &n, &x);
    s += a;
  }
  return s;
}

double signed_volume(const Mesh& m, const Positions& p) {
  double s = 0.0;
  for (size_t f = 0; f < m.faces.size(); ++f) {
    const auto c = corners(p, m.faces[f]);
    s += dot(c[0], cross(c[1], c[2])) / 6.0;
  }
  return s;
}

std::vector<double> area_gradient(const Mesh& m, const Positions& p) {
  std::vector<double> g(3 * (size_t)m.nv, 0.0);
  for (size_t f = 0; f < m.faces.size(); ++f) {
    const auto c = corners(p, m.faces[f]);
    double a;
    V3 nrm, x;
    face_quantities(c, &a, &nrm, &x);
    const auto e = opposite_edges(c);
    for (int k = 0; k < 3; ++k) {
      const V3 d = scale(cross(nrm, e[k]), 0.5);
      const size_t v = m.faces[f][k];
      for (int q = 0; q < 3; ++q) g[3 * v + q] += d[q];
    }
  }
  return g;
}

std::vector<double> volume_gradient(const Mesh& m, const Positions& p) {
  std::vector<double> g(3 * (size_t)m.nv, 0.0);
  for (size_t f = 0; f < m.faces.size(); ++f) {
    const auto c = corners(p, m.faces[f]);
    for (int k = 0; k < 3; ++k) {
      const V3 d = cross(c[(k + 1) % 3], c[(k + 2) % 3]);
      const size_t v = m.faces[f][k];
      for (int q = 0; q < 3; ++q) g[3 * v + q] += d[q] / 6.0;
    }
  }
  return g;
}

double row_value(const Mesh& m, const Positions& p, int kind) {
  return kind == ROW_AREA ? total_area(m, p) : signed_volume(m, p);
}

std::vector<double> row_gradient(const Mesh& m, const Positions& p, int kind) {
  return kind == ROW_AREA ? area_gradient(m, p) : volume_gradient(m, p);
}

double row_tolerance(double target) {
  const double t = std::fabs(target);
  return t > 0.0 ? 1e-8 * t : 1e-8;
}

// C_N v: the row's entries in column order (zero entries add nothing)
static double row_dot(const std::vector<double>& c, const double* v) {
  double acc = 0.0;
  for (size_t i = 0; i < c.size(); ++i)
    if (c[i] != 0.0) acc += c[i] * v[i];
  return acc;
}

void schur_build(Schur* S, std::vector<std::vector<double>> c, std::vector<std::vector<double>> z) {
  const size_t k = c.size();
  S->c = std::move(c);
  S->z = std::move(z);
  S->s.assign(k * k, 0.0);
  for (size_t r = 0; r < k; ++r)
    for (size_t q = 0; q < k; ++q) S->s[r * k + q] = -row_dot(S->c[r], S->z[q].data());
  // factor (schur.rs:185)
  const std::vector<double>& s = S->s;
  auto n = [&](size_t i, size_t j) { return -0.5 * (s[i * k + j] + s[j * k + i]); };
  S->d.assign(k, 1.0);
  for (size_t i = 0; i < k; ++i) {
    const double nii = n(i, i);
    S->d[i] = (nii > 0.0 && std::isfinite(nii)) ? 1.0 / std::sqrt(nii) : 0.0;
  }
  std::vector<double> w(k * k);
  for (size_t i = 0; i < k; ++i)
    for (size_t j = 0; j < k; ++j) w[i * k + j] = n(i, j) * S->d[i] * S->d[j];
  std::vector<size_t> remaining, kept;
  for (size_t i = 0; i < k; ++i)
    if (S->d[i] > 0.0) remaining.push_back(i);
  std::vector<std::vector<double>> lcols;
  double min_pivot = 1.0;
  while (!remaining.empty()) {
    size_t pos = 0;
    for (size_t q = 1; q < remaining.size(); ++q)  // the largest remaining pivot, the first on ties
      if (!(w[remaining[pos] * k + remaining[pos]] >= w[remaining[q] * k + remaining[q]])) pos = q;
    const size_t p = remaining[pos];
    const double piv = w[p * k + p];
    if (!(piv > 1e-10)) break;
    min_pivot = std::min(min_pivot, piv);
    remaining.erase(remaining.begin() + (long)pos);
    const double sq = std::sqrt(piv);
    std::vector<double> col(k, 0.0);
    col[p] = sq;
    for (size_t i : remaining) col[i] = w[i * k + p] / sq;
    for (size_t i : remaining)
      for (size_t j : remaining) w[i * k + j] -= col[i] * col[j];
    kept.push_back(p);
    lcols.push_back(col);
  }
  const size_t mk = kept.size();
  S->l.assign(mk * mk, 0.0);
  for (size_t b = 0; b < mk; ++b)
    for (size_t a = 0; a < mk; ++a) S->l[a * mk + b] = lcols[b][kept[a]];
  S->kept = kept;
  S->min_pivot = (mk == 0 && k > 0) ? 0.0 : min_pivot;
}

std::vector<double> schur_solve_n(const Schur& S, const std::vector<double>& v) {
  const size_t k = S.c.size(), mk = S.kept.size();
  std::vector<double> y(mk);
  for (size_t a = 0; a < mk; ++a) y[a] = S.d[S.kept[a]] * v[S.kept[a]];
  for (size_t a = 0; a < mk; ++a) {
    double t = y[a];
    for (size_t b = 0; b < a; ++b) t -= S.l[a * mk + b] * y[b];
    y[a] = t / S.l[a * mk + a];
  }
  for (size_t a = mk; a-- > 0;) {
    double t = y[a];
    for (size_t b = a + 1; b < mk; ++b) t -= S.l[b * mk + a] * y[b];
    y[a] = t / S.l[a * mk + a];
  }
  std::vector<double> out(k, 0.0);
  for (size_t a = 0; a < mk; ++a) out[S.kept[a]] = S.d[S.kept[a]] * y[a];
  return out;
}

std::vector<double> schur_combine(const Schur& S, const std::vector<double>& c) {
  const size_t nd = S.z.empty() ? 0 : S.z[0].size();
  std::vector<double> out(nd, 0.0);
  for (size_t r = 0; r < c.size(); ++r)
    if (c[r] != 0.0)
      for (size_t i = 0; i < nd; ++i) out[i] += c[r] * S.z[r][i];
  return out;
}

std::vector<double> schur_apply_c(const Schur& S, const double* v) {
  std::vector<double> out(S.c.size());
  for (size_t r = 0; r < S.c.size(); ++r) out[r] = row_dot(S.c[r], v);
  return out;
}

std::vector<double> schur_direction(const Schur& S, const std::vector<double>& y) {
  const std::vector<double> zc = schur_combine(S, schur_solve_n(S, schur_apply_c(S, y.data())));
  std::vector<double> x(y.size());
  for (size_t i = 0; i < y.size(); ++i) x[i] = y[i] - zc[i];
  return x;
}

std::vector<double> schur_correction(const Schur& S, const std::vector<double>& phi) {
  return schur_combine(S, schur_solve_n(S, phi));
}

}  // namespace orc
