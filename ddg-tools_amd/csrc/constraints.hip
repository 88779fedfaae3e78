// Nonlinear constraint rows and their Schur complement on the host (see constraints.hpp).
#include <algorithm>
#include <cmath>

#include "constraints.hpp"

#pragma clang fp contract(off)

namespace ddgrs {

namespace {

struct Tri {
  double p[3][3];
};

Tri corners(const HostMesh& m, const double* pos, int64_t f) {
  Tri t;
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 3; ++k) t.p[c][k] = pos[3 * (size_t)m.faces[3 * f + c] + k];
  return t;
}

void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// area and unit normal (RS/src/geometry.rs:33 face_quantities)
double face_area(const Tri& t, double* nrm) {
  double e1[3], e2[3], n[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = t.p[1][k] - t.p[0][k];
    e2[k] = t.p[2][k] - t.p[0][k];
  }
  cross(e1, e2, n);
  const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nrm) {
    const double inv = 1.0 / len;
    for (int k = 0; k < 3; ++k) nrm[k] = n[k] * inv;
  }
  return 0.5 * len;
}

double row_dot(const std::vector<double>& c, const double* v) {
  double acc = 0.0;
  for (size_t i = 0; i < c.size(); ++i)
    if (c[i] != 0.0) acc += c[i] * v[i];
  return acc;
}

}  // namespace

double row_value(const HostMesh& m, const double* pos, int kind) {
  double s = 0.0;
  for (int64_t f = 0; f < m.nf; ++f) {
    const Tri t = corners(m, pos, f);
    if (kind == kRowArea) {
      s += face_area(t, nullptr);
    } else {
      double c[3];
      cross(t.p[1], t.p[2], c);
      s += (t.p[0][0] * c[0] + t.p[0][1] * c[1] + t.p[0][2] * c[2]) / 6.0;
    }
  }
  return s;
}

std::vector<double> row_gradient(const HostMesh& m, const double* pos, int kind) {
  std::vector<double> g(3 * (size_t)m.nv, 0.0);
  for (int64_t f = 0; f < m.nf; ++f) {
    const Tri t = corners(m, pos, f);
    double nrm[3];
    if (kind == kRowArea) face_area(t, nrm);
    for (int c = 0; c < 3; ++c) {
      double d[3];
      if (kind == kRowArea) {
        // ½ N × e_opp(c), e_opp = p[c+2] − p[c+1] (geometry.rs:42 opposite_edges, :50)
        double e[3];
        for (int k = 0; k < 3; ++k) e[k] = t.p[(c + 2) % 3][k] - t.p[(c + 1) % 3][k];
        cross(nrm, e, d);
        for (int k = 0; k < 3; ++k) d[k] = d[k] * 0.5;
      } else {
        cross(t.p[(c + 1) % 3], t.p[(c + 2) % 3], d);
        for (int k = 0; k < 3; ++k) d[k] = d[k] / 6.0;
      }
      const size_t v = m.faces[3 * f + c];
      for (int k = 0; k < 3; ++k) g[3 * v + k] += d[k];
    }
  }
  return g;
}

double row_tolerance(double target) {
  const double t = std::fabs(target);
  return t > 0.0 ? 1e-8 * t : 1e-8;
}

void area_barycenter(const HostMesh& m, const double* pos, double b[3]) {
  std::vector<double> va((size_t)m.nv, 0.0);
  for (int64_t f = 0; f < m.nf; ++f) {
    const double a = face_area(corners(m, pos, f), nullptr);
    for (int c = 0; c < 3; ++c) va[m.faces[3 * f + c]] += a / 3.0;
  }
  double s[3] = {0.0, 0.0, 0.0}, tot = 0.0;
  for (int64_t v = 0; v < m.nv; ++v) {
    for (int k = 0; k < 3; ++k) s[k] += va[v] * pos[3 * v + k];
    tot += va[v];
  }
  for (int k = 0; k < 3; ++k) b[k] = s[k] * (1.0 / tot);
}

void SchurRows::build(std::vector<std::vector<double>> c_rows, std::vector<std::vector<double>> z_cols) {
  const size_t k = c_rows.size();
  c = std::move(c_rows);
  z = std::move(z_cols);
  s.assign(k * k, 0.0);
  for (size_t r = 0; r < k; ++r)
    for (size_t q = 0; q < k; ++q) s[r * k + q] = -row_dot(c[r], z[q].data());
  auto n = [&](size_t i, size_t j) { return -0.5 * (s[i * k + j] + s[j * k + i]); };
  d.assign(k, 1.0);
  for (size_t i = 0; i < k; ++i) {
    const double nii = n(i, i);
    d[i] = (nii > 0.0 && std::isfinite(nii)) ? 1.0 / std::sqrt(nii) : 0.0;
  }
  std::vector<double> w(k * k);
  for (size_t i = 0; i < k; ++i)
    for (size_t j = 0; j < k; ++j) w[i * k + j] = n(i, j) * d[i] * d[j];
  std::vector<size_t> remaining;
  for (size_t i = 0; i < k; ++i)
    if (d[i] > 0.0) remaining.push_back(i);
  std::vector<std::vector<double>> lcols;
  kept.clear();
  double mp = 1.0;
  while (!remaining.empty()) {
    size_t pos = 0;
    for (size_t q = 1; q < remaining.size(); ++q)
      if (!(w[remaining[pos] * k + remaining[pos]] >= w[remaining[q] * k + remaining[q]])) pos = q;
    const size_t p = remaining[pos];
    const double piv = w[p * k + p];
    if (!(piv > 1e-10)) break;  // SCHUR_DROP_TOLERANCE
    mp = std::min(mp, piv);
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
  l.assign(mk * mk, 0.0);
  for (size_t b = 0; b < mk; ++b)
    for (size_t a = 0; a < mk; ++a) l[a * mk + b] = lcols[b][kept[a]];
  min_pivot = (mk == 0 && k > 0) ? 0.0 : mp;
}

std::vector<double> SchurRows::apply_c(const double* v) const {
  std::vector<double> out(c.size());
  for (size_t r = 0; r < c.size(); ++r) out[r] = row_dot(c[r], v);
  return out;
}

std::vector<double> SchurRows::solve_n(const std::vector<double>& v) const {
  const size_t k = c.size(), mk = kept.size();
  std::vector<double> y(mk);
  for (size_t a = 0; a < mk; ++a) y[a] = d[kept[a]] * v[kept[a]];
  for (size_t a = 0; a < mk; ++a) {
    double t = y[a];
    for (size_t b = 0; b < a; ++b) t -= l[a * mk + b] * y[b];
    y[a] = t / l[a * mk + a];
  }
  for (size_t a = mk; a-- > 0;) {
    double t = y[a];
    for (size_t b = a + 1; b < mk; ++b) t -= l[b * mk + a] * y[b];
    y[a] = t / l[a * mk + a];
  }
  std::vector<double> out(k, 0.0);
  for (size_t a = 0; a < mk; ++a) out[kept[a]] = d[kept[a]] * y[a];
  return out;
}

std::vector<double> SchurRows::combine(const std::vector<double>& cf) const {
  const size_t nd = z.empty() ? 0 : z[0].size();
  std::vector<double> out(nd, 0.0);
  for (size_t r = 0; r < cf.size(); ++r)
    if (cf[r] != 0.0)
      for (size_t i = 0; i < nd; ++i) out[i] += cf[r] * z[r][i];
  return out;
}

std::vector<double> SchurRows::direction(const std::vector<double>& y) const {
  const std::vector<double> zc = combine(solve_n(apply_c(y.data())));
  std::vector<double> x(y.size());
  for (size_t i = 0; i < y.size(); ++i) x[i] = y[i] - zc[i];
  return x;
}

std::vector<double> SchurRows::correction(const std::vector<double>& phi) const { return combine(solve_n(phi)); }

}  // namespace ddgrs
