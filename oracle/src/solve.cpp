// FP64 oracle — TEST INFRASTRUCTURE ONLY: the Hs saddle solver and one descent step.
//
// Restates:
//   crates/ddg-repulsive-surfaces/src/geometry.rs:221 cotan_stiffness, :194 vertex_areas,
//       mean_edge_length
//   crates/ddg-repulsive-surfaces/src/constraints/mod.rs:556 jacobian (the barycenter rows)
//   crates/ddg-repulsive-surfaces/src/solve/dense.rs:529 preconditioner_epsilon
//   crates/ddg-repulsive-surfaces/src/solve/precond.rs  LaplacianSaddle :157 (per-axis scalar
//       saddle, multipliers last), fractional_symbol_ratio :561, SparsePreconditioner::apply_fine
//       :1026 (augmented form, Symbol scale, vertex pairs, no refinement)
//   crates/ddg-repulsive-surfaces/src/solve/iterative.rs:284 HierarchicalOperator::apply_batch
//   crates/ddg-repulsive-surfaces/src/solve/gmres.rs:311-583 GmresRun (right preconditioning,
//       x0 = 0, MGS fused with the next dot in chunks of 4096, Givens, explicit residual per cycle)
//   crates/ddg-repulsive-surfaces/src/step/session.rs:749 step_inner (energy, direction, Armijo
//       line search with the StartRule kinds of :1604, max_halvings, the trial records, the
//       barycenter projection and F after it), with remeshing and CCD off; the objective's exact
//       or Barnes–Hut evaluation (step/objective.rs:151, energy/exact.rs:61 energy_value).
//
// The L′ factor is this oracle's own: a nested-dissection order by coordinate bisection and the
// up-looking LDLᵀ of T. Davis ("Algorithm 849: A concise sparse Cholesky factorization package",
// ACM TOMS 31(4), 2005).  The reference factors with faer 0.24.4 (AMD order, not vendored); the
// factor is the same matrix factorization up to round-off, which the Lean `precond` fixtures pin
// at 1e-10 (tests/oracle.rs:45-48).
//
// Pins: the Lean `rs-icosa-precond` fixture (ε, L′ solves, P⁻¹ r), the Lean `rs-octa-saddle-none`
// fixture (the direction x = A′⁻¹ [g; 0] and ⟨g, x⟩ with the dense metric), an independent numpy
// assembly of K and L′, GMRES against a dense solve.  session_step: the Lean `rs-icosa-step-stuck`
// fixture (barycenter only, exact energy, dense metric = χ 0, absolute start, three halvings:
// F₀, ⟨g, x⟩, ‖x‖, ℓ̄₀, τ₀, every trial's τ, outcome and F, the stuck result and F after the
// projection, tests/oracle.rs:557-640 tolerances), plus the Armijo / energy / barycenter
// properties of accepted steps.
#include <algorithm>
#include <chrono>
#include <numeric>

#include "oracle.hpp"

namespace orc {

Csr csr_from_triplets(size_t nrows, size_t ncols, const std::vector<Triplet>& t) {
  std::vector<size_t> idx(t.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return t[a].r != t[b].r ? t[a].r < t[b].r : t[a].c < t[b].c;
  });
  Csr m;
  m.nrows = nrows;
  m.ncols = ncols;
  m.ptr.assign(nrows + 1, 0);
  for (size_t q = 0; q < idx.size();) {
    const Triplet& a = t[idx[q]];
    double v = 0.0;
    size_t q2 = q;
    while (q2 < idx.size() && t[idx[q2]].r == a.r && t[idx[q2]].c == a.c) v += t[idx[q2++]].v;
    m.idx.push_back(a.c);
    m.val.push_back(v);
    m.ptr[a.r + 1]++;
    q = q2;
  }
  for (size_t i = 0; i < nrows; ++i) m.ptr[i + 1] += m.ptr[i];
  return m;
}

std::vector<Triplet> cotan_triplets(const Mesh& m, const Positions& p) {
  std::vector<Triplet> trip;
  trip.reserve(12 * m.faces.size());
  for (auto& t : m.faces) {
    auto q = corners(p, t);
    for (int c = 0; c < 3; ++c) {
      uint32_t a = t[(c + 1) % 3], b = t[(c + 2) % 3];
      V3 u = sub(q[(c + 1) % 3], q[c]);
      V3 w = sub(q[(c + 2) % 3], q[c]);
      double cot = dot(u, w) / norm(cross(u, w));
      double h = 0.5 * cot;
      trip.push_back({a, b, -h});
      trip.push_back({b, a, -h});
      trip.push_back({a, a, h});
      trip.push_back({b, b, h});
    }
  }
  return trip;
}

double mean_edge_length(const Mesh& m, const Positions& p) {
  std::vector<std::pair<uint32_t, uint32_t>> e;
  for (auto& t : m.faces)
    for (int c = 0; c < 3; ++c) {
      uint32_t a = t[c], b = t[(c + 1) % 3];
      e.push_back({std::min(a, b), std::max(a, b)});
    }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  double s = 0.0;
  for (auto [a, b] : e) s += norm(sub(p[a], p[b]));
  return s / (double)e.size();
}

// Nested dissection by coordinate bisection: domains ordered before their separator.
static void nd_order(const std::vector<uint32_t>& verts, const Positions& p, const Csr& g,
                     std::vector<int>& mark, int tag, std::vector<uint32_t>& out) {
  if (verts.size() <= 64) {
    out.insert(out.end(), verts.begin(), verts.end());
    return;
  }
  V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v : verts)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[v][k]);
      hi[k] = std::max(hi[k], p[v][k]);
    }
  int ax = 0;
  for (int k = 1; k < 3; ++k)
    if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
  std::vector<uint32_t> s = verts;
  std::sort(s.begin(), s.end(), [&](uint32_t a, uint32_t b) {
    return p[a][ax] != p[b][ax] ? p[a][ax] < p[b][ax] : a < b;
  });
  size_t half = s.size() / 2;
  // side: tag for A, tag+1 for B
  for (size_t i = 0; i < s.size(); ++i) mark[s[i]] = i < half ? tag : tag + 1;
  std::vector<uint32_t> A, B, S;
  for (size_t i = 0; i < half; ++i) {
    uint32_t v = s[i];
    bool touches = false;
    for (size_t q = g.ptr[v]; q < g.ptr[v + 1]; ++q)
      if (mark[g.idx[q]] == tag + 1) touches = true;
    (touches ? S : A).push_back(v);
  }
  for (size_t i = half; i < s.size(); ++i) B.push_back(s[i]);
  for (uint32_t v : S) mark[v] = -1;
  nd_order(A, p, g, mark, tag + 2, out);
  nd_order(B, p, g, mark, tag + 2, out);
  out.insert(out.end(), S.begin(), S.end());
}

std::vector<uint32_t> nested_dissection(const Csr& k, const Positions& p) {
  std::vector<uint32_t> all(k.nrows);
  std::iota(all.begin(), all.end(), 0);
  std::vector<int> mark(k.nrows, -1);
  std::vector<uint32_t> out;
  out.reserve(k.nrows);
  nd_order(all, p, k, mark, 0, out);
  return out;
}

bool ldlt_factor(const Csr& a, const std::vector<uint32_t>& perm, Ldlt* f, Error* err) {
  // `a` symmetric with both triangles stored (CSR rows = CSC columns).
  const size_t n = a.nrows;
  f->n = n;
  f->perm = perm;
  std::vector<uint32_t> pinv(n);
  for (size_t i = 0; i < n; ++i) pinv[perm[i]] = (uint32_t)i;
  std::vector<int64_t> parent(n), flag(n);
  std::vector<uint64_t> lnz(n);
  for (size_t k = 0; k < n; ++k) {
    parent[k] = -1;
    flag[k] = (int64_t)k;
    lnz[k] = 0;
    size_t kk = perm[k];
    for (size_t q = a.ptr[kk]; q < a.ptr[kk + 1]; ++q) {
      size_t i = pinv[a.idx[q]];
      if (i < k)
        for (; flag[i] != (int64_t)k; i = (size_t)parent[i]) {
          if (parent[i] == -1) parent[i] = (int64_t)k;
          lnz[i]++;
          flag[i] = (int64_t)k;
        }
    }
  }
  f->lp.assign(n + 1, 0);
  for (size_t k = 0; k < n; ++k) f->lp[k + 1] = f->lp[k] + lnz[k];
  f->li.assign(f->lp[n], 0);
  f->lx.assign(f->lp[n], 0.0);
  f->d.assign(n, 0.0);
  std::vector<double> y(n, 0.0);
  std::vector<size_t> pattern(n);
  for (size_t k = 0; k < n; ++k) {
    y[k] = 0.0;
    size_t top = n;
    flag[k] = (int64_t)k;
    lnz[k] = 0;
    size_t kk = perm[k];
    for (size_t q = a.ptr[kk]; q < a.ptr[kk + 1]; ++q) {
      size_t i = pinv[a.idx[q]];
      if (i <= k) {
        y[i] += a.val[q];
        size_t len = 0;
        for (; flag[i] != (int64_t)k; i = (size_t)parent[i]) {
          pattern[len++] = i;
          flag[i] = (int64_t)k;
        }
        while (len > 0) pattern[--top] = pattern[--len];
      }
    }
    f->d[k] = y[k];
    y[k] = 0.0;
    for (; top < n; ++top) {
      size_t i = pattern[top];
      double yi = y[i];
      y[i] = 0.0;
      uint64_t p2 = f->lp[i] + lnz[i];
      for (uint64_t q = f->lp[i]; q < p2; ++q) y[f->li[q]] -= f->lx[q] * yi;
      double l_ki = yi / f->d[i];
      f->d[k] -= l_ki * yi;
      f->li[p2] = (uint32_t)k;
      f->lx[p2] = l_ki;
      lnz[i]++;
    }
    if (f->d[k] == 0.0 || !std::isfinite(f->d[k])) {
      err->status = SINGULAR_SADDLE;
      err->a = (int64_t)k;
      err->detail = "zero pivot in the L' factorization";
      return false;
    }
  }
  f->parent.assign(parent.begin(), parent.end());
  return true;
}

// x = A⁻¹ b for one column in the original numbering.
void ldlt_solve(const Ldlt& f, const double* b, double* x) {
  const size_t n = f.n;
  std::vector<double> y(n);
  for (size_t i = 0; i < n; ++i) y[i] = b[f.perm[i]];
  for (size_t j = 0; j < n; ++j)
    for (uint64_t q = f.lp[j]; q < f.lp[j + 1]; ++q) y[f.li[q]] -= f.lx[q] * y[j];
  for (size_t j = 0; j < n; ++j) y[j] /= f.d[j];
  for (size_t j = n; j-- > 0;)
    for (uint64_t q = f.lp[j]; q < f.lp[j + 1]; ++q) y[j] -= f.lx[q] * y[f.li[q]];
  for (size_t i = 0; i < n; ++i) x[f.perm[i]] = y[i];
}

// precond.rs:528-564: `4 I(σ) I(2 − s)`, I(t) = π Γ(1−t) / (t 4^t Γ(1+t)).
double fractional_symbol_ratio(double s) {
  auto I = [](double t) { return M_PI * std::tgamma(1.0 - t) / (t * std::pow(4.0, t) * std::tgamma(1.0 + t)); };
  return 4.0 * I(s - 1.0) * I(2.0 - s);
}

bool saddle_build(const Mesh& m, const Positions& p, const Exponents& e, const SaddleOptions& o, int threads,
                  Saddle* out, Error* err) {
  if (!face_metric_build(m, p, e, o.chi, o.leaf_size, o.leaf_pairs, threads, &out->metric, err)) return false;
  const size_t nv = (size_t)m.nv;
  out->nv = nv;
  out->w_tp = o.w_tp;
  out->c_lo = o.c_lo;
  out->mass = vertex_areas(m, out->metric.geom.areas);
  // Global barycenter rows a_v / Σ a (constraints/mod.rs:556), one scalar row shared by the axes.
  double tot = 0.0;
  for (size_t v = 0; v < nv; ++v) tot += out->mass[v];
  out->crow.resize(nv);
  for (size_t v = 0; v < nv; ++v) out->crow[v] = out->mass[v] / tot;
  // K = W + ε M (precond.rs:705-760)
  auto trip = cotan_triplets(m, p);
  Csr w = csr_from_triplets(nv, nv, trip);
  double trw = 0.0, trm = 0.0;
  for (size_t i = 0; i < nv; ++i)
    for (size_t q = w.ptr[i]; q < w.ptr[i + 1]; ++q)
      if (w.idx[q] == i) trw += w.val[q];
  for (size_t v = 0; v < nv; ++v) trm += out->mass[v];
  out->eps = o.eps_rel * trw / trm;
  for (size_t v = 0; v < nv; ++v) trip.push_back({(uint32_t)v, (uint32_t)v, out->eps * out->mass[v]});
  out->k = csr_from_triplets(nv, nv, trip);
  // L′ = [[K, c1ᵀ], [c1, 0]], order: nested dissection of K, then the multiplier.
  std::vector<Triplet> lt = trip;
  for (size_t v = 0; v < nv; ++v) {
    lt.push_back({(uint32_t)nv, (uint32_t)v, out->crow[v]});
    lt.push_back({(uint32_t)v, (uint32_t)nv, out->crow[v]});
  }
  Csr lp = csr_from_triplets(nv + 1, nv + 1, lt);
  auto order = nested_dissection(out->k, p);
  order.push_back((uint32_t)nv);
  if (!ldlt_factor(lp, order, &out->lprime, err)) return false;
  out->lprime_matrix = std::move(lp);
  const double s = e.s();
  if (!vertex_fractional_build(p, out->mass, 2.0 - s, o.chi, o.leaf_size, o.leaf_pairs, threads, &out->frac, err))
    return false;
  out->frac_scale = o.scale == 1 ? 1.0 : 1.0 / fractional_symbol_ratio(s) / o.w_tp;
  out->form = o.form;
  out->refine = o.refine;
  return true;
}

// A′ [x; λ] (iterative.rs:284): w_TP (B̃ + c_lo B̃⁰) x + C_Qᵀ λ; C_Q x.
void saddle_apply(const Saddle& sd, const double* x, double* y, int threads) {
  const size_t nv = sd.nv, nd = 3 * nv;
  face_metric_apply(sd.metric, sd.w_tp, sd.w_tp * sd.c_lo, 1, x, y, threads);
  for (int k = 0; k < 3; ++k) {
    double lam = x[nd + k];
    double cx = 0.0;
    for (size_t v = 0; v < nv; ++v) {
      y[3 * v + k] += sd.crow[v] * lam;
      cx += sd.crow[v] * x[3 * v + k];
    }
    y[nd + k] = cx;
  }
}

// L′⁻¹ r for a full vector (3 nv + 3), the three axes as scalar saddles.
static void lprime_solve(const Saddle& sd, const double* r, double* x) {
  const size_t nv = sd.nv, nd = 3 * nv;
  std::vector<double> b(nv + 1), y(nv + 1);
  for (int k = 0; k < 3; ++k) {
    for (size_t v = 0; v < nv; ++v) b[v] = r[3 * v + k];
    b[nv] = r[nd + k];
    ldlt_solve(sd.lprime, b.data(), y.data());
    for (size_t v = 0; v < nv; ++v) x[3 * v + k] = y[v];
    x[nd + k] = y[nv];
  }
}

// L′⁻¹ r with one step of iterative refinement when the saddle asks for it (precond.rs:1031
// solve_many_on(.., refine)): x ← x + L′⁻¹ (r − L′ x), L′ the assembled scalar saddle per axis.
static void lprime_solve_opt(const Saddle& sd, const double* r, double* x) {
  lprime_solve(sd, r, x);
  if (!sd.refine) return;
  const size_t nv = sd.nv, nd = 3 * nv, n = nd + 3;
  std::vector<double> res(n), dx(n);
  const Csr& L = sd.lprime_matrix;
  for (int a = 0; a < 3; ++a)
    for (size_t i = 0; i <= nv; ++i) {
      double acc = 0.0;
      for (size_t q = L.ptr[i]; q < L.ptr[i + 1]; ++q) {
        const size_t j = L.idx[q];
        acc += L.val[q] * (j < nv ? x[3 * j + a] : x[nd + a]);
      }
      const size_t t = i < nv ? 3 * i + a : nd + a;
      res[t] = r[t] - acc;
    }
  lprime_solve(sd, res.data(), dx.data());
  for (size_t i = 0; i < n; ++i) x[i] += dx[i];
}

// precond.rs:1026 apply_fine: y = L′⁻¹ r, v = L^V y on the positions and 0 on the multipliers;
// augmented form (default): v ← s v − (K ⊗ I₃) y, z = L′⁻¹ v + y; congruence form: z = L′⁻¹ v.
void precond_apply(const Saddle& sd, const double* r, double* z, int threads) {
  const size_t nv = sd.nv, nd = 3 * nv, n = nd + 3;
  std::vector<double> y(n), v(n, 0.0);
  lprime_solve_opt(sd, r, y.data());
  vertex_fractional_apply(sd.frac, 1, y.data(), v.data(), threads);
  if (sd.form == 0) {
    if (sd.frac_scale != 1.0)
      for (size_t i = 0; i < nd; ++i) v[i] *= sd.frac_scale;
    for (size_t j = 0; j < nv; ++j) {
      double acc[3] = {0.0, 0.0, 0.0};
      for (size_t q = sd.k.ptr[j]; q < sd.k.ptr[j + 1]; ++q)
        for (int a = 0; a < 3; ++a) acc[a] += sd.k.val[q] * y[3 * (size_t)sd.k.idx[q] + a];
      for (int a = 0; a < 3; ++a) v[3 * j + a] -= acc[a];
    }
  }
  lprime_solve_opt(sd, v.data(), z);
  if (sd.form == 0)
    for (size_t i = 0; i < n; ++i) z[i] += y[i];
}

namespace {
constexpr size_t kChunk = 4096;
double dot_seq(const double* a, const double* b, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
double vdot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t c = 0; c < a.size(); c += kChunk) s += dot_seq(&a[c], &b[c], std::min(kChunk, a.size() - c));
  return s;
}
double vnorm(const std::vector<double>& a) { return std::sqrt(vdot(a, a)); }
// w −= h v, then w · next (chunked)
double axpy_dot(std::vector<double>& w, double h, const std::vector<double>& v, const std::vector<double>* next) {
  double s = 0.0;
  for (size_t c = 0; c < w.size(); c += kChunk) {
    size_t e = std::min(w.size(), c + kChunk);
    for (size_t i = c; i < e; ++i) w[i] -= h * v[i];
    if (next) s += dot_seq(&w[c], &(*next)[c], e - c);
  }
  return s;
}
}  // namespace

// gmres.rs:311-583, right preconditioning.
GmresReport gmres_right(const std::function<void(const double*, double*)>& op,
                        const std::function<void(const double*, double*)>& prec, const std::vector<double>& b,
                        const GmresOptions& o, std::vector<double>* x_out) {
  const size_t n = b.size();
  const size_t m = std::min(o.restart, std::max<size_t>(n, 1));
  GmresReport rep;
  std::vector<double> x(n, 0.0);
  double bnorm = vnorm(b);
  if (bnorm == 0.0) {
    rep.converged = true;
    *x_out = x;
    return rep;
  }
  std::vector<std::vector<double>> basis, h;
  std::vector<double> cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0), r(n), tmp(n), w(n);
  for (;;) {
    // explicit residual
    if (rep.cycles > 0) {
      op(x.data(), tmp.data());
      rep.operator_applications++;
      for (size_t i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
    } else {
      r = b;
    }
    double beta = vnorm(r);
    rep.relative_residual = beta / bnorm;
    double rel = beta / bnorm;
    rep.cycle_residuals.push_back(rel);
    if (!std::isfinite(rel)) break;
    if (rel <= o.tolerance) {
      rep.converged = true;
      break;
    }
    if (rep.iterations >= o.max_iterations) break;
    rep.cycles++;
    basis.clear();
    h.clear();
    basis.push_back(std::vector<double>(n));
    for (size_t i = 0; i < n; ++i) basis[0][i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    size_t j = 0, k_end = 0;
    for (;;) {
      prec(basis[j].data(), tmp.data());
      op(tmp.data(), w.data());
      rep.operator_applications++;
      rep.preconditioner_applications++;
      std::vector<double> col(j + 2, 0.0);
      double hij = vdot(w, basis[0]);
      for (size_t i = 0; i <= j; ++i) {
        col[i] = hij;
        hij = axpy_dot(w, col[i], basis[i], i < j ? &basis[i + 1] : nullptr);
      }
      double hn = vnorm(w);
      col[j + 1] = hn;
      for (size_t i = 0; i < j; ++i) {
        double t = cs[i] * col[i] + sn[i] * col[i + 1];
        col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
        col[i] = t;
      }
      double denom = std::hypot(col[j], col[j + 1]);
      if (denom == 0.0) {
        cs[j] = 1.0;
        sn[j] = 0.0;
      } else {
        cs[j] = col[j] / denom;
        sn[j] = col[j + 1] / denom;
      }
      col[j] = denom;
      col[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] *= cs[j];
      h.push_back(col);
      rep.iterations++;
      double est = std::fabs(g[j + 1]) / bnorm;
      rep.history.push_back(est);
      if (hn == 0.0 || !std::isfinite(hn) || est <= o.tolerance || rep.iterations >= o.max_iterations || j + 1 >= m) {
        k_end = j + 1;
        break;
      }
      basis.push_back(std::vector<double>(n));
      for (size_t i = 0; i < n; ++i) basis[j + 1][i] = w[i] / hn;
      j++;
    }
    std::vector<double> y(k_end, 0.0);
    for (size_t i = k_end; i-- > 0;) {
      double s = g[i];
      for (size_t l = i + 1; l < k_end; ++l) s -= h[l][i] * y[l];
      y[i] = h[i][i] != 0.0 ? s / h[i][i] : 0.0;
    }
    std::vector<double> u(n, 0.0);
    for (size_t l = 0; l < k_end; ++l)
      for (size_t i = 0; i < n; ++i) u[i] += y[l] * basis[l][i];
    prec(u.data(), tmp.data());
    rep.preconditioner_applications++;
    for (size_t i = 0; i < n; ++i) x[i] += tmp[i];
  }
  *x_out = std::move(x);
  return rep;
}

// Area-weighted barycenter (constraints/mod.rs:448 `barycenter`, vertex areas :487 `areas`).
static V3 barycenter_of(const Mesh& m, const Positions& p) {
  FaceGeometry gg = face_geometry(m, p);
  auto va = vertex_areas(m, gg.areas);
  double tot = 0.0;
  V3 bc{0, 0, 0};
  for (size_t v = 0; v < (size_t)m.nv; ++v) {
    for (int k = 0; k < 3; ++k) bc[k] += va[v] * p[v][k];
    tot += va[v];
  }
  return scale(bc, 1.0 / tot);
}

// session.rs:1604 `start_tau`.
static double start_tau(const SessionOptions& o, double last_tau, double xnorm, double max_disp, double l0) {
  switch (o.start_kind) {
    case START_ABSOLUTE: return o.start_value;
    case START_MAXDISP: return o.start_value * l0 / max_disp;
    case START_REFERENCE: return 2.0 / xnorm;
    default: return last_tau > 0.0 ? 2.0 * last_tau : o.start_value * l0 / max_disp;
  }
}

bool session_step(Session& s, int threads, StepReport* rep, Error* err) {
  const Mesh& m = s.mesh;
  const size_t nv = (size_t)m.nv, nd = 3 * nv;
  const bool exact = s.opts.energy_exact;
  auto clock = std::chrono::steady_clock::now();
  auto lap = [&](int k) {
    const auto now = std::chrono::steady_clock::now();
    rep->seconds[k] = std::chrono::duration<double>(now - clock).count();
    clock = now;
  };
  // 1. trees, lists, F and dF (objective.rs:71 build, :151 evaluate)
  BhTree bh;
  double f0;
  std::vector<V3> gv;
  if (exact) {
    if (!exact_energy_and_differential(m, s.positions, s.exps, threads, &f0, &gv, nullptr, err)) return false;
  } else {
    if (!bh_build(m, s.positions, s.opts.theta, s.opts.saddle.leaf_size, threads, &bh, err)) return false;
    if (!bh_energy_and_differential(bh, s.exps, threads, &f0, &gv, err)) return false;
  }
  // F at trial positions, value only; a geometric failure is a rejected (non-finite) trial
  auto value_at = [&](const Positions& q) {
    Error e2;
    if (exact) {
      if (!validate_positions(m, q, &e2)) return (double)NAN;
      return exact_energy_value(face_geometry(m, q), s.exps, threads);
    }
    if (!bh_refit(bh, q, &e2)) return (double)NAN;
    return bh_energy(bh, s.exps, threads);
  };
  std::vector<double> g(nd);
  for (size_t v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k) g[3 * v + k] = gv[v][k];
  lap(0);
  // 2. the solver and the direction (session.rs:820-860): with nonlinear rows, the Schur columns
  // z_r = A′⁻¹ [C_N row r; 0] and y = A′⁻¹ [g; 0] (schur.rs:101 with_direction: one batch, the
  // same results as separate solves), then x = y − Z N⁻¹ C_N y; without them x = y
  Saddle sd;
  if (!saddle_build(m, s.positions, s.exps, s.opts.saddle, threads, &sd, err)) return false;
  lap(1);
  auto op = [&](const double* x, double* y) { saddle_apply(sd, x, y, threads); };
  auto pc = [&](const double* r, double* z) { precond_apply(sd, r, z, threads); };
  auto solve = [&](const std::vector<double>& b, std::vector<double>* out, GmresReport* gr) {
    std::vector<double> rhs(nd + 3, 0.0), sol;
    std::copy(b.begin(), b.end(), rhs.begin());
    *gr = gmres_right(op, pc, rhs, s.opts.gmres, &sol);
    if (!gr->converged && !(gr->relative_residual <= 1e-1)) {
      err->status = SOLVER_DIVERGED;
      err->detail = "GMRES diverged";
      return false;
    }
    out->assign(sol.begin(), sol.begin() + nd);
    return true;
  };
  const size_t kn = s.opts.rows.size();
  Schur schur;
  std::vector<std::vector<double>> crow(kn), zcol(kn);
  for (size_t r = 0; r < kn; ++r) {
    crow[r] = row_gradient(m, s.positions, s.opts.rows[r]);
    GmresReport gr;
    if (!solve(crow[r], &zcol[r], &gr)) return false;
  }
  std::vector<double> x;
  if (!solve(g, &x, &rep->gmres)) return false;
  if (kn) {
    schur_build(&schur, std::move(crow), std::move(zcol));
    x = schur_direction(schur, x);
  }
  rep->min_pivot = schur.min_pivot;
  lap(2);
  double slope = 0.0, xinf = 0.0, xn2 = 0.0, maxd = 0.0, gn2 = 0.0;
  for (size_t i = 0; i < nd; ++i) {
    slope += g[i] * x[i];
    gn2 += g[i] * g[i];
    xinf = std::max(xinf, std::fabs(x[i]));
    xn2 += x[i] * x[i];
  }
  for (size_t v = 0; v < nv; ++v)
    maxd = std::max(maxd, std::sqrt(x[3 * v] * x[3 * v] + x[3 * v + 1] * x[3 * v + 1] + x[3 * v + 2] * x[3 * v + 2]));
  const double xnorm = std::sqrt(xn2);
  // 3. Armijo line search (session.rs:906, :958-1069)
  double tau0 = start_tau(s.opts, s.last_tau, xnorm, maxd, s.l0);
  const double tau_min = s.opts.tau_min_rel * s.l0 / xinf;
  // the CCD cap (session.rs:917-956): the band of the faces that intersect at the start, then
  // τ₀ ← min(τ₀, ⅔ τ_max) of first_contact_with_band along −x
  rep->has_ccd = false;
  if (s.opts.ccd) {
    size_t np = 0, nb = 0;
    const std::vector<uint8_t> band = intersection_band(m, s.positions, &np, &nb);
    if (std::isfinite(tau0) && tau0 > 0.0) {
      CcdOptions co = s.opts.ccd_opts;
      if (std::isnan(co.reference_length)) co.reference_length = s.l0;
      Positions vel(nv);
      for (size_t v = 0; v < nv; ++v) vel[v] = {-x[3 * v], -x[3 * v + 1], -x[3 * v + 2]};
      if (!first_contact(m, s.positions, vel, tau0, co, nb ? &band : nullptr, &rep->ccd, err)) return false;
      if (rep->ccd.has_contact) tau0 = fmin_(tau0, 2.0 / 3.0 * rep->ccd.tau_max);
      rep->has_ccd = true;
    }
  }
  rep->energy0 = f0;
  rep->slope = slope;
  rep->gradient_norm = std::sqrt(gn2);
  rep->direction_norm = xnorm;
  rep->direction_max = xinf;
  rep->tau0 = tau0;
  rep->accepted = false;
  rep->trials = rep->halvings = 0;
  rep->trial.clear();
  Positions trial(nv);
  double tau = tau0;
  if (slope > 0.0 && std::isfinite(tau0) && tau0 > 0.0) {
    for (;;) {
      if (!(tau >= tau_min) || (s.opts.max_halvings >= 0 && (long)rep->halvings > s.opts.max_halvings)) break;
      for (size_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) trial[v][k] = s.positions[v][k] - tau * x[3 * v + k];
      const double ft = value_at(trial);
      const bool ok = std::isfinite(ft) && ft <= f0 - s.opts.armijo_c * tau * slope;
      rep->trial.push_back({tau, std::isfinite(ft) ? ft : NAN, !std::isfinite(ft) ? 2 : ok ? 0 : 1});
      rep->trials++;
      if (ok) {
        rep->accepted = true;
        rep->energy1 = ft;
        break;
      }
      tau *= 0.5;
      rep->halvings++;
    }
  }
  if (!rep->accepted) rep->energy1 = f0;
  lap(3);
  // 4. projection (schur.rs:495 project): the positional projection (constraints/mod.rs:657:
  // translate by −Φ, Φ the barycenter's offset from its target, areas at the moved positions),
  // then Newton rounds f ← f − Z N⁻¹ Φ_N, each followed by the positional projection and kept only
  // if it lowers max_r |Φ_r| / tol_r over the kept rows; then F at the projected positions
  auto values = [&](const Positions& q) {
    std::vector<double> phi(3 + kn);
    const V3 b = barycenter_of(m, q);
    for (int k = 0; k < 3; ++k) phi[k] = b[k] - s.barycenter_target[k];
    for (size_t r = 0; r < kn; ++r) phi[3 + r] = row_value(m, q, s.opts.rows[r]) - s.row_targets[r];
    return phi;
  };
  auto project_positional = [&](Positions& q) {
    const V3 b = barycenter_of(m, q);
    for (size_t v = 0; v < nv; ++v)
      for (int k = 0; k < 3; ++k) q[v][k] -= b[k] - s.barycenter_target[k];
  };
  Positions np = rep->accepted ? trial : s.positions;
  rep->phi_before = values(np);
  project_positional(np);
  std::vector<double> phi = values(np);
  rep->newton_rounds = 0;
  if (kn && !schur.kept.empty()) {
    std::vector<size_t> kept = schur.kept;
    std::sort(kept.begin(), kept.end());
    auto measure = [&](const std::vector<double>& ph) {
      double mm = 0.0;
      for (size_t i : kept) {
        const double v = std::fabs(ph[3 + i]) / row_tolerance(s.row_targets[i]);
        if (std::isnan(v) || std::isnan(mm))
          mm = NAN;
        else
          mm = std::max(mm, v);
      }
      return mm;
    };
    double mcur = measure(phi);
    const bool until = !s.opts.newton_fixed;
    while (rep->newton_rounds < s.opts.newton_count) {
      if (until && mcur <= 1.0) break;
      const std::vector<double> phin(phi.begin() + 3, phi.end());
      const std::vector<double> h = schur_correction(schur, phin);
      Positions t2 = np;
      for (size_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) t2[v][k] -= h[3 * v + k];
      project_positional(t2);
      const std::vector<double> pt = values(t2);
      const double mt = measure(pt);
      if (!(mt < mcur)) break;
      np = t2;
      phi = pt;
      mcur = mt;
      rep->newton_rounds++;
    }
  }
  rep->phi = phi;
  rep->energy_after = value_at(np);
  lap(4);
  s.last_phi_before = rep->phi_before;
  s.last_phi = rep->phi;
  s.last_has_ccd = rep->has_ccd;
  s.last_ccd = rep->ccd;
  s.last_newton_rounds = rep->newton_rounds;
  s.last_min_pivot = rep->min_pivot;
  s.positions = np;
  if (rep->accepted) s.last_tau = tau;
  rep->tau = rep->accepted ? tau : 0.0;
  return true;
}

bool session_create(const Mesh& m, const Positions& p, const Exponents& e, const SessionOptions& o, Session* s,
                    Error* err) {
  if (!validate_positions(m, p, err)) return false;
  s->mesh = m;
  s->positions = p;
  s->exps = e;
  s->opts = o;
  s->l0 = mean_edge_length(m, p);
  s->barycenter_target = barycenter_of(m, p);
  s->row_targets.clear();
  for (size_t r = 0; r < o.rows.size(); ++r) {
    const double t = r < o.targets.size() ? o.targets[r] : NAN;
    s->row_targets.push_back(std::isnan(t) ? row_value(m, p, o.rows[r]) : t);
  }
  return true;
}

}  // namespace orc
