#include <algorithm>
// FP64 oracle — TEST INFRASTRUCTURE ONLY: a flat C interface for the Python test harness
// (oracle/__init__.py loads it through ctypes).  Arrays are flat: faces `3·nf` u32, positions and
// differentials `3·nv` f64 interleaved (`3 v + axis`), per-face partials `7·nf`
// (∂/∂a, ∂/∂X, ∂/∂N).
#include <cstring>

#include "oracle.hpp"

using namespace orc;

namespace {
Mesh make_mesh(int64_t nv, int64_t nf, const uint32_t* faces) {
  Mesh m;
  m.nv = nv;
  m.faces.resize(nf);
  for (int64_t f = 0; f < nf; ++f) m.faces[f] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
  return m;
}
Positions make_pos(int64_t nv, const double* p) {
  Positions out(nv);
  for (int64_t v = 0; v < nv; ++v) out[v] = {p[3 * v], p[3 * v + 1], p[3 * v + 2]};
  return out;
}
int fail(const Error& e, char* msg, int64_t len, int64_t* culprit) {
  if (msg && len > 0) {
    std::strncpy(msg, e.detail.c_str(), (size_t)len - 1);
    msg[len - 1] = 0;
  }
  if (culprit) {
    culprit[0] = e.a;
    culprit[1] = e.b;
  }
  return e.status;
}
bool make_exps(double a, double b, Exponents* e, Error* err) {
  std::string why;
  if (!Exponents::make(a, b, e, &why)) {
    err->status = INVALID_EXPONENTS;
    err->detail = why;
    return false;
  }
  return true;
}
void write_partials(const FacePartials& fp, double* out) {
  for (size_t f = 0; f < fp.d_area.size(); ++f) {
    out[7 * f] = fp.d_area[f];
    for (int k = 0; k < 3; ++k) {
      out[7 * f + 1 + k] = fp.d_bary[f][k];
      out[7 * f + 4 + k] = fp.d_normal[f][k];
    }
  }
}
}  // namespace

extern "C" {

int orc_exact(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, double alpha, double beta,
              int threads, double* energy, double* diff, double* partials, int64_t* culprit, char* msg,
              int64_t msglen) {
  Error err;
  Exponents e;
  if (!make_exps(alpha, beta, &e, &err)) return fail(err, msg, msglen, culprit);
  Mesh m = make_mesh(nv, nf, faces);
  Positions p = make_pos(nv, pos);
  std::vector<V3> d;
  FacePartials fp;
  if (!exact_energy_and_differential(m, p, e, threads, energy, &d, &fp, &err)) return fail(err, msg, msglen, culprit);
  for (int64_t v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k) diff[3 * v + k] = d[v][k];
  if (partials) write_partials(fp, partials);
  return OK;
}

void* orc_bh_build(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, double theta,
                   int64_t leaf_size, int threads, int* status, char* msg, int64_t msglen) {
  Error err;
  auto* bh = new BhTree();
  if (!bh_build(make_mesh(nv, nf, faces), make_pos(nv, pos), theta, (size_t)leaf_size, threads, bh, &err)) {
    delete bh;
    *status = fail(err, msg, msglen, nullptr);
    return nullptr;
  }
  *status = OK;
  return bh;
}

void orc_bh_free(void* h) { delete static_cast<BhTree*>(h); }

int orc_bh_refit(void* h, const double* pos) {
  auto* bh = static_cast<BhTree*>(h);
  Error err;
  if (!bh_refit(*bh, make_pos(bh->mesh.nv, pos), &err)) return err.status;
  return OK;
}

// counts: [nodes, leaves, lumped entries, exact entries, pairs_exact, partner slots]
void orc_bh_counts(void* h, int64_t* out) {
  auto* bh = static_cast<BhTree*>(h);
  out[0] = (int64_t)bh->tree.bvh.nodes.size();
  out[1] = (int64_t)bh->tree.leaves.size();
  out[2] = (int64_t)bh->lists.lumped.size();
  out[3] = (int64_t)bh->lists.exact.size();
  out[4] = (int64_t)bh->lists.pairs_exact;
  out[5] = (int64_t)bh->lists.partner_off.back();
}

// The tree: order (nf), per node start/end/right, and per node
// [area, centroid(3), projector(6), bbox lo(3), bbox hi(3), radius] (17 f64).
void orc_bh_tree(void* h, uint32_t* order, uint32_t* start, uint32_t* end, uint32_t* right, double* cl) {
  auto* bh = static_cast<BhTree*>(h);
  const auto& t = bh->tree;
  std::memcpy(order, t.bvh.order.data(), t.bvh.order.size() * 4);
  for (size_t n = 0; n < t.bvh.nodes.size(); ++n) {
    start[n] = t.bvh.nodes[n].start;
    end[n] = t.bvh.nodes[n].end;
    right[n] = t.bvh.nodes[n].right;
    const Cluster& c = t.clusters[n];
    double* o = cl + 17 * n;
    o[0] = c.area;
    for (int k = 0; k < 3; ++k) o[1 + k] = c.centroid[k];
    for (int k = 0; k < 6; ++k) o[4 + k] = c.projector[k];
    for (int k = 0; k < 3; ++k) o[10 + k] = c.bbox.lo[k];
    for (int k = 0; k < 3; ++k) o[13 + k] = c.bbox.hi[k];
    o[16] = c.radius;
  }
}

void orc_bh_lists(void* h, uint64_t* lumped_off, uint32_t* lumped, uint64_t* exact_off, uint32_t* exact) {
  auto* bh = static_cast<BhTree*>(h);
  const auto& L = bh->lists;
  std::memcpy(lumped_off, L.lumped_off.data(), L.lumped_off.size() * 8);
  std::memcpy(lumped, L.lumped.data(), L.lumped.size() * 4);
  std::memcpy(exact_off, L.exact_off.data(), L.exact_off.size() * 8);
  std::memcpy(exact, L.exact.data(), L.exact.size() * 4);
}

int orc_bh_energy(void* h, double alpha, double beta, int threads, double* energy) {
  Error err;
  Exponents e;
  if (!make_exps(alpha, beta, &e, &err)) return err.status;
  *energy = bh_energy(*static_cast<BhTree*>(h), e, threads);
  return OK;
}

int orc_bh_energy_and_differential(void* h, double alpha, double beta, int threads, double* energy,
                                   double* diff, double* partials) {
  Error err;
  Exponents e;
  if (!make_exps(alpha, beta, &e, &err)) return err.status;
  auto* bh = static_cast<BhTree*>(h);
  if (partials) write_partials(bh_face_partials(*bh, e, threads), partials);
  std::vector<V3> d;
  if (!bh_energy_and_differential(*bh, e, threads, energy, &d, &err)) return err.status;
  for (int64_t v = 0; v < bh->mesh.nv; ++v)
    for (int k = 0; k < 3; ++k) diff[3 * v + k] = d[v][k];
  return OK;
}

// ---- hierarchical metric ----

void* orc_metric_build(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, double alpha, double beta,
                       double chi, int64_t leaf_size, int64_t leaf_pairs, int threads, int* status, char* msg,
                       int64_t msglen) {
  Error err;
  Exponents e;
  if (!make_exps(alpha, beta, &e, &err)) {
    *status = fail(err, msg, msglen, nullptr);
    return nullptr;
  }
  auto* fm = new FaceMetric();
  if (!face_metric_build(make_mesh(nv, nf, faces), make_pos(nv, pos), e, chi, (size_t)leaf_size, (size_t)leaf_pairs,
                         threads, fm, &err)) {
    delete fm;
    *status = fail(err, msg, msglen, nullptr);
    return nullptr;
  }
  *status = OK;
  return fm;
}

void orc_metric_free(void* h) { delete static_cast<FaceMetric*>(h); }

void orc_metric_apply(void* h, double w_hi, double w_lo, int64_t nb, const double* xs, double* ys, int threads) {
  face_metric_apply(*static_cast<FaceMetric*>(h), w_hi, w_lo, (size_t)nb, xs, ys, threads);
}

// [admissible_blocks, inadmissible_blocks, admissible_pairs, near_entries, nodes, depth, near_total]
void orc_metric_stats(void* h, int64_t* out) {
  const auto& s = static_cast<FaceMetric*>(h)->bct.stats;
  out[0] = (int64_t)s.admissible_blocks;
  out[1] = (int64_t)s.inadmissible_blocks;
  out[2] = (int64_t)s.admissible_pairs;
  out[3] = (int64_t)s.near_entries;
  out[4] = (int64_t)s.nodes;
  out[5] = (int64_t)s.depth;
  out[6] = (int64_t)static_cast<FaceMetric*>(h)->bct.near_total;
}

// H̃a per face (2 per face: hi, lo), original order.
// A block tree's node arrays and grouped lists (block_tree.rs:129-209); any pointer may be null;
// counts4 = [elements, nodes, admissible blocks, inadmissible blocks].
static void block_lists(const BlockTree& t, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end,
                        uint32_t* right, uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr, uint32_t* near_col) {
  const size_t nn = t.start.size();
  if (counts4) {
    counts4[0] = (int64_t)t.order.size();
    counts4[1] = (int64_t)nn;
    counts4[2] = (int64_t)t.far_col.size();
    counts4[3] = (int64_t)t.near_col.size();
  }
  auto put = [](uint32_t* out, const std::vector<uint32_t>& v) {
    if (out) std::copy(v.begin(), v.end(), out);
  };
  auto put64 = [](uint32_t* out, const std::vector<uint64_t>& v) {
    if (out)
      for (size_t i = 0; i < v.size(); ++i) out[i] = (uint32_t)v[i];
  };
  put(order, t.order);
  put(start, t.start);
  put(end, t.end);
  put(right, t.right);
  put64(far_ptr, t.far_ptr);
  put(far_col, t.far_col);
  put64(near_ptr, t.near_ptr);
  put(near_col, t.near_col);
}

void orc_metric_blocks(void* h, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end, uint32_t* right,
                       uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr, uint32_t* near_col) {
  block_lists(static_cast<FaceMetric*>(h)->bct, counts4, order, start, end, right, far_ptr, far_col, near_ptr, near_col);
}

void orc_vfrac_blocks(void* h, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end, uint32_t* right,
                      uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr, uint32_t* near_col) {
  block_lists(static_cast<VertexFractional*>(h)->bct, counts4, order, start, end, right, far_ptr, far_col, near_ptr,
              near_col);
}

void orc_metric_ha(void* h, double* out) {
  auto* fm = static_cast<FaceMetric*>(h);
  std::memcpy(out, fm->ha.data(), fm->ha.size() * 8);
}

void* orc_vfrac_build(int64_t nv, const double* pos, const double* mass, double tau, double chi, int64_t leaf_size,
                      int64_t leaf_pairs, int threads, int* status) {
  Error err;
  auto* vf = new VertexFractional();
  std::vector<double> m(mass, mass + nv);
  if (!vertex_fractional_build(make_pos(nv, pos), m, tau, chi, (size_t)leaf_size, (size_t)leaf_pairs, threads, vf, &err)) {
    delete vf;
    *status = err.status;
    return nullptr;
  }
  *status = OK;
  return vf;
}

void orc_vfrac_free(void* h) { delete static_cast<VertexFractional*>(h); }

void orc_vfrac_apply(void* h, int64_t nb, const double* xs, double* ys, int threads) {
  vertex_fractional_apply(*static_cast<VertexFractional*>(h), (size_t)nb, xs, ys, threads);
}

void orc_vertex_areas(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, double* out) {
  Mesh m = make_mesh(nv, nf, faces);
  FaceGeometry g = face_geometry(m, make_pos(nv, pos));
  auto va = vertex_areas(m, g.areas);
  std::memcpy(out, va.data(), va.size() * 8);
}

// ---- saddle solver, preconditioner, GMRES, descent step ----

void* orc_saddle_build(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, double alpha, double beta,
                       double chi, int64_t leaf_size, int64_t leaf_pairs, double eps_rel, double c_lo, int threads,
                       int* status, int64_t form, int64_t scale, int64_t refine) {
  Error err;
  Exponents e;
  if (!make_exps(alpha, beta, &e, &err)) {
    *status = err.status;
    return nullptr;
  }
  SaddleOptions o;
  o.chi = chi;
  o.leaf_size = (size_t)leaf_size;
  o.leaf_pairs = (size_t)leaf_pairs;
  o.eps_rel = eps_rel;
  o.c_lo = c_lo;
  o.form = (int)form;
  o.scale = (int)scale;
  o.refine = refine != 0;
  auto* sd = new Saddle();
  if (!saddle_build(make_mesh(nv, nf, faces), make_pos(nv, pos), e, o, threads, sd, &err)) {
    delete sd;
    *status = err.status;
    return nullptr;
  }
  *status = OK;
  return sd;
}

void orc_saddle_free(void* h) { delete static_cast<Saddle*>(h); }
double orc_saddle_eps(void* h) { return static_cast<Saddle*>(h)->eps; }
double orc_saddle_frac_scale(void* h) { return static_cast<Saddle*>(h)->frac_scale; }
void orc_saddle_apply(void* h, const double* x, double* y, int threads) { saddle_apply(*static_cast<Saddle*>(h), x, y, threads); }
void orc_precond_apply(void* h, const double* r, double* z, int threads) { precond_apply(*static_cast<Saddle*>(h), r, z, threads); }

// x = L′⁻¹ r for a full vector (3 nv + 3): the scalar saddle per axis.
void orc_lprime_solve(void* h, const double* r, double* x) {
  auto* sd = static_cast<Saddle*>(h);
  const size_t nv = sd->nv, nd = 3 * nv;
  std::vector<double> b(nv + 1), y(nv + 1);
  for (int k = 0; k < 3; ++k) {
    for (size_t v = 0; v < nv; ++v) b[v] = r[3 * v + k];
    b[nv] = r[nd + k];
    ldlt_solve(sd->lprime, b.data(), y.data());
    for (size_t v = 0; v < nv; ++v) x[3 * v + k] = y[v];
    x[nd + k] = y[nv];
  }
}

// report: [iterations, cycles, converged, relative residual, operator applications, preconditioner applications];
// with history outputs: at most cap_c cycle residuals and cap_h Arnoldi estimates (gmres.rs:214-217),
// counts = [cycle residuals, estimates]
void orc_saddle_solve_history(void* h, const double* rhs, double* x, int64_t restart, double tol, int64_t maxit,
                              int threads, double* report, double* cycle_residuals, int64_t cap_c, double* history,
                              int64_t cap_h, int64_t* counts) {
  auto* sd = static_cast<Saddle*>(h);
  const size_t n = 3 * sd->nv + 3;
  std::vector<double> b(rhs, rhs + n), sol;
  GmresOptions o;
  o.restart = (size_t)restart;
  o.tolerance = tol;
  o.max_iterations = (size_t)maxit;
  auto op = [&](const double* a, double* y) { saddle_apply(*sd, a, y, threads); };
  auto pc = [&](const double* r, double* z) { precond_apply(*sd, r, z, threads); };
  GmresReport rep = gmres_right(op, pc, b, o, &sol);
  std::memcpy(x, sol.data(), n * 8);
  report[0] = (double)rep.iterations;
  report[1] = (double)rep.cycles;
  report[2] = rep.converged ? 1.0 : 0.0;
  report[3] = rep.relative_residual;
  report[4] = (double)rep.operator_applications;
  report[5] = (double)rep.preconditioner_applications;
  if (counts) {
    counts[0] = (int64_t)rep.cycle_residuals.size();
    counts[1] = (int64_t)rep.history.size();
    for (int64_t i = 0; cycle_residuals && i < cap_c && i < counts[0]; ++i) cycle_residuals[i] = rep.cycle_residuals[i];
    for (int64_t i = 0; history && i < cap_h && i < counts[1]; ++i) history[i] = rep.history[i];
  }
}

void orc_saddle_solve(void* h, const double* rhs, double* x, int64_t restart, double tol, int64_t maxit, int threads,
                      double* report) {
  orc_saddle_solve_history(h, rhs, x, restart, tol, maxit, threads, report, nullptr, 0, nullptr, 0, nullptr);
}

// opts_f: [θ, χ, tolerance, start value, Armijo c, τ_min rel, c_lo, target of row 1, of row 2
// (NaN: the value at creation)];
// opts_i: [restart, max iterations, exact energy, start kind, max halvings (< 0: none), leaf size,
// row 1 and row 2 after the barycenter (0 none, 1 area, 2 volume), Newton (0: until tolerance,
// 1: fixed), Newton rounds]
void* orc_session_create(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, double alpha, double beta,
                         const double* opts_f, const int64_t* opts_i, int* status) {
  Error err;
  Exponents e;
  if (!make_exps(alpha, beta, &e, &err)) {
    *status = err.status;
    return nullptr;
  }
  SessionOptions o;
  o.theta = opts_f[0];
  o.saddle.chi = opts_f[1];
  o.gmres.tolerance = opts_f[2];
  o.start_value = opts_f[3];
  o.armijo_c = opts_f[4];
  o.tau_min_rel = opts_f[5];
  o.saddle.c_lo = opts_f[6];
  o.gmres.restart = (size_t)opts_i[0];
  o.gmres.max_iterations = (size_t)opts_i[1];
  o.energy_exact = opts_i[2] != 0;
  o.start_kind = (int)opts_i[3];
  o.max_halvings = (long)opts_i[4];
  o.saddle.leaf_size = (size_t)opts_i[5];
  for (int r = 0; r < 2; ++r)
    if (opts_i[6 + r] == ROW_AREA || opts_i[6 + r] == ROW_VOLUME) {
      o.rows.push_back((int)opts_i[6 + r]);
      o.targets.push_back(opts_f[7 + r]);
    }
  o.newton_fixed = opts_i[8] != 0;
  o.newton_count = (int)opts_i[9];
  o.ccd = opts_i[10] != 0;
  o.ccd_opts.gap = opts_f[9];
  o.ccd_opts.reference_length = opts_f[10];
  o.ccd_opts.close_factor = opts_f[11];
  o.ccd_opts.max_iterations = (uint32_t)opts_i[11];
  o.ccd_opts.leaf_size = (size_t)opts_i[12];
  if (o.ccd) {
    Error cerr;
    if (!ccd_validate(o.ccd_opts, &cerr)) {
      *status = cerr.status;
      return nullptr;
    }
  }
  auto* s = new Session();
  if (!session_create(make_mesh(nv, nf, faces), make_pos(nv, pos), e, o, s, &err)) {
    delete s;
    *status = err.status;
    return nullptr;
  }
  *status = OK;
  return s;
}

void orc_session_free(void* h) { delete static_cast<Session*>(h); }

// report: [energy0, energy1, slope, tau, accepted, trials, gmres iterations, gmres converged,
// gmres relres, energy after the projection, τ₀, ‖g‖, ‖x‖, ‖x‖_∞, halvings, ℓ̄₀, then the wall
// seconds of trees + F, dF / the solver's build / GMRES / the line search / the projection];
// trials (cap × 3): per trial [τ, F, outcome (0 accepted, 1 Armijo failed, 2 non-finite)]
int orc_session_step(void* h, int threads, double* report, double* trials, int64_t cap) {
  Error err;
  StepReport rep;
  auto* s = static_cast<Session*>(h);
  if (!session_step(*s, threads, &rep, &err)) return err.status;
  report[0] = rep.energy0;
  report[1] = rep.energy1;
  report[2] = rep.slope;
  report[3] = rep.tau;
  report[4] = rep.accepted ? 1.0 : 0.0;
  report[5] = (double)rep.trials;
  report[6] = (double)rep.gmres.iterations;
  report[7] = rep.gmres.converged ? 1.0 : 0.0;
  report[8] = rep.gmres.relative_residual;
  report[9] = rep.energy_after;
  report[10] = rep.tau0;
  report[11] = rep.gradient_norm;
  report[12] = rep.direction_norm;
  report[13] = rep.direction_max;
  report[14] = (double)rep.halvings;
  report[15] = s->l0;
  for (int k = 0; k < 5; ++k) report[16 + k] = rep.seconds[k];
  for (size_t t = 0; t < rep.trial.size() && (int64_t)t < cap; ++t) {
    trials[3 * t] = rep.trial[t].tau;
    trials[3 * t + 1] = rep.trial[t].energy;
    trials[3 * t + 2] = (double)rep.trial[t].outcome;
  }
  return OK;
}

// out: [rows (3 + nonlinear), Newton rounds, smallest kept Schur pivot, Φ before the projection
// (rows), Φ after it (rows)] of the last step; cap doubles at most
void orc_session_constraints(void* h, double* out, int64_t cap) {
  auto* s = static_cast<Session*>(h);
  std::vector<double> v{(double)s->last_phi.size(), (double)s->last_newton_rounds, s->last_min_pivot};
  v.insert(v.end(), s->last_phi_before.begin(), s->last_phi_before.end());
  v.insert(v.end(), s->last_phi.begin(), s->last_phi.end());
  for (size_t i = 0; i < v.size() && (int64_t)i < cap; ++i) out[i] = v[i];
}

// The Schur complement of k rows (c: k dense rows of nd, z: the k columns A′⁻¹ C_Nᵀ, position
// blocks): x = y − Z N⁻¹ C_N y, h = Z N⁻¹ Φ_N, S (k × k as computed), the smallest kept pivot.
void orc_schur(int64_t k, int64_t nd, const double* c, const double* z, const double* y, const double* phi,
               double* x_out, double* h_out, double* s_out, double* min_pivot) {
  std::vector<std::vector<double>> cr(k), zc(k);
  for (int64_t r = 0; r < k; ++r) {
    cr[r].assign(c + r * nd, c + (r + 1) * nd);
    zc[r].assign(z + r * nd, z + (r + 1) * nd);
  }
  Schur S;
  schur_build(&S, cr, zc);
  const auto x = schur_direction(S, std::vector<double>(y, y + nd));
  const auto h = schur_correction(S, std::vector<double>(phi, phi + k));
  std::copy(x.begin(), x.end(), x_out);
  std::copy(h.begin(), h.end(), h_out);
  std::copy(S.s.begin(), S.s.end(), s_out);
  *min_pivot = S.min_pivot;
}

// Values (rows of kind 1 area, 2 volume) and their gradients (3 nv each) at positions.
void orc_constraint_rows(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, const int64_t* kinds,
                         int64_t k, double* values, double* grads) {
  const Mesh m = make_mesh(nv, nf, faces);
  const Positions p = make_pos(nv, pos);
  for (int64_t r = 0; r < k; ++r) {
    values[r] = row_value(m, p, (int)kinds[r]);
    const auto g = row_gradient(m, p, (int)kinds[r]);
    std::copy(g.begin(), g.end(), grads + r * 3 * nv);
  }
}

void orc_session_positions(void* h, double* out) {
  auto* s = static_cast<Session*>(h);
  for (size_t v = 0; v < s->positions.size(); ++v)
    for (int k = 0; k < 3; ++k) out[3 * v + k] = s->positions[v][k];
}

double orc_fractional_symbol_ratio(double s) { return fractional_symbol_ratio(s); }

// D_f u and the average ū per face (metric/mod.rs:200 `average`, :210 `gradient` with the
// coefficients of :120 `df_coefficients`); avg: nf, grad: 3·nf.
void orc_face_operators(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, const double* u,
                        double* avg, double* grad) {
  Mesh m = make_mesh(nv, nf, faces);
  Positions p = make_pos(nv, pos);
  for (int64_t f = 0; f < nf; ++f) {
    const auto& t = m.faces[f];
    avg[f] = (u[t[0]] + u[t[1]] + u[t[2]]) / 3.0;
    const auto g = df_coefficients(corners(p, t));
    double out[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) out[r] += u[t[c]] * g[c][r];
    for (int r = 0; r < 3; ++r) grad[3 * f + r] = out[r];
  }
}

// ---- CCD -------------------------------------------------------------------------------------

// The last step's CCD report in the layout of orc_ccd_first_contact; returns 1 when it ran.
int orc_session_ccd(void* h, double* out_f, int64_t* out_i) {
  auto* s = static_cast<Session*>(h);
  const CcdReport& r = s->last_ccd;
  out_f[0] = r.tau_max;
  out_f[1] = r.gap_abs;
  out_f[2] = r.contact.tau;
  const int64_t vals[10] = {r.has_contact, r.contact.kind, r.contact.a, r.contact.b, (int64_t)r.iterations,
                            (int64_t)r.pairs_tested, r.windows, (int64_t)r.close_pairs, (int64_t)r.touching_pairs,
                            (int64_t)r.exempt_pairs};
  for (int k = 0; k < 10; ++k) out_i[k] = vals[k];
  return s->last_has_ccd ? 1 : 0;
}

// opts_f: gap, reference_length (NaN: none), close_factor; opts_i: max_iterations, leaf_size.
// band: nv flags or NULL.  out_f: tau_max, gap_abs, contact tau; out_i: has_contact, kind, a, b,
// iterations, pairs_tested, windows, close_pairs, touching_pairs, exempt_pairs.
int orc_ccd_first_contact(int64_t nv, int64_t nf, const uint32_t* faces, const double* start, const double* vel,
                          double tau_end, const double* opts_f, const int64_t* opts_i, const uint8_t* band,
                          double* out_f, int64_t* out_i) {
  Mesh m = make_mesh(nv, nf, faces);
  Positions s = make_pos(nv, start), u = make_pos(nv, vel);
  CcdOptions o;
  o.gap = opts_f[0];
  o.reference_length = opts_f[1];
  o.close_factor = opts_f[2];
  o.max_iterations = (uint32_t)opts_i[0];
  o.leaf_size = (size_t)opts_i[1];
  std::vector<uint8_t> b;
  if (band) b.assign(band, band + nv);
  CcdReport r;
  Error err;
  if (!first_contact(m, s, u, tau_end, o, band ? &b : nullptr, &r, &err)) return err.status;
  out_f[0] = r.tau_max;
  out_f[1] = r.gap_abs;
  out_f[2] = r.contact.tau;
  const int64_t vals[10] = {r.has_contact, r.contact.kind, r.contact.a, r.contact.b, (int64_t)r.iterations,
                            (int64_t)r.pairs_tested, r.windows, (int64_t)r.close_pairs, (int64_t)r.touching_pairs,
                            (int64_t)r.exempt_pairs};
  for (int k = 0; k < 10; ++k) out_i[k] = vals[k];
  return 0;
}

// The sorted intersecting face pairs: at most cap written to pairs (2 each); returns the count.
int64_t orc_intersecting_pairs(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, uint32_t* pairs,
                               int64_t cap) {
  const auto out = intersecting_pairs(make_mesh(nv, nf, faces), make_pos(nv, pos));
  for (size_t i = 0; i < out.size() && (int64_t)i < cap; ++i) {
    pairs[2 * i] = out[i][0];
    pairs[2 * i + 1] = out[i][1];
  }
  return (int64_t)out.size();
}

// The band flags (nv); counts2 = [intersecting pairs, vertices in the band].
void orc_intersection_band(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, uint8_t* flags,
                           int64_t* counts2) {
  size_t np = 0, nb = 0;
  const auto b = intersection_band(make_mesh(nv, nf, faces), make_pos(nv, pos), &np, &nb);
  std::copy(b.begin(), b.end(), flags);
  counts2[0] = (int64_t)np;
  counts2[1] = (int64_t)nb;
}

double orc_min_distance(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos) {
  return min_distance(make_mesh(nv, nf, faces), make_pos(nv, pos));
}

// kind 0: point p (x[0..3]) to triangle x[3..12]; 1: segment x[0..6] to segment x[6..12];
// 2: point to segment x[3..9]; 3: segment x[0..6] meets triangle x[6..15] (1.0 or 0.0);
// 4: triangles x[0..9] and x[9..18] intersect.
double orc_ccd_primitive(int kind, const double* x) {
  auto v = [&](int i) { return V3{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; };
  switch (kind) {
    case 0: return point_triangle_distance(v(0), v(1), v(2), v(3));
    case 1: return segment_segment_distance(v(0), v(1), v(2), v(3));
    case 2: return point_segment_distance(v(0), v(1), v(2));
    case 3: return segment_meets_triangle(v(0), v(1), {v(2), v(3), v(4)}) ? 1.0 : 0.0;
    default: return triangles_intersect({v(0), v(1), v(2)}, {v(3), v(4), v(5)}) ? 1.0 : 0.0;
  }
}


// ---- remeshing -------------------------------------------------------------------------------
// Options: f = [split, collapse, rho, max_growth], i = [rounds, flip_cap_per_vertex, order,
// split on, collapse on, flip on, smooth on].  Report (47 int64): splits, collapses, flips,
// splits_deferred, smoothing moves / halved / refused / skipped, flip_cap_hits, n_vertices, n_faces,
// guard refused / tested / exempt, then refused splits, collapses, flips (11 kinds each, in the
// order of mesh.rs:75 RefusalKind::ALL).
namespace {
RemeshOptions remesh_options(const double* f, const int64_t* i) {
  RemeshOptions o;
  o.split = f[0];
  o.collapse = f[1];
  o.rho = f[2];
  o.max_growth = f[3];
  o.rounds = (int)i[0];
  o.flip_cap_per_vertex = (size_t)i[1];
  o.order = (int)i[2];
  o.en_split = i[3] != 0;
  o.en_collapse = i[4] != 0;
  o.en_flip = i[5] != 0;
  o.en_smooth = i[6] != 0;
  return o;
}
void write_remesh_report(const RemeshReport& r, int64_t* out) {
  const int64_t head[14] = {(int64_t)r.splits,          (int64_t)r.collapses,         (int64_t)r.flips,
                            (int64_t)r.splits_deferred, (int64_t)r.smoothing_moves,   (int64_t)r.smoothing_halved,
                            (int64_t)r.smoothing_refused, (int64_t)r.smoothing_skipped, (int64_t)r.flip_cap_hits,
                            (int64_t)r.n_vertices,      (int64_t)r.n_faces,           (int64_t)r.guard_refused,
                            (int64_t)r.guard_tested,    (int64_t)r.guard_exempt};
  for (int k = 0; k < 14; ++k) out[k] = head[k];
  for (int k = 0; k < 11; ++k) {
    out[14 + k] = (int64_t)r.refused_splits[k];
    out[25 + k] = (int64_t)r.refused_collapses[k];
    out[36 + k] = (int64_t)r.refused_flips[k];
  }
}
}  // namespace

void* orc_remesh(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, int64_t npin, const uint32_t* pins,
                 double l0, const double* opts_f, const int64_t* opts_i, double guard_gap, const uint8_t* band,
                 int64_t* report, int* status, char* msg, int64_t msglen) {
  Error err;
  std::vector<uint32_t> pv(pins, pins + npin);
  std::vector<uint8_t> bv;
  if (band) bv.assign(band, band + nv);
  auto* out = new Remeshed();
  RemeshReport rep;
  if (!remesh(make_mesh(nv, nf, faces), make_pos(nv, pos), pv, l0, remesh_options(opts_f, opts_i), guard_gap,
              band ? &bv : nullptr, out, &rep, &err)) {
    delete out;
    *status = fail(err, msg, msglen, nullptr);
    return nullptr;
  }
  write_remesh_report(rep, report);
  *status = OK;
  return out;
}

// Single operation (op 0 split, 1 collapse, 2 flip) then finish; *refusal the RemeshRefusal.
void* orc_remesh_op(int64_t nv, int64_t nf, const uint32_t* faces, const double* pos, int64_t npin,
                    const uint32_t* pins, int op, uint32_t a, uint32_t b, int delaunay, int foldover, int* refusal,
                    int* reordered, int* status, char* msg, int64_t msglen) {
  Error err;
  std::vector<uint32_t> pv(pins, pins + npin);
  EditMesh* m = nullptr;
  if (!edit_mesh_new(make_mesh(nv, nf, faces), make_pos(nv, pos), pv, &m, &err)) {
    *status = fail(err, msg, msglen, nullptr);
    return nullptr;
  }
  bool ro = false;
  *refusal = edit_mesh_op(*m, op, a, b, delaunay != 0, foldover != 0, &ro);
  *reordered = ro ? 1 : 0;
  auto* out = new Remeshed();
  const bool ok = edit_mesh_finish(*m, out, &err);
  edit_mesh_free(m);
  if (!ok) {
    delete out;
    *status = fail(err, msg, msglen, nullptr);
    return nullptr;
  }
  *status = OK;
  return out;
}

// sizes: nv, nf, entries of the vertex map, pins
void orc_remeshed_sizes(void* h, int64_t* sizes) {
  auto* r = (Remeshed*)h;
  sizes[0] = (int64_t)r->positions.size();
  sizes[1] = (int64_t)r->faces.size();
  sizes[2] = (int64_t)r->map_sources.size();
  sizes[3] = (int64_t)r->pinned.size();
}

void orc_remeshed_read(void* h, double* pos, uint32_t* faces, uint32_t* parents, uint32_t* map_offsets,
                       uint32_t* map_sources, double* map_weights, uint32_t* pins) {
  auto* r = (Remeshed*)h;
  for (size_t v = 0; v < r->positions.size(); ++v)
    for (int k = 0; k < 3; ++k) pos[3 * v + k] = r->positions[v][k];
  for (size_t f = 0; f < r->faces.size(); ++f) {
    for (int k = 0; k < 3; ++k) faces[3 * f + k] = r->faces[f][k];
    parents[f] = r->face_parents[f];
  }
  std::copy(r->map_offsets.begin(), r->map_offsets.end(), map_offsets);
  std::copy(r->map_sources.begin(), r->map_sources.end(), map_sources);
  std::copy(r->map_weights.begin(), r->map_weights.end(), map_weights);
  std::copy(r->pinned.begin(), r->pinned.end(), pins);
}

void orc_remeshed_free(void* h) { delete (Remeshed*)h; }

// Geometry predicates of the remesher (geom.rs): kind 0 circumcenter (x: 9 → out 3), 1 Delaunay
// test (x: p q c d, 12 → out[0] 0/1), 2 folds (x: before, after, 6 → out[0] 0/1), 3 triangle
// distance (x: 18 → out[0]).
void orc_remesh_geom(int kind, const double* x, double* out) {
  auto v = [&](int i) { return V3{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; };
  switch (kind) {
    case 0: {
      const V3 c = remesh_circumcenter({v(0), v(1), v(2)});
      for (int k = 0; k < 3; ++k) out[k] = c[k];
      break;
    }
    case 1:
      out[0] = remesh_violates_delaunay(v(0), v(1), v(2), v(3)) ? 1.0 : 0.0;
      break;
    case 2:
      out[0] = remesh_folds(v(0), v(1)) ? 1.0 : 0.0;
      break;
    default:
      out[0] = remesh_triangle_distance({v(0), v(1), v(2)}, {v(3), v(4), v(5)});
  }
}

}  // extern "C"
