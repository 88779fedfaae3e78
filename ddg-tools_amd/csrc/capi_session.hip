// C ABI of the step's saddle solver and of the descent session.
#include <cmath>
#include <cstring>

#include "capi_util.hpp"
#include "robust.hpp"
#include "session.hpp"

using namespace ddgrs;

struct ddgrs_saddle {
  ddgrs_surface* s = nullptr;
  FaceMetricDev fm;
  Preconditioner pc;
  SaddleOp op;
  DeviceVecs vec;
  GmresResult last;  // the last solve's report, with its cycle residuals and history
};

struct ddgrs_session {
  ddgrs_surface* s = nullptr;
  Session ses;
};

namespace {
void check_saddle_options(const ddgrs_saddle_options& o) {
  if (!(std::isfinite(o.chi) && o.chi >= 0.0)) invalid("chi", std::to_string(o.chi), "a finite number >= 0 (0: exact products)");
  if (o.leaf_size < 1) invalid("leaf_size", std::to_string(o.leaf_size), "at least 1");
  if (o.leaf_pairs < 0) invalid("leaf_pairs", std::to_string(o.leaf_pairs), "non-negative");
  if (!(std::isfinite(o.w_tp) && o.w_tp > 0.0)) invalid("w_tp", std::to_string(o.w_tp), "a finite weight > 0");
  if (!(std::isfinite(o.c_lo) && o.c_lo >= 0.0)) invalid("c_lo", std::to_string(o.c_lo), "a finite weight >= 0");
  if (!(std::isfinite(o.eps_rel) && o.eps_rel > 0.0)) invalid("eps_rel", std::to_string(o.eps_rel), "a finite number > 0");
  if (o.form != 0 && o.form != 1) invalid("form", std::to_string(o.form), "0 (augmented) or 1 (congruence)");
  if (o.scale != 0 && o.scale != 1) invalid("scale", std::to_string(o.scale), "0 (symbol) or 1 (unit)");
  if (o.refine != 0 && o.refine != 1) invalid("refine", std::to_string(o.refine), "0 or 1");
}

// gmres.rs:168 GmresOptions::validate
void check_gmres(int64_t restart, double tol, int64_t maxit) {
  if (restart < 1) invalid("gmres.restart", std::to_string(restart), "at least 1");
  if (!(std::isfinite(tol) && tol > 0.0)) invalid("gmres.tolerance", std::to_string(tol), "a finite number > 0");
  if (maxit < 1) invalid("gmres.max_iterations", std::to_string(maxit), "at least 1");
}

template <class F>
void batch(ddgrs_ctx* ctx, int64_t nb, int64_t n, const double* xs, double* ys, F&& apply) {
  DevBuf<double> dx(n), dy(n);
  for (int64_t q = 0; q < nb; ++q) {
    HIPCHECK(hipMemcpyAsync(dx.ptr, xs + q * n, n * 8, hipMemcpyHostToDevice, ctx->stream));
    apply(dx.ptr, dy.ptr);
    HIPCHECK(hipMemcpyAsync(ys + q * n, dy.ptr, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  ctx->sync();
}

void copy_history(const GmresResult& r, double* cycle_residuals, int64_t cap_cycles, double* history,
                  int64_t cap_history, int64_t* counts2) {
  counts2[0] = (int64_t)r.cycle_residuals.size();
  counts2[1] = (int64_t)r.history.size();
  for (int64_t i = 0; cycle_residuals && i < cap_cycles && i < counts2[0]; ++i) cycle_residuals[i] = r.cycle_residuals[i];
  for (int64_t i = 0; history && i < cap_history && i < counts2[1]; ++i) history[i] = r.history[i];
}

void fill_report(const GmresResult& r, ddgrs_gmres_report* out) {
  out->iterations = (int64_t)r.iterations;
  out->cycles = (int64_t)r.cycles;
  out->operator_applications = (int64_t)r.operator_applications;
  out->preconditioner_applications = (int64_t)r.preconditioner_applications;
  out->converged = r.converged ? 1 : 0;
  out->extended = r.extended ? 1 : 0;
  out->relative_residual = r.relative_residual;
}
}  // namespace

extern "C" {

void ddgrs_saddle_options_default(ddgrs_saddle_options* o) {
  if (!o) return;
  o->chi = 0.5;
  o->leaf_size = 8;
  o->leaf_pairs = 64;
  o->w_tp = 1.0;
  o->c_lo = 1.0;
  o->eps_rel = 1e-10;
  o->form = 0;
  o->scale = 0;
  o->refine = 0;
}

int ddgrs_saddle_build(ddgrs_surface* s, double alpha, double beta, const ddgrs_saddle_options* o, ddgrs_saddle** out,
                       ddgrs_error* err) {
  return guarded(err, "saddle_build", [&] {
    need(s, "surface");
    need(o, "options");
    need(out, "out");
    *out = nullptr;
    Exps e = make_exps(alpha, beta);
    check_saddle_options(*o);
    auto h = std::make_unique<ddgrs_saddle>();
    h->s = s;
    const double s_exp = (e.beta64 - 2.0) / e.alpha64;
    Topology t;
    Clusters c;
    build_face_tree(s, (size_t)o->leaf_size, &t, &c);
    h->fm.w_hi = o->w_tp;
    h->fm.w_lo = o->w_tp * o->c_lo;
    h->fm.build(s, t, c, o->chi, (size_t)o->leaf_pairs, s_exp);
    h->pc.form = o->form;
    h->pc.unit_scale = o->scale == 1;
    h->pc.refine = o->refine == 1;
    h->pc.build(s, s_exp, o->w_tp, o->eps_rel, o->chi, (size_t)o->leaf_size, (size_t)o->leaf_pairs);
    h->op.fm = &h->fm;
    h->op.crow = h->pc.crow_dev;
    h->op.nv = (size_t)s->nv;
    h->vec.init(3 * (size_t)s->nv + 3);
    surface_retain(s);
    *out = h.release();
  });
}

void ddgrs_saddle_destroy(ddgrs_saddle* h) { destroy_handle(h); }

int ddgrs_saddle_info(ddgrs_saddle* h, double* out2, ddgrs_error* err) {
  return guarded(err, "saddle_info", [&] {
    need(h, "saddle");
    need(out2, "out");
    out2[0] = h->pc.eps_value(h->s->ctx);
    out2[1] = h->pc.scale();
  });
}

int ddgrs_saddle_apply_batch(ddgrs_saddle* h, int64_t nb, const double* xs, double* ys, ddgrs_error* err) {
  return guarded(err, "saddle_apply", [&] {
    need(h, "saddle");
    need(xs, "xs");
    need(ys, "ys");
    ddgrs_ctx* ctx = h->s->ctx;
    batch(ctx, nb, 3 * h->s->nv + 3, xs, ys, [&](const double* x, double* y) { h->op.apply(ctx, x, y); });
  });
}

int ddgrs_saddle_precond_batch(ddgrs_saddle* h, int64_t nb, const double* rs, double* zs, ddgrs_error* err) {
  return guarded(err, "saddle_precond", [&] {
    need(h, "saddle");
    need(rs, "rs");
    need(zs, "zs");
    ddgrs_ctx* ctx = h->s->ctx;
    batch(ctx, nb, 3 * h->s->nv + 3, rs, zs, [&](const double* r, double* z) { h->pc.apply(ctx, r, z); });
  });
}

int ddgrs_saddle_lprime_batch(ddgrs_saddle* h, int64_t nb, const double* rs, double* xs, ddgrs_error* err) {
  return guarded(err, "saddle_lprime", [&] {
    need(h, "saddle");
    need(rs, "rs");
    need(xs, "xs");
    ddgrs_ctx* ctx = h->s->ctx;
    batch(ctx, nb, 3 * h->s->nv + 3, rs, xs, [&](const double* r, double* x) { h->pc.lprime_solve(ctx, r, x); });
  });
}

namespace {
// One right-preconditioned GMRES solve of A′ x = b on device vectors (x starts at zero); the
// extended products take over when the FP32 products stall.
GmresResult saddle_gmres(ddgrs_saddle* h, const double* b, double* x, const GmresSettings& gs) {
  ddgrs_ctx* ctx = h->s->ctx;
  const size_t n = 3 * (size_t)h->s->nv + 3;
  auto A = [&](const double* u, double* y) { h->op.apply(ctx, u, y); };
  auto P = [&](const double* r, double* z) { h->pc.apply(ctx, r, z); };
  FaceMetricDev& fm = *h->op.fm;
  fm.extended = h->pc.vf.extended = false;
  auto stall = [&]() {
    if (fm.extended) return false;
    fm.extended = h->pc.vf.extended = true;
    return true;
  };
  GmresResult r = gmres_device(ctx, n, A, P, b, x, gs, h->vec, stall);
  fm.extended = h->pc.vf.extended = false;
  return r;
}
}  // namespace

int ddgrs_saddle_solve(ddgrs_saddle* h, const double* b, double* x, int64_t restart, double tolerance,
                       int64_t max_iterations, ddgrs_gmres_report* report, ddgrs_error* err) {
  return guarded(err, "saddle_solve", [&] {
    need(h, "saddle");
    need(b, "b");
    need(x, "x");
    check_gmres(restart, tolerance, max_iterations);
    ddgrs_ctx* ctx = h->s->ctx;
    const size_t n = 3 * (size_t)h->s->nv + 3;
    DevBuf<double> db(n), dx(n);
    upload(ctx, db, b, n);
    GmresResult r = saddle_gmres(h, db.ptr, dx.ptr, GmresSettings{(size_t)restart, tolerance, (size_t)max_iterations});
    download(ctx, x, (const double*)dx.ptr, n);
    ctx->sync();
    h->last = r;
    if (report) fill_report(r, report);
  });
}

int ddgrs_saddle_solve_many(ddgrs_saddle* h, int64_t nb, const double* bs, double* xs, int64_t restart,
                            double tolerance, int64_t max_iterations, ddgrs_gmres_report* reports, ddgrs_error* err) {
  return guarded(err, "saddle_solve_many", [&] {
    need(h, "saddle");
    if (nb < 0) invalid("nb", std::to_string(nb), "non-negative");
    if (nb == 0) return;
    need(bs, "bs");
    need(xs, "xs");
    check_gmres(restart, tolerance, max_iterations);
    ddgrs_ctx* ctx = h->s->ctx;
    const size_t n = 3 * (size_t)h->s->nv + 3;
    DevBuf<double> db(n * (size_t)nb), dx(n * (size_t)nb);
    upload(ctx, db, bs, n * (size_t)nb);
    const GmresSettings gs{(size_t)restart, tolerance, (size_t)max_iterations};
    for (int64_t i = 0; i < nb; ++i) {
      GmresResult r = saddle_gmres(h, db.ptr + n * (size_t)i, dx.ptr + n * (size_t)i, gs);
      if (reports) fill_report(r, reports + i);
      h->last = std::move(r);
    }
    download(ctx, xs, (const double*)dx.ptr, n * (size_t)nb);
    ctx->sync();
  });
}

int ddgrs_saddle_gmres_history(ddgrs_saddle* h, double* cycle_residuals, int64_t cap_cycles, double* history,
                               int64_t cap_history, int64_t* counts2, ddgrs_error* err) {
  return guarded(err, "saddle_gmres_history", [&] {
    need(h, "saddle");
    need(counts2, "counts");
    copy_history(h->last, cycle_residuals, cap_cycles, history, cap_history, counts2);
  });
}

void ddgrs_step_options_default(ddgrs_step_options* o) {
  if (!o) return;
  o->alpha = 6.0;
  o->beta = 12.0;
  o->energy_exact = 0;
  o->theta = 0.5;
  ddgrs_saddle_options_default(&o->saddle);
  o->restart = 50;
  o->tolerance = 1e-6;
  o->max_iterations = 200;
  o->armijo_c = 1e-2;
  o->tau_min_rel = 1e-12;
  o->max_halvings = -1;
  o->start_kind = DDGRS_START_GROW_PREVIOUS;
  o->start_value = 1.0;
  o->robust = 1;
  o->retry_max_iterations = 0;
  o->dense_max_vertices = kDenseFallbackMaxVertices;
  o->diverged_residual = 1e-1;
  o->rows[0] = o->rows[1] = 0;
  o->row_targets[0] = o->row_targets[1] = NAN;
  o->newton_fixed = 0;
  o->newton_count = 5;
  o->ccd = 0;
  ddgrs_ccd_options_default(&o->ccd_options);
}

int ddgrs_session_create(ddgrs_surface* s, const ddgrs_step_options* o, ddgrs_session** out, ddgrs_error* err) {
  return guarded(err, "session_create", [&] {
    need(s, "surface");
    need(o, "options");
    need(out, "out");
    *out = nullptr;
    (void)make_exps(o->alpha, o->beta);
    check_saddle_options(o->saddle);
    check_gmres(o->restart, o->tolerance, o->max_iterations);
    if (!o->energy_exact && !(std::isfinite(o->theta) && o->theta >= 0.0))
      invalid("theta", std::to_string(o->theta), "a finite number >= 0");
    if (!(std::isfinite(o->armijo_c) && o->armijo_c > 0.0 && o->armijo_c < 1.0))
      invalid("armijo_c", std::to_string(o->armijo_c), "in (0, 1)");
    if (!(std::isfinite(o->tau_min_rel) && o->tau_min_rel > 0.0))
      invalid("tau_min_rel", std::to_string(o->tau_min_rel), "a finite number > 0");
    if (o->start_kind < DDGRS_START_GROW_PREVIOUS || o->start_kind > DDGRS_START_REFERENCE)
      invalid("start_kind", std::to_string(o->start_kind), "0 (grow previous), 1 (absolute), 2 (max displacement) or 3 (reference)");
    if (o->start_kind != DDGRS_START_REFERENCE && !(std::isfinite(o->start_value) && o->start_value > 0.0))
      invalid("start_value", std::to_string(o->start_value), "a finite number > 0");
    StepOptions so;
    so.alpha = o->alpha;
    so.beta = o->beta;
    so.energy_exact = o->energy_exact != 0;
    so.theta = o->theta;
    so.leaf_size = (size_t)o->saddle.leaf_size;
    so.chi = o->saddle.chi;
    so.leaf_pairs = (size_t)o->saddle.leaf_pairs;
    so.w_tp = o->saddle.w_tp;
    so.c_lo = o->saddle.c_lo;
    so.eps_rel = o->saddle.eps_rel;
    so.pc_form = o->saddle.form;
    so.pc_unit_scale = o->saddle.scale == 1;
    so.pc_refine = o->saddle.refine == 1;
    so.restart = (size_t)o->restart;
    so.tolerance = o->tolerance;
    so.max_iterations = (size_t)o->max_iterations;
    so.armijo_c = o->armijo_c;
    so.tau_min_rel = o->tau_min_rel;
    so.max_halvings = (long)o->max_halvings;
    so.start_kind = o->start_kind;
    so.start_value = o->start_value;
    if (o->retry_max_iterations < 0)
      invalid("retry_max_iterations", std::to_string(o->retry_max_iterations), "0 (the default cap) or more");
    if (o->dense_max_vertices < 0)
      invalid("dense_max_vertices", std::to_string(o->dense_max_vertices), "non-negative");
    so.robust = o->robust != 0;
    so.retry_max_iterations = (long)o->retry_max_iterations;
    so.dense_max_vertices = o->dense_max_vertices;
    if (!(o->diverged_residual >= 0.0 && std::isfinite(o->diverged_residual)))
      invalid("diverged_residual", std::to_string(o->diverged_residual), "a finite number >= 0");
    so.diverged_residual = o->diverged_residual;
    for (int r = 0; r < 2; ++r) {
      if (o->rows[r] < 0 || o->rows[r] > 2)
        invalid("rows", std::to_string(o->rows[r]), "0 (none), 1 (total area) or 2 (enclosed volume)");
      if (!std::isnan(o->row_targets[r]) && !std::isfinite(o->row_targets[r]))
        invalid("row_targets", std::to_string(o->row_targets[r]), "a finite target or NaN (the value at creation)");
      so.rows[r] = o->rows[r];
      so.row_targets[r] = o->row_targets[r];
    }
    if (o->rows[0] != 0 && o->rows[0] == o->rows[1])
      invalid("rows", std::to_string(o->rows[1]), "each nonlinear row at most once");
    if (o->newton_count < 0) invalid("newton_count", std::to_string(o->newton_count), "non-negative");
    so.newton_fixed = o->newton_fixed != 0;
    so.newton_count = o->newton_count;
    if (o->ccd != 0 && o->ccd != 1) invalid("ccd", std::to_string(o->ccd), "0 or 1");
    so.ccd = o->ccd == 1;
    if (so.ccd) {
      const ddgrs_ccd_options& c = o->ccd_options;
      if (c.max_iterations < 1 || c.max_iterations > UINT32_MAX)
        invalid("ccd.max_iterations", std::to_string(c.max_iterations), "at least 1");
      if (c.leaf_size < 1 || c.leaf_size > (1 << 20)) invalid("ccd.leaf_size", std::to_string(c.leaf_size), "at least 1");
      so.ccd_opts.gap = c.gap;
      so.ccd_opts.reference_length = c.reference_length;
      so.ccd_opts.close_factor = c.close_factor;
      so.ccd_opts.max_iterations = (uint32_t)c.max_iterations;
      so.ccd_opts.leaf_size = (uint32_t)c.leaf_size;
      check_ccd_options(so.ccd_opts);
    }
    auto h = std::make_unique<ddgrs_session>();
    h->s = s;
    h->ses.init(s, so);
    surface_retain(s);
    *out = h.release();
  });
}

void ddgrs_session_destroy(ddgrs_session* h) { destroy_handle(h); }

int ddgrs_session_step(ddgrs_session* h, ddgrs_step_report* report, ddgrs_error* err) {
  return guarded(err, "session_step", [&] {
    need(h, "session");
    ddgrs_step_report local;
    h->ses.step(report ? report : &local);
  });
}

int ddgrs_session_trials(ddgrs_session* h, int64_t capacity, double* tau, double* energy, int32_t* outcome,
                         int64_t* count, ddgrs_error* err) {
  return guarded(err, "session_trials", [&] {
    need(h, "session");
    need(count, "count");
    const auto& tr = h->ses.trials;
    *count = (int64_t)tr.size();
    for (int64_t i = 0; i < capacity && i < (int64_t)tr.size(); ++i) {
      if (tau) tau[i] = tr[i].tau;
      if (energy) energy[i] = tr[i].energy;
      if (outcome) outcome[i] = tr[i].outcome;
    }
  });
}

int ddgrs_session_gmres_history(ddgrs_session* h, double* cycle_residuals, int64_t cap_cycles, double* history,
                                int64_t cap_history, int64_t* counts2, ddgrs_error* err) {
  return guarded(err, "session_gmres_history", [&] {
    need(h, "session");
    need(counts2, "counts");
    copy_history(h->ses.last_gmres, cycle_residuals, cap_cycles, history, cap_history, counts2);
  });
}

int ddgrs_schur(int64_t k, int64_t nd, const double* c, const double* z, const double* y, const double* phi,
                double* x, double* h, double* s, double* min_pivot, ddgrs_error* err) {
  return guarded(err, "schur", [&] {
    if (k < 0 || k > 64) invalid("k", std::to_string(k), "0 to 64 rows");
    if (nd < 0) invalid("nd", std::to_string(nd), "non-negative");
    need(c, "c");
    need(z, "z");
    need(y, "y");
    need(phi, "phi");
    need(x, "x");
    need(h, "h");
    std::vector<std::vector<double>> cr((size_t)k), zc((size_t)k);
    for (int64_t r = 0; r < k; ++r) {
      cr[r].assign(c + r * nd, c + (r + 1) * nd);
      zc[r].assign(z + r * nd, z + (r + 1) * nd);
    }
    SchurRows S;
    S.build(std::move(cr), std::move(zc));
    const auto xv = S.direction(std::vector<double>(y, y + nd));
    const auto hv = S.correction(std::vector<double>(phi, phi + k));
    std::copy(xv.begin(), xv.end(), x);
    std::copy(hv.begin(), hv.end(), h);
    if (s) std::copy(S.s.begin(), S.s.end(), s);
    if (min_pivot) *min_pivot = S.min_pivot;
  });
}

int ddgrs_session_constraints(ddgrs_session* h, double* phi_before, double* phi, int64_t capacity, int64_t* nrows,
                              ddgrs_error* err) {
  return guarded(err, "session_constraints", [&] {
    need(h, "session");
    need(nrows, "nrows");
    const auto& a = h->ses.last_phi_before;
    const auto& b = h->ses.last_phi;
    *nrows = (int64_t)b.size();
    for (int64_t i = 0; i < capacity && i < (int64_t)b.size(); ++i) {
      if (phi_before) phi_before[i] = a[i];
      if (phi) phi[i] = b[i];
    }
  });
}

int ddgrs_session_fallbacks(ddgrs_session* h, char* text, int64_t capacity, int64_t* count, ddgrs_error* err) {
  return guarded(err, "session_fallbacks", [&] {
    need(h, "session");
    need(count, "count");
    const auto& ev = h->ses.fallbacks;
    *count = (int64_t)ev.size();
    if (!text || capacity <= 0) return;
    std::string all;
    for (const auto& e : ev) all += e + "\n";
    const size_t m = std::min<size_t>(all.size(), (size_t)capacity - 1);
    std::memcpy(text, all.data(), m);
    text[m] = '\0';
  });
}

int ddgrs_session_set_learned(ddgrs_session* h, ddgrs_learned* l, int32_t mode, ddgrs_error* err) {
  return guarded(err, "session_set_learned", [&] {
    need(h, "session");
    if (mode < 0 || mode > 2) invalid("mode", std::to_string(mode), "0 (plain), 1 (record pairs) or 2 ((I + Q) P⁻¹)");
    if (!l || mode == 0) {
      h->ses.learned = nullptr;
      h->ses.learned_mode = 0;
      return;
    }
    LearnedModel* m = learned_model(l);
    if (m->nv != h->s->nv || m->faces != h->s->faces_h)
      invalid("learned", "a network of another mesh", "the session's surface mesh");
    if (mode == 2 && !(m->scale > 0.f)) invalid("learned", "untrained", "a trained network (or a set scale) for mode 2");
    h->ses.learned = m;
    h->ses.learned_mode = mode;
    h->ses.learned_misses = 0;
  });
}

int ddgrs_session_set_state(ddgrs_session* h, const double* in5, ddgrs_error* err) {
  return guarded(err, "session_set_state", [&] {
    need(h, "session");
    need(in5, "state");
    for (int k = 0; k < 5; ++k)
      if (!std::isfinite(in5[k])) invalid("state", std::to_string(in5[k]), "finite values");
    if (!(in5[0] > 0.0)) invalid("mean_edge_length0", std::to_string(in5[0]), "positive");
    if (in5[1] < 0.0) invalid("last_tau", std::to_string(in5[1]), "non-negative");
    h->ses.l0 = in5[0];
    h->ses.last_tau = in5[1];
    for (int k = 0; k < 3; ++k) h->ses.target[k] = in5[2 + k];
  });
}

int ddgrs_session_row_targets(ddgrs_session* h, double* out2, ddgrs_error* err) {
  return guarded(err, "session_row_targets", [&] {
    need(h, "session");
    need(out2, "out");
    out2[0] = out2[1] = NAN;
    size_t j = 0;
    for (int r = 0; r < 2; ++r)
      if (h->ses.opts.rows[r] == kRowArea || h->ses.opts.rows[r] == kRowVolume)
        out2[r] = j < h->ses.nl_targets.size() ? h->ses.nl_targets[j++] : NAN;
  });
}

int ddgrs_session_state(ddgrs_session* h, double* out5, ddgrs_error* err) {
  return guarded(err, "session_state", [&] {
    need(h, "session");
    need(out5, "out");
    out5[0] = h->ses.l0;
    out5[1] = h->ses.last_tau;
    for (int k = 0; k < 3; ++k) out5[2 + k] = h->ses.target[k];
  });
}

}  // extern "C"
