// C ABI of the Barnes–Hut evaluator (include/ddgrs.h): build, refit, energy, gradient, and the
// tree / list export used to check the structures against the reference's build.
#include <cmath>

#include "bh.hpp"
#include "capi_util.hpp"

using namespace ddgrs;

struct ddgrs_bh {
  ddgrs_surface* s = nullptr;
  BhEval ev;
};

extern "C" {

int ddgrs_bh_build(ddgrs_surface* s, double theta, int64_t leaf_size, ddgrs_bh** out, ddgrs_error* err) {
  return guarded(err, "bh_build", [&] {
    need(s, "surface");
    need(out, "out");
    *out = nullptr;
    if (!(std::isfinite(theta) && theta >= 0.0)) invalid("theta", std::to_string(theta), "finite and non-negative");
    if (leaf_size < 1) invalid("leaf_size", std::to_string(leaf_size), "at least 1");
    auto h = std::make_unique<ddgrs_bh>();
    h->s = s;
    h->ev.build(s, theta, (size_t)leaf_size);
    surface_retain(s);
    *out = h.release();
  });
}

void ddgrs_bh_destroy(ddgrs_bh* h) { destroy_handle(h); }

int ddgrs_bh_refit(ddgrs_bh* h, const double* positions, ddgrs_error* err) {
  return guarded(err, "bh_refit", [&] {
    need(h, "bh");
    ddgrs_surface* s = h->s;
    if (positions) upload(s->ctx, s->positions, positions, (size_t)(3 * s->nv));
    compute_face_geometry(s, s->packs, false);
    h->ev.refresh(s);
  });
}

int ddgrs_bh_energy(ddgrs_bh* h, double alpha, double beta, double* energy, ddgrs_error* err) {
  return guarded(err, "bh_energy", [&] {
    need(h, "bh");
    need(energy, "energy");
    Exps e = make_exps(alpha, beta);
    ddgrs_ctx* ctx = h->s->ctx;
    double* d = ctx->scratch.f64(1);
    h->ev.energy(ctx, e, d);
    double* pin = (double*)ctx->staging(8);
    download(ctx, pin, (const double*)d, 1);
    ctx->sync();
    *energy = pin[0];
  });
}

int ddgrs_bh_energy_and_differential(ddgrs_bh* h, double alpha, double beta, double* energy, double* differential,
                                     double* face_partials, ddgrs_error* err) {
  return guarded(err, "bh_energy_and_differential", [&] {
    need(h, "bh");
    need(energy, "energy");
    Exps e = make_exps(alpha, beta);
    ddgrs_surface* s = h->s;
    ddgrs_ctx* ctx = s->ctx;
    const size_t nf = (size_t)s->nf, nv = (size_t)s->nv;
    double* buf = ctx->scratch.f64(7 * nf + 3 * nv + 1);
    double *partials = buf, *diff = buf + 7 * nf, *total = diff + 3 * nv;
    h->ev.gradient(s, e, total, partials);
    if (differential) pull_back(s, partials, diff);
    double* pin = (double*)ctx->staging(8);
    download(ctx, pin, (const double*)total, 1);
    if (differential) download(ctx, differential, (const double*)diff, 3 * nv);
    if (face_partials) download(ctx, face_partials, (const double*)partials, 7 * nf);
    ctx->sync();
    *energy = pin[0];
    bool finite = std::isfinite(*energy);
    if (finite && differential)
      for (size_t i = 0; i < 3 * nv && finite; ++i) finite = std::isfinite(differential[i]);
    if (!finite) throw Error(DDGRS_NON_FINITE_ENERGY, "bh_energy_and_differential", "non-finite Barnes-Hut energy or differential");
  });
}

int ddgrs_bh_counts(ddgrs_bh* h, int64_t* out, ddgrs_error* err) {
  return guarded(err, "bh_counts", [&] {
    need(h, "bh");
    need(out, "out");
    const auto& ev = h->ev;
    out[0] = (int64_t)ev.topo.n_nodes();
    out[1] = (int64_t)ev.topo.leaves.size();
    out[2] = (int64_t)ev.n_lumped;
    out[3] = (int64_t)ev.n_exact;
    out[4] = (int64_t)ev.pairs_exact;
    const BhLists L = ev.download_lists(h->s->ctx);
    int64_t slots = 0;
    for (uint32_t a : L.exact) slots += ev.topo.end[a] - ev.topo.start[a];
    out[5] = slots;
  });
}

int ddgrs_bh_tree(ddgrs_bh* h, uint32_t* order, uint32_t* start, uint32_t* end, uint32_t* right, double* clusters,
                  ddgrs_error* err) {
  return guarded(err, "bh_tree", [&] {
    need(h, "bh");
    const auto& t = h->ev.topo;
    const auto& c = h->ev.clusters;
    const size_t nn = t.n_nodes();
    if (order) std::copy(t.order.begin(), t.order.end(), order);
    if (start) std::copy(t.start.begin(), t.start.end(), start);
    if (end) std::copy(t.end.begin(), t.end.end(), end);
    if (right) std::copy(t.right.begin(), t.right.end(), right);
    if (clusters)
      for (size_t a = 0; a < nn; ++a) {
        double* o = clusters + 17 * a;
        o[0] = c.mass[a];
        for (int k = 0; k < 3; ++k) o[1 + k] = c.centroid[3 * a + k];
        for (int k = 0; k < 6; ++k) o[4 + k] = c.projector.empty() ? 0.0 : c.projector[6 * a + k];
        for (int k = 0; k < 3; ++k) o[10 + k] = c.lo[3 * a + k];
        for (int k = 0; k < 3; ++k) o[13 + k] = c.hi[3 * a + k];
        o[16] = c.radius[a];
      }
  });
}

int ddgrs_bh_lists(ddgrs_bh* h, uint32_t* lumped_off, uint32_t* lumped, uint32_t* exact_off, uint32_t* exact,
                   ddgrs_error* err) {
  return guarded(err, "bh_lists", [&] {
    need(h, "bh");
    const BhLists L = h->ev.download_lists(h->s->ctx);
    if (lumped_off) std::copy(L.lumped_off.begin(), L.lumped_off.end(), lumped_off);
    if (lumped) std::copy(L.lumped.begin(), L.lumped.end(), lumped);
    if (exact_off) std::copy(L.exact_off.begin(), L.exact_off.end(), exact_off);
    if (exact) std::copy(L.exact.begin(), L.exact.end(), exact);
  });
}

}  // extern "C"
