// C ABI of the continuous collision detection (crates/ddg-repulsive-surfaces/src/ccd.rs).
#include <chrono>
#include <cmath>
#include <cstring>

#include "capi_util.hpp"
#include "ccd.hpp"

using namespace ddgrs;

namespace {

CcdOpts to_opts(const ddgrs_ccd_options& o) {
  CcdOpts c;
  c.gap = o.gap;
  c.reference_length = o.reference_length;
  c.close_factor = o.close_factor;
  if (o.max_iterations < 1 || o.max_iterations > UINT32_MAX)
    invalid("ccd.max_iterations", std::to_string(o.max_iterations), "at least 1");
  if (o.leaf_size < 1 || o.leaf_size > (1 << 20)) invalid("ccd.leaf_size", std::to_string(o.leaf_size), "at least 1");
  c.max_iterations = (uint32_t)o.max_iterations;
  c.leaf_size = (uint32_t)o.leaf_size;
  check_ccd_options(c);
  return c;
}

// Surface::validate_positions (RS/src/surface.rs:513) on host positions.
void validate_positions(ddgrs_surface* s, const double* p) {
  for (int64_t v = 0; v < s->nv; ++v)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(p[3 * v + k])) {
        Error e(DDGRS_NON_FINITE_POSITION, "validate_positions", "vertex " + std::to_string(v) + " has a non-finite position");
        e.a = v;
        throw e;
      }
  for (int64_t f = 0; f < s->nf; ++f) {
    const double* a = p + 3 * (size_t)s->faces_h[3 * f];
    const double* b = p + 3 * (size_t)s->faces_h[3 * f + 1];
    const double* c = p + 3 * (size_t)s->faces_h[3 * f + 2];
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double area = 0.5 * len, inv = 1.0 / len;
    const bool ok = area > 0.0 && std::isfinite(area) && std::isfinite(n[0] * inv) && std::isfinite(n[1] * inv) &&
                    std::isfinite(n[2] * inv);
    if (!ok) {
      Error e(DDGRS_DEGENERATE_FACE, "validate_positions", "face " + std::to_string(f) + " is degenerate");
      e.a = f;
      e.value = area;
      throw e;
    }
  }
}

// The positions to use on the host and the device: the caller's (validated, copied up) or the
// surface's own.
struct Positions {
  std::vector<double> host;
  DevBuf<double> dev;
  const double* d = nullptr;
  const double* h = nullptr;
};

void positions_of(ddgrs_surface* s, const double* given, Positions* P) {
  const size_t n = 3 * (size_t)s->nv;
  if (given) {
    validate_positions(s, given);
    upload(s->ctx, P->dev, given, n);
    P->d = P->dev.ptr;
    P->h = given;
  } else {
    P->host.resize(n);
    download(s->ctx, P->host.data(), (const double*)s->positions.ptr, n);
    s->ctx->sync();
    P->d = s->positions.ptr;
    P->h = P->host.data();
  }
}

}  // namespace

extern "C" {

void ddgrs_ccd_options_default(ddgrs_ccd_options* o) {
  if (!o) return;
  o->gap = 0.05;
  o->reference_length = NAN;
  o->max_iterations = 100;
  o->close_factor = 0.5;
  o->leaf_size = 8;
}

int ddgrs_ccd_first_contact(ddgrs_surface* s, const double* start, const double* velocity, double tau_end,
                            const ddgrs_ccd_options* o, const uint8_t* band, ddgrs_ccd_report* report,
                            ddgrs_error* err) {
  return guarded(err, "ccd_first_contact", [&] {
    need(s, "surface");
    need(velocity, "velocity");
    need(o, "options");
    need(report, "report");
    const auto t0 = std::chrono::steady_clock::now();
    const CcdOpts c = to_opts(*o);
    const size_t n = 3 * (size_t)s->nv;
    for (size_t i = 0; i < n; ++i)
      if (!std::isfinite(velocity[i])) invalid("velocity", "non-finite", "finite components");
    if (!(std::isfinite(tau_end) && tau_end >= 0.0)) invalid("tau_end", std::to_string(tau_end), "finite and non-negative");
    ddgrs_ctx* ctx = s->ctx;
    Positions P;
    positions_of(s, start, &P);
    CcdDev d;
    d.set_mesh(ctx, s->faces_h, (uint32_t)s->nv);
    DevBuf<double> vel;
    upload(ctx, vel, velocity, n);
    DevBuf<uint8_t> bd;
    bool any = false;
    if (band) {
      for (int64_t v = 0; v < s->nv && !any; ++v) any = band[v] != 0;
      if (any) upload(ctx, bd, band, (size_t)s->nv);
    }
    const CcdResult r = d.first_contact(ctx, s->faces.ptr, P.d, P.h, vel.ptr, 1.0, tau_end, c, any ? bd.ptr : nullptr);
    std::memset(report, 0, sizeof *report);
    report->tau_max = r.tau_max;
    report->has_contact = r.has_contact ? 1 : 0;
    report->contact_kind = r.kind;
    report->contact_a = r.a;
    report->contact_b = r.b;
    report->gap_abs = r.gap_abs;
    report->iterations = (int64_t)r.iterations;
    report->pairs_tested = (int64_t)r.pairs_tested;
    report->close_pairs = (int64_t)r.close_pairs;
    report->touching_pairs = (int64_t)r.touching_pairs;
    report->exempt_pairs = (int64_t)r.exempt_pairs;
    report->windows = (int32_t)r.windows;
    report->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int ddgrs_ccd_intersecting_pairs(ddgrs_surface* s, const double* positions, uint32_t* pairs, int64_t cap,
                                 int64_t* count, ddgrs_error* err) {
  return guarded(err, "ccd_intersecting_pairs", [&] {
    need(s, "surface");
    need(count, "count");
    if (cap < 0) invalid("cap", std::to_string(cap), "non-negative");
    if (cap > 0) need(pairs, "pairs");
    ddgrs_ctx* ctx = s->ctx;
    Positions P;
    positions_of(s, positions, &P);
    CcdDev d;
    d.set_mesh(ctx, s->faces_h, (uint32_t)s->nv);
    const size_t np = d.find_intersecting_pairs(ctx, s->faces.ptr, P.d);
    std::vector<uint64_t> keys(np);
    if (np) {
      std::vector<uint32_t> raw(2 * np);
      download(ctx, raw.data(), (const uint32_t*)d.pairs.ptr, 2 * np);
      ctx->sync();
      for (size_t i = 0; i < np; ++i) keys[i] = (uint64_t)raw[2 * i] << 32 | raw[2 * i + 1];
      std::sort(keys.begin(), keys.end());
    }
    for (size_t i = 0; i < np && (int64_t)i < cap; ++i) {
      pairs[2 * i] = (uint32_t)(keys[i] >> 32);
      pairs[2 * i + 1] = (uint32_t)keys[i];
    }
    *count = (int64_t)np;
  });
}

int ddgrs_ccd_band(ddgrs_surface* s, const double* positions, uint8_t* flags, int64_t* counts2, ddgrs_error* err) {
  return guarded(err, "ccd_band", [&] {
    need(s, "surface");
    need(flags, "flags");
    need(counts2, "counts");
    ddgrs_ctx* ctx = s->ctx;
    Positions P;
    positions_of(s, positions, &P);
    CcdDev d;
    d.set_mesh(ctx, s->faces_h, (uint32_t)s->nv);
    const size_t np = d.band_flags(ctx, s->faces.ptr, P.d);
    download(ctx, flags, (const uint8_t*)d.band.ptr, (size_t)s->nv);
    ctx->sync();
    int64_t nb = 0;
    for (int64_t v = 0; v < s->nv; ++v) nb += flags[v] != 0;
    counts2[0] = (int64_t)np;
    counts2[1] = nb;
  });
}

int ddgrs_surface_edges(ddgrs_surface* s, uint32_t* edges, int64_t cap, int64_t* ne, ddgrs_error* err) {
  return guarded(err, "surface_edges", [&] {
    need(s, "surface");
    need(ne, "ne");
    CcdDev d;
    d.set_mesh(s->ctx, s->faces_h, (uint32_t)s->nv);
    *ne = (int64_t)d.ne;
    if (edges) std::memcpy(edges, d.edges_h.data(), sizeof(uint32_t) * 2 * std::min<size_t>(d.ne, (size_t)std::max<int64_t>(cap, 0)));
  });
}

}  // extern "C"
