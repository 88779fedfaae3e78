// C ABI of the remeshing pass (crates/ddg-repulsive-surfaces/src/remesh/).
#include <cmath>

#include "capi_util.hpp"
#include "remesh.hpp"

using namespace ddgrs;

namespace {

RemeshInput make_input(int64_t nv, int64_t nf, const uint32_t* faces, const double* positions, int64_t npinned,
                       const uint32_t* pinned) {
  need(faces, "faces");
  need(positions, "positions");
  if (npinned < 0) invalid("npinned", std::to_string(npinned), "non-negative");
  if (npinned > 0) need(pinned, "pinned");
  RemeshInput in;
  in.nv = nv;
  in.nf = nf;
  in.faces = faces;
  in.positions = positions;
  in.pinned.assign(pinned, pinned + npinned);
  return in;
}

}  // namespace

extern "C" {

// RemeshOptions::default (RS/src/remesh/mod.rs:131); no collapse guard.
DDGRS_API void ddgrs_remesh_options_default(ddgrs_remesh_options* o) {
  if (!o) return;
  o->split = 1.5;
  o->collapse = 0.5;
  o->rounds = 5;
  o->rho = 0.5;
  o->flip_cap_per_vertex = 100;
  o->order = 0;
  o->split_on = o->collapse_on = o->flip_on = o->smooth_on = 1;
  o->max_growth = 2.0;
  o->guard_gap = 0.0;
}

DDGRS_API int ddgrs_remesh(ddgrs_ctx* ctx, int64_t nv, int64_t nf, const uint32_t* faces, const double* positions,
                           int64_t npinned, const uint32_t* pinned, double l0, const ddgrs_remesh_options* o,
                           const uint8_t* band, ddgrs_remeshed** out, ddgrs_remesh_report* report, ddgrs_error* err) {
  return guarded(err, "ddgrs_remesh", [&] {
    need(ctx, "ctx");
    need(o, "options");
    need(out, "out");
    *out = nullptr;
    RemeshInput in = make_input(nv, nf, faces, positions, npinned, pinned);
    in.band = band;
    HIPCHECK(hipSetDevice(ctx->device));
    auto r = std::make_unique<ddgrs_remeshed>();
    ddgrs_remesh_report rep{};
    remesh_pass(ctx, in, l0, *o, r.get(), &rep);
    if (report) *report = rep;
    *out = r.release();
  });
}

DDGRS_API int ddgrs_remesh_op(ddgrs_ctx* ctx, int64_t nv, int64_t nf, const uint32_t* faces, const double* positions,
                              int64_t npinned, const uint32_t* pinned, int32_t op, uint32_t a, uint32_t b,
                              int32_t delaunay, int32_t foldover, int32_t* refusal, int32_t* reordered,
                              ddgrs_remeshed** out, ddgrs_error* err) {
  return guarded(err, "ddgrs_remesh_op", [&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = nullptr;
    const RemeshInput in = make_input(nv, nf, faces, positions, npinned, pinned);
    HIPCHECK(hipSetDevice(ctx->device));
    auto r = std::make_unique<ddgrs_remeshed>();
    bool reo = false;
    const int k = remesh_single_op(ctx, in, op, a, b, delaunay != 0, foldover != 0, &reo, r.get());
    if (refusal) *refusal = k;
    if (reordered) *reordered = reo ? 1 : 0;
    *out = r.release();
  });
}

DDGRS_API int ddgrs_remeshed_sizes(const ddgrs_remeshed* r, int64_t* sizes4, ddgrs_error* err) {
  return guarded(err, "ddgrs_remeshed_sizes", [&] {
    need(r, "remeshed");
    need(sizes4, "sizes4");
    sizes4[0] = (int64_t)r->positions.size() / 3;
    sizes4[1] = (int64_t)r->faces.size() / 3;
    sizes4[2] = (int64_t)r->map_sources.size();
    sizes4[3] = (int64_t)r->pinned.size();
  });
}

DDGRS_API int ddgrs_remeshed_read(const ddgrs_remeshed* r, double* positions, uint32_t* faces, uint32_t* face_parents,
                                  uint32_t* map_offsets, uint32_t* map_sources, double* map_weights, uint32_t* pinned,
                                  ddgrs_error* err) {
  return guarded(err, "ddgrs_remeshed_read", [&] {
    need(r, "remeshed");
    auto copy = [](auto* dst, const auto& v) {
      if (dst) std::copy(v.begin(), v.end(), dst);
    };
    copy(positions, r->positions);
    copy(faces, r->faces);
    copy(face_parents, r->parents);
    copy(map_offsets, r->map_offsets);
    copy(map_sources, r->map_sources);
    copy(map_weights, r->map_weights);
    copy(pinned, r->pinned);
  });
}

DDGRS_API void ddgrs_remeshed_destroy(ddgrs_remeshed* r) { delete r; }

}  // extern "C"
