// C ABI of the hierarchical metric B̃, B̃⁰ and of the vertex-pair fractional operator.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "bh.hpp"
#include "blocktree.hpp"
#include "capi_util.hpp"

using namespace ddgrs;

struct ddgrs_metric {
  ddgrs_surface* s = nullptr;
  FaceMetricDev fm;
};

struct ddgrs_vfrac {
  ddgrs_surface* s = nullptr;
  VertexFractionalDev vf;
};

namespace {
void check_tree_options(double chi, int64_t leaf_size, int64_t leaf_pairs) {
  if (!(std::isfinite(chi) && chi >= 0.0)) invalid("chi", std::to_string(chi), "a finite number >= 0 (0: exact products)");
  if (leaf_size < 1) invalid("leaf_size", std::to_string(leaf_size), "at least 1");
  if (leaf_pairs < 0) invalid("leaf_pairs", std::to_string(leaf_pairs), "non-negative");
}

// DDGRS_BLOCKS_ON_HOST=1 builds the block structures and panels with the host planner instead of
// on the device (the two are bitwise equal; the switch exists to check that).
bool blocks_on_host() {
  const char* e = std::getenv("DDGRS_BLOCKS_ON_HOST");
  return e && e[0] == '1';
}

// DDGRS_PRODUCTS_EXTENDED=1 applies the metric and L̃^V with the extended near field (FP64
// differences and sums), which the solves switch to when the FP32 products' rounding limits them.
bool products_extended() {
  const char* e = std::getenv("DDGRS_PRODUCTS_EXTENDED");
  return e && e[0] == '1';
}

// Runs `apply(x_dev, y_dev)` on nb host vectors of length n.
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
}  // namespace

extern "C" {

int ddgrs_metric_build(ddgrs_surface* s, double alpha, double beta, double chi, int64_t leaf_size, int64_t leaf_pairs,
                       ddgrs_metric** out, ddgrs_error* err) {
  return guarded(err, "metric_build", [&] {
    need(s, "surface");
    need(out, "out");
    *out = nullptr;
    Exps e = make_exps(alpha, beta);
    check_tree_options(chi, leaf_size, leaf_pairs);
    auto h = std::make_unique<ddgrs_metric>();
    h->s = s;
    Topology t;
    Clusters c;
    build_face_tree(s, (size_t)leaf_size, &t, &c);
    const double s_exp = (e.beta64 - 2.0) / e.alpha64;
    if (blocks_on_host()) {
      std::vector<double> area((size_t)s->nf);
      download(s->ctx, area.data(), (const double*)s->area.ptr, area.size());
      s->ctx->sync();
      h->fm.build(s, plan_face_metric(s->faces_h, (size_t)s->nv, t, c, chi, (size_t)leaf_pairs, area), s_exp);
    } else {
      h->fm.build(s, t, c, chi, (size_t)leaf_pairs, s_exp);
    }
    surface_retain(s);
    *out = h.release();
  });
}

void ddgrs_metric_destroy(ddgrs_metric* h) { destroy_handle(h); }

int ddgrs_metric_apply_batch(ddgrs_metric* h, double w_hi, double w_lo, int64_t nb, const double* xs, double* ys,
                             ddgrs_error* err) {
  return guarded(err, "metric_apply", [&] {
    need(h, "metric");
    need(xs, "xs");
    need(ys, "ys");
    if (!(std::isfinite(w_hi) && std::isfinite(w_lo))) invalid("metric weights", "non-finite", "finite weights");
    h->fm.w_hi = w_hi;
    h->fm.w_lo = w_lo;
    ddgrs_ctx* ctx = h->s->ctx;
    h->fm.extended = products_extended();
    batch(ctx, nb, 3 * h->s->nv, xs, ys, [&](const double* x, double* y) { h->fm.apply(ctx, x, y); });
    h->fm.extended = false;
  });
}

int ddgrs_metric_stats(ddgrs_metric* h, int64_t* out, ddgrs_error* err) {
  return guarded(err, "metric_stats", [&] {
    need(h, "metric");
    need(out, "out");
    const auto& b = h->fm.stats;
    out[0] = (int64_t)b.far_blocks;
    out[1] = (int64_t)b.near_blocks;
    out[2] = (int64_t)b.admissible_pairs;
    out[3] = (int64_t)b.near_entries;
    out[4] = (int64_t)h->fm.topo.n_nodes();
    out[5] = (int64_t)b.depth;
    out[6] = (int64_t)b.near_total;
  });
}

namespace {
// Copies out a block tree's node arrays and grouped lists (any pointer may be null; counts4 =
// [elements, nodes, admissible blocks, near blocks]).
void export_lists(ddgrs_ctx* ctx, BlockTreeDev& bt, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end,
                  uint32_t* right, uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr, uint32_t* near_col) {
  std::vector<uint32_t> v[8];
  bt.lists(ctx, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
  if (counts4) {
    counts4[0] = bt.n;
    counts4[1] = bt.nn;
    counts4[2] = bt.n_far;
    counts4[3] = bt.n_near;
  }
  uint32_t* outs[8] = {order, start, end, right, far_ptr, far_col, near_ptr, near_col};
  for (int k = 0; k < 8; ++k)
    if (outs[k] && !v[k].empty()) std::memcpy(outs[k], v[k].data(), v[k].size() * sizeof(uint32_t));
}
}  // namespace

int ddgrs_metric_blocks(ddgrs_metric* h, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end,
                        uint32_t* right, uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr, uint32_t* near_col,
                        ddgrs_error* err) {
  return guarded(err, "metric_blocks", [&] {
    need(h, "metric");
    export_lists(h->s->ctx, h->fm.bt, counts4, order, start, end, right, far_ptr, far_col, near_ptr, near_col);
  });
}

int ddgrs_vfrac_build(ddgrs_surface* s, double tau, double chi, int64_t leaf_size, int64_t leaf_pairs,
                      ddgrs_vfrac** out, ddgrs_error* err) {
  return guarded(err, "vfrac_build", [&] {
    need(s, "surface");
    need(out, "out");
    *out = nullptr;
    if (!std::isfinite(tau)) invalid("tau", std::to_string(tau), "a finite order");
    check_tree_options(chi, leaf_size, leaf_pairs);
    auto h = std::make_unique<ddgrs_vfrac>();
    h->s = s;
    std::vector<double> pos, mass;
    vertex_areas_host(s, &pos, &mass);
    GpuBvhBuilder b;
    Topology vt = build_vertex_tree(s, (size_t)leaf_size, b);
    if (blocks_on_host())
      h->vf.build(s->ctx, plan_vertex_fractional(pos, mass, tau, chi, (size_t)leaf_pairs, std::move(vt)));
    else
      h->vf.build(s->ctx, pos, mass, tau, chi, (size_t)leaf_pairs, std::move(vt));
    surface_retain(s);
    *out = h.release();
  });
}

void ddgrs_vfrac_destroy(ddgrs_vfrac* h) { destroy_handle(h); }

int ddgrs_vfrac_blocks(ddgrs_vfrac* h, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end,
                       uint32_t* right, uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr, uint32_t* near_col,
                       ddgrs_error* err) {
  return guarded(err, "vfrac_blocks", [&] {
    need(h, "vfrac");
    export_lists(h->s->ctx, h->vf.bt, counts4, order, start, end, right, far_ptr, far_col, near_ptr, near_col);
  });
}

int ddgrs_vfrac_apply_batch(ddgrs_vfrac* h, int64_t nb, const double* xs, double* ys, ddgrs_error* err) {
  return guarded(err, "vfrac_apply", [&] {
    need(h, "vfrac");
    need(xs, "xs");
    need(ys, "ys");
    ddgrs_ctx* ctx = h->s->ctx;
    h->vf.extended = products_extended();
    batch(ctx, nb, 3 * h->s->nv, xs, ys, [&](const double* x, double* y) { h->vf.apply(ctx, x, y); });
    h->vf.extended = false;
  });
}

}  // extern "C"
