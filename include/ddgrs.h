/*
 * ddgrs — MI355X (gfx950) implementation of the Hs-descent hot path of ddg-repulsive-surfaces.
 *
 * C ABI only: opaque handles, plain pointers and sizes, no C++ or torch types.  Every entry point
 * returns a status (0 = ok) and fills `ddgrs_error` on failure; nothing aborts.  All arrays at the
 * boundary are the reference's: faces are `3·nf` u32 (caller's vertex order), positions and
 * differentials are `3·nv` f64 interleaved `[3 v + axis]` (crates/ddg-repulsive-surfaces/src/metric/
 * hierarchical.rs:39-42), saddle vectors are `3·nv + k_Q` f64.  Input slices are borrowed for the
 * duration of the call; outputs are caller-allocated and overwritten (solve/gmres.rs:36-40).  Device
 * copies belong to the handle until it is destroyed.  One context is one in-order HIP stream; calls on
 * a context are serialized by the caller (gpu/cuda.rs:111-114).
 *
 * Each declaration names the reference interface it replaces (path relative to the reference root,
 * `RS/` = crates/ddg-repulsive-surfaces/).
 */
#ifndef DDGRS_H
#define DDGRS_H

#include <stdint.h>

#if defined(__GNUC__)
#define DDGRS_API __attribute__((visibility("default")))
#else
#define DDGRS_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the kinds mirror RsError (RS/src/error.rs). */
enum ddgrs_status {
  DDGRS_OK = 0,
  DDGRS_INVALID_OPTION = 1,     /* RsError::InvalidOption{name, value, requirement} */
  DDGRS_INVALID_EXPONENTS = 2,  /* RsError::InvalidExponents */
  DDGRS_NON_FINITE_ENERGY = 3,  /* RsError::NonFiniteEnergy{faces} */
  DDGRS_SOLVER_DIVERGED = 4,    /* RsError::SolverDiverged{residual} */
  DDGRS_SINGULAR_SADDLE = 5,    /* RsError::SingularSaddle{pivot} */
  DDGRS_DEGENERATE_FACE = 6,    /* RsError::DegenerateFace{face, area} */
  DDGRS_NON_FINITE_POSITION = 7,/* RsError::NonFinitePosition{vertex} */
  DDGRS_GPU = 8                 /* RsError::Gpu{operation, detail} (RS/src/gpu/cuda.rs:19-24) */
};

typedef struct ddgrs_error {
  int32_t status;
  char operation[64];
  char detail[256];
  int64_t a, b; /* e.g. the face pair of NonFiniteEnergy, the face of DegenerateFace */
  double value; /* e.g. the residual of SolverDiverged */
} ddgrs_error;

typedef struct ddgrs_ctx ddgrs_ctx;
typedef struct ddgrs_surface ddgrs_surface;
typedef struct ddgrs_bh ddgrs_bh;
typedef struct ddgrs_metric ddgrs_metric;
typedef struct ddgrs_vfrac ddgrs_vfrac;

/* ---- context ------------------------------------------------------------------------------- */

/* Replaces GpuContext::new / CudaBackend::new (RS/src/gpu/mod.rs:151, RS/src/gpu/cuda.rs:106),
 * selected by Device::Cuda(CudaOptions{ordinal, determinism}) (RS/src/gpu/mod.rs:47-87).
 * determinism: 1 = run-to-run bitwise reproducible (the only level; results never depend on the
 * launch shape). */
DDGRS_API int ddgrs_ctx_create(int32_t ordinal, int32_t determinism, ddgrs_ctx** out, ddgrs_error* err);
DDGRS_API void ddgrs_ctx_destroy(ddgrs_ctx* ctx);
/* Replaces gpu_info() (crates/ddg-py/src/python/repulsive_surfaces.rs:683). */
DDGRS_API int ddgrs_device_info(ddgrs_ctx* ctx, char* name, int64_t name_len, int64_t* compute_units,
                      int64_t* memory_bytes, ddgrs_error* err);
/* Blocks until the context's stream is idle. */
DDGRS_API int ddgrs_ctx_synchronize(ddgrs_ctx* ctx, ddgrs_error* err);
/* Kernel timing: when enabled, every launch is bracketed by HIP events on the context's stream;
 * ddgrs_ctx_kernel_times reports per kernel name the launch count and the summed milliseconds. */
DDGRS_API int ddgrs_ctx_set_profiling(ddgrs_ctx* ctx, int32_t enabled, ddgrs_error* err);
DDGRS_API int ddgrs_ctx_kernel_times(ddgrs_ctx* ctx, int64_t capacity, char* names /* capacity*64 */,
                           int64_t* launches, double* milliseconds, int64_t* count, ddgrs_error* err);

/* ---- surface (device-resident mesh and positions) ------------------------------------------ */

/* The validated triangle mesh and its positions on the device (RS/src/surface.rs:132 Surface::new,
 * :513 validate_positions). */
DDGRS_API int ddgrs_surface_create(ddgrs_ctx* ctx, int64_t nv, int64_t nf, const uint32_t* faces,
                         const double* positions, ddgrs_surface** out, ddgrs_error* err);
DDGRS_API void ddgrs_surface_destroy(ddgrs_surface* s);
/* Moves the surface (validated as RS/src/surface.rs:513). */
DDGRS_API int ddgrs_surface_set_positions(ddgrs_surface* s, const double* positions, ddgrs_error* err);
DDGRS_API int ddgrs_surface_get_positions(ddgrs_surface* s, double* positions, ddgrs_error* err);

/* ---- energy --------------------------------------------------------------------------------- */

/* The exact all-pairs tangent-point energy and its differential `∂E/∂f_v` (3·nv, optional),
 * FP32 pair kernel with compensated accumulation.  Replaces energy() with EnergyEval::Exact
 * (RS/src/energy/mod.rs:168 → RS/src/energy/exact.rs:146 energy_and_differential).
 * face_partials (optional, 7·nf): (∂/∂a_S, ∂/∂X_S, ∂/∂N_S) as exact.rs:87 face_partials. */
DDGRS_API int ddgrs_exact_energy_and_differential(ddgrs_surface* s, double alpha, double beta, double* energy,
                                        double* differential, double* face_partials, ddgrs_error* err);

/* ---- Barnes–Hut --------------------------------------------------------------------------- */

/* Builds the face BVH (leaf size, split rule of RC/src/spatial/bvh.rs:126) and the frozen
 * interaction lists for θ at the surface's current positions.  Replaces BhTree::build
 * (RS/src/energy/barnes_hut.rs:156) + attach_gpu (:265). */
DDGRS_API int ddgrs_bh_build(ddgrs_surface* s, double theta, int64_t leaf_size, ddgrs_bh** out, ddgrs_error* err);
DDGRS_API void ddgrs_bh_destroy(ddgrs_bh* bh);
/* Moves to new positions (3·nv, or NULL for the surface's current ones) keeping the tree, boxes,
 * radii and lists; refreshes the face data and A_C, X_C.  Replaces BhTree::refit (:204). */
DDGRS_API int ddgrs_bh_refit(ddgrs_bh* bh, const double* positions, ddgrs_error* err);
/* Ẽ on the frozen lists (value only, line-search trials).  Replaces BhTree::try_energy (:350). */
DDGRS_API int ddgrs_bh_energy(ddgrs_bh* bh, double alpha, double beta, double* energy, ddgrs_error* err);
/* Ẽ and its exact differential for the frozen lists (3·nv, optional) and the per-face partials
 * (7·nf, optional).  Replaces BhTree::energy_and_differential (:387) / try_face_partials (:421). */
DDGRS_API int ddgrs_bh_energy_and_differential(ddgrs_bh* bh, double alpha, double beta, double* energy,
                                               double* differential, double* face_partials, ddgrs_error* err);
/* [nodes, leaves, lumped entries, exact blocks, pairs_exact, partner slots]
 * (InteractionLists::n_lumped / pairs_exact, :94-107). */
DDGRS_API int ddgrs_bh_counts(ddgrs_bh* bh, int64_t* out6, ddgrs_error* err);
/* The tree (any pointer may be NULL): order (nf), per node start/end/right, and per node 17 f64
 * [A, X(3), P̄(6: xx xy xz yy yz zz), box lo(3), box hi(3), r] (RS/src/spatial/mod.rs:80 Cluster). */
DDGRS_API int ddgrs_bh_tree(ddgrs_bh* bh, uint32_t* order, uint32_t* start, uint32_t* end, uint32_t* right,
                            double* clusters, ddgrs_error* err);
/* The lists: per leaf ordinal CSR offsets (leaves+1) and node indices (InteractionLists :61). */
DDGRS_API int ddgrs_bh_lists(ddgrs_bh* bh, uint32_t* lumped_off, uint32_t* lumped, uint32_t* exact_off,
                             uint32_t* exact, ddgrs_error* err);

/* ---- hierarchical metric ------------------------------------------------------------------- */

/* The block cluster tree of B, B⁰ on a face BVH at the surface's current positions: separation χ,
 * leaf size, leaf_pairs (RS/src/metric/hierarchical.rs:65 ClusterTreeOptions), near field and
 * H̃a on the device.  Replaces FaceMetricTree::build / build_on_gpu (hierarchical.rs:296,374). */
DDGRS_API int ddgrs_metric_build(ddgrs_surface* s, double alpha, double beta, double chi, int64_t leaf_size,
                                 int64_t leaf_pairs, ddgrs_metric** out, ddgrs_error* err);
DDGRS_API void ddgrs_metric_destroy(ddgrs_metric* m);
/* ys_k = (w_hi B̃ + w_lo B̃⁰) xs_k for nb interleaved 3·nv vectors.  Replaces
 * FaceMetricTree::apply_batch (hierarchical.rs:540) / GpuFaceMetric::apply_batch (gpu/mod.rs:318). */
DDGRS_API int ddgrs_metric_apply_batch(ddgrs_metric* m, double w_hi, double w_lo, int64_t nb, const double* xs,
                                       double* ys, ddgrs_error* err);
/* [admissible blocks, inadmissible blocks, admissible pairs, near entries, nodes, depth, stored
 * near entries] (BctStats, RS/src/spatial/block_tree.rs:401). */
DDGRS_API int ddgrs_metric_stats(ddgrs_metric* m, int64_t* out7, ddgrs_error* err);

/* The vertex-pair L̃^V_τ (dual areas of the surface) on a vertex BVH.  Replaces
 * VertexFractionalTree::for_surface / build_on_gpu (hierarchical.rs:1157,1092). */
DDGRS_API int ddgrs_vfrac_build(ddgrs_surface* s, double tau, double chi, int64_t leaf_size, int64_t leaf_pairs,
                                ddgrs_vfrac** out, ddgrs_error* err);
DDGRS_API void ddgrs_vfrac_destroy(ddgrs_vfrac* v);
/* Replaces VertexFractionalTree::apply_batch (hierarchical.rs:1274). */
DDGRS_API int ddgrs_vfrac_apply_batch(ddgrs_vfrac* v, int64_t nb, const double* xs, double* ys, ddgrs_error* err);

#ifdef __cplusplus
}
#endif
#endif /* DDGRS_H */
