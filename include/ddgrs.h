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
 * Lifetimes: a context lives until it and every surface on it are destroyed, a surface until it and
 * every handle built on it (Barnes–Hut, metric, saddle, session) are destroyed; the destroy calls may
 * come in any order.
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
typedef struct ddgrs_saddle ddgrs_saddle;
typedef struct ddgrs_session ddgrs_session;
typedef struct ddgrs_learned ddgrs_learned;

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

/* The same as ddgrs_exact_energy_and_differential (same kernels, same results bit for bit), but
 * allocating nothing: no counterpart in the reference; for callers that cannot allocate (tight
 * serving loops, code that must not call hipMalloc).  The kernels' scratch memory is one 52 MiB
 * array per device, resident until the process ends, so the mesh is limited to 131,072 faces and
 * 131,072 vertices (larger: DDGRS_INVALID_OPTION, before any work).  Calls that overlap on one
 * device share that array and race; serialize them.  Outputs are host memory as in the allocating
 * entry point (the copies are made straight into them). */
DDGRS_API int ddgrs_exact_energy_and_differential_noalloc(ddgrs_surface* s, double alpha, double beta, double* energy,
                                                          double* differential, double* face_partials,
                                                          ddgrs_error* err);

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

/* The block cluster tree and its block lists, for the bit-exact check against the CPU build
 * (RS/src/spatial/block_tree.rs:129 build_impl: blocks level by level, then grouped by row node,
 * creation order kept, :198 group_by_row).  Any array pointer may be NULL (call with all NULL to
 * read the sizes first): counts4 = [elements, nodes, admissible blocks, inadmissible blocks];
 * order (elements: BVH position → element), start, end, right (nodes; pre-order numbering),
 * far_ptr and near_ptr (nodes + 1: CSR offsets per row node), far_col (admissible blocks) and
 * near_col (inadmissible blocks: dense near field): the column nodes. */
DDGRS_API int ddgrs_metric_blocks(ddgrs_metric* m, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end,
                                  uint32_t* right, uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr,
                                  uint32_t* near_col, ddgrs_error* err);

/* The vertex-pair L̃^V_τ (dual areas of the surface) on a vertex BVH.  Replaces
 * VertexFractionalTree::for_surface / build_on_gpu (hierarchical.rs:1157,1092). */
DDGRS_API int ddgrs_vfrac_build(ddgrs_surface* s, double tau, double chi, int64_t leaf_size, int64_t leaf_pairs,
                                ddgrs_vfrac** out, ddgrs_error* err);
DDGRS_API void ddgrs_vfrac_destroy(ddgrs_vfrac* v);
/* Replaces VertexFractionalTree::apply_batch (hierarchical.rs:1274). */
DDGRS_API int ddgrs_vfrac_apply_batch(ddgrs_vfrac* v, int64_t nb, const double* xs, double* ys, ddgrs_error* err);
/* The vertex tree and its block lists, as ddgrs_metric_blocks (the vertex BVH and the blocks of
 * VertexFractionalTree::build, hierarchical.rs:1017). */
DDGRS_API int ddgrs_vfrac_blocks(ddgrs_vfrac* v, int64_t* counts4, uint32_t* order, uint32_t* start, uint32_t* end,
                                 uint32_t* right, uint32_t* far_ptr, uint32_t* far_col, uint32_t* near_ptr,
                                 uint32_t* near_col, ddgrs_error* err);

/* ---- the step's saddle solver --------------------------------------------------------------- */

/* Options of the hierarchical saddle solver (RS/src/solve/iterative.rs:350 HierarchicalOptions with
 * RS/src/solve/precond.rs:466 PreconditionerOptions; the metric weights of SobolevWeights). */
typedef struct ddgrs_saddle_options {
  double chi;         /* block cluster tree separation (0: exact products) */
  int64_t leaf_size;  /* face and vertex BVH leaf size */
  int64_t leaf_pairs; /* blocks with at most this many pairs are dense */
  double w_tp;        /* weight of B̃ */
  double c_lo;        /* weight of B̃⁰ relative to w_tp */
  double eps_rel;     /* K = W + ε M with ε = eps_rel · tr W / tr M */
  /* PreconditionerOptions::form: 0 Augmented, P⁻¹ = L′⁻¹[(s L^V − K) y; 0] + y with y = L′⁻¹ r;
   * 1 Congruence, P⁻¹ = L′⁻¹[L^V y; 0] (RS/src/solve/precond.rs:1026 apply_fine). */
  int32_t form;
  /* PreconditionerOptions::scale: 0 Symbol, s = 1/(w_tp 4 I(σ) I(2−s)); 1 Unit, s = 1
   * (precond.rs:502 FractionalScale).  No effect on the congruence form. */
  int32_t scale;
  /* PreconditionerOptions::refine: 1 adds one step of iterative refinement to every L′ solve. */
  int32_t refine;
} ddgrs_saddle_options;

/* The GMRES report (RS/src/solve/gmres.rs:200 GmresReport). */
typedef struct ddgrs_gmres_report {
  int64_t iterations, cycles, operator_applications, preconditioner_applications;
  int32_t converged;
  /* 1 when the solve switched its products to the extended near field (FP64 differences and sums)
   * because the FP32 products' rounding kept the explicit residual above the tolerance; no
   * counterpart in the reference, whose products are FP64 throughout. */
  int32_t extended;
  double relative_residual; /* explicit ‖b − A x‖ / ‖b‖ at the last cycle start */
} ddgrs_gmres_report;

/* Fills the reference defaults (χ 0.5, leaf 8, 64 pairs, w 1, c_lo 1, ε 1e-10, augmented form,
 * symbol scale, no refinement). */
DDGRS_API void ddgrs_saddle_options_default(ddgrs_saddle_options* o);
/* The solver of A′ = [[w (B̃ + c_lo B̃⁰), C_Qᵀ], [C_Q, 0]] with the global barycenter rows C_Q
 * (k_Q = 3) at the surface's current positions: the metric tree, L′ (sparse LDLᵀ in a nested
 * dissection order, solved on the device) and L̃^V_{2−s}.  Replaces hierarchical_saddle
 * (RS/src/solve/iterative.rs:455) with SparsePreconditioner::new (RS/src/solve/precond.rs:686). */
DDGRS_API int ddgrs_saddle_build(ddgrs_surface* s, double alpha, double beta, const ddgrs_saddle_options* o,
                                 ddgrs_saddle** out, ddgrs_error* err);
DDGRS_API void ddgrs_saddle_destroy(ddgrs_saddle* h);
/* [ε, fractional scale s_frac] (precond.rs:931 fractional_scale). */
DDGRS_API int ddgrs_saddle_info(ddgrs_saddle* h, double* out2, ddgrs_error* err);
/* y = A′ x for nb vectors of 3·nv + 3 (RS/src/solve/iterative.rs:284 apply_batch). */
DDGRS_API int ddgrs_saddle_apply_batch(ddgrs_saddle* h, int64_t nb, const double* xs, double* ys, ddgrs_error* err);
/* z = P⁻¹ r, the augmented preconditioner (RS/src/solve/precond.rs:1099 apply_batch → :1026). */
DDGRS_API int ddgrs_saddle_precond_batch(ddgrs_saddle* h, int64_t nb, const double* rs, double* zs, ddgrs_error* err);
/* x = L′⁻¹ r (RS/src/solve/precond.rs:384 LaplacianSaddle::solve). */
DDGRS_API int ddgrs_saddle_lprime_batch(ddgrs_saddle* h, int64_t nb, const double* rs, double* xs, ddgrs_error* err);
/* x = A′⁻¹ b by right-preconditioned restarted GMRES on the device (RS/src/solve/gmres.rs:783
 * GmresSaddle::solve, options gmres.rs:138). */
DDGRS_API int ddgrs_saddle_solve(ddgrs_saddle* h, const double* b, double* x, int64_t restart, double tolerance,
                                 int64_t max_iterations, ddgrs_gmres_report* report, ddgrs_error* err);
/* x_i = A′⁻¹ b_i for nb right-hand sides of 3·nv + 3 each (bs, xs row-major nb × (3·nv + 3)), one
 * report per right-hand side (reports may be NULL).  Replaces GmresSaddle::solve_many
 * (RS/src/solve/gmres.rs:801 → gmres_batch_on): every solve is the one ddgrs_saddle_solve makes
 * for its right-hand side, bit for bit; the history entry point then reports the last of them. */
DDGRS_API int ddgrs_saddle_solve_many(ddgrs_saddle* h, int64_t nb, const double* bs, double* xs, int64_t restart,
                                      double tolerance, int64_t max_iterations, ddgrs_gmres_report* reports,
                                      ddgrs_error* err);
/* The last solve's stopping measure (relative, explicit) at every cycle start and at the end, and
 * its Arnoldi estimate after every iteration (RS/src/solve/gmres.rs:214-217
 * GmresReport::cycle_residuals, ::history): at most cap_cycles / cap_history values are written
 * (either array may be NULL); counts2 = [cycle residuals, history entries]. */
DDGRS_API int ddgrs_saddle_gmres_history(ddgrs_saddle* h, double* cycle_residuals, int64_t cap_cycles, double* history,
                                         int64_t cap_history, int64_t* counts2, ddgrs_error* err);

/* ---- continuous collision detection ------------------------------------------------------- */

/* CcdOptions (RS/src/ccd.rs:86). */
typedef struct ddgrs_ccd_options {
  double gap;              /* the gap relative to the reference length (0.05) */
  double reference_length; /* ℓ̄; NaN (None): the mean edge length of the start positions */
  int64_t max_iterations;  /* advances per pair and window (100) */
  double close_factor;     /* a pair closer than the gap at t = 0 may close to this fraction (0.5) */
  int64_t leaf_size;       /* leaf size of the broad-phase box trees (8) */
} ddgrs_ccd_options;

/* CcdReport (RS/src/ccd.rs:214).  contact_kind 0: Primitive::VertexTriangle{vertex a, triangle b};
 * 1: Primitive::EdgeEdge{a < b} (ids of the sorted unique edge list, ddgrs_surface_edges). */
typedef struct ddgrs_ccd_report {
  double tau_max;
  int32_t has_contact, contact_kind;
  int64_t contact_a, contact_b;
  double gap_abs;
  int64_t iterations, pairs_tested, close_pairs, touching_pairs, exempt_pairs;
  int32_t windows;
  double seconds;
} ddgrs_ccd_report;

/* Fills CcdOptions::default(). */
DDGRS_API void ddgrs_ccd_options_default(ddgrs_ccd_options* o);
/* The largest τ_max ≤ tau_end such that start + t·velocity keeps every non-adjacent vertex–triangle
 * and edge–edge pair of the surface at distance ≥ gap·ℓ̄ on [0, τ_max] (pairs closer at t = 0: at
 * close_factor of their distance; touching pairs: free), broad and narrow phase on the device.
 * start: 3·nv positions, or NULL for the surface's current ones; velocity: 3·nv; band: nv flags
 * (IntersectionBand, ddgrs_ccd_band) or NULL.  Every pair's advance, the windows and the counts
 * are the reference's, so the report equals its report.  Replaces first_contact_with_band
 * (RS/src/ccd.rs:781) for the mesh alone (static obstacles are not on this path). */
DDGRS_API int ddgrs_ccd_first_contact(ddgrs_surface* s, const double* start, const double* velocity, double tau_end,
                                      const ddgrs_ccd_options* o, const uint8_t* band, ddgrs_ccd_report* report,
                                      ddgrs_error* err);
/* The face pairs [f, g], f < g, that intersect at the positions (NULL: the surface's current), sorted;
 * at most cap are written to pairs (2 each), *count is their number.  Replaces intersecting_pairs
 * (RS/src/ccd.rs:1229). */
DDGRS_API int ddgrs_ccd_intersecting_pairs(ddgrs_surface* s, const double* positions, uint32_t* pairs, int64_t cap,
                                           int64_t* count, ddgrs_error* err);
/* IntersectionBand::new (RS/src/ccd.rs:1292): flags (nv) of the vertices of intersecting faces
 * widened by one ring; counts2 = [intersecting pairs, vertices in the band]. */
DDGRS_API int ddgrs_ccd_band(ddgrs_surface* s, const double* positions, uint8_t* flags, int64_t* counts2,
                             ddgrs_error* err);
/* The sorted unique edges (a < b) of the surface (Surface::edges, RS/src/surface.rs:564): *ne edges,
 * at most cap written to edges (2 each; edges may be NULL). */
DDGRS_API int ddgrs_surface_edges(ddgrs_surface* s, uint32_t* edges, int64_t cap, int64_t* ne, ddgrs_error* err);

/* ---- the descent session -------------------------------------------------------------------- */

/* The line search's first step (RS/src/step/options.rs StartRule; session.rs:1604 start_tau). */
enum ddgrs_start_kind {
  DDGRS_START_GROW_PREVIOUS = 0, /* GrowPrevious{factor 2, first MaxDisplacement{start_value}} */
  DDGRS_START_ABSOLUTE = 1,      /* Absolute{tau = start_value} */
  DDGRS_START_MAX_DISPLACEMENT = 2, /* MaxDisplacement{lengths = start_value}: start_value ℓ̄₀ / max_v |x_v| */
  DDGRS_START_REFERENCE = 3      /* Reference: 2 / ‖x‖ */
};

/* Options of a step (RS/src/step/session.rs FlowOptions with the Hs method, remeshing and CCD
 * off, the global barycenter constraint; the default start is GrowPrevious{2, MaxDisplacement{1}}). */
typedef struct ddgrs_step_options {
  double alpha, beta;
  int32_t energy_exact;    /* 1: exact all pairs; 0: Barnes–Hut */
  double theta;            /* Barnes–Hut θ */
  ddgrs_saddle_options saddle;
  int64_t restart;
  double tolerance;
  int64_t max_iterations;
  double armijo_c;         /* sufficient decrease constant */
  double tau_min_rel;      /* τ_min = tau_min_rel · ℓ̄₀ / ‖x‖_∞ */
  int64_t max_halvings;    /* < 0: no cap (LineSearch::max_halvings) */
  int32_t start_kind;      /* enum ddgrs_start_kind */
  double start_value;      /* τ of Absolute, the lengths of MaxDisplacement (GrowPrevious's first) */
  /* The recovery chain of a diverged solve (RS/src/step/direction.rs:259-338 RobustHierarchical):
   * robust 1 retries with right GMRES at restart 60 and retry_max_iterations (0: at least 800, four
   * times max_iterations), then solves by a dense LU of the exact metric's saddle system on meshes
   * of at most dense_max_vertices (6000) vertices; robust 0 returns SolverDiverged at once.
   * diverged_residual: a solve that ends unconverged with its explicit relative residual above it
   * has diverged (the reference's 0.1, RS/src/solve/gmres.rs). */
  int32_t robust;
  int64_t retry_max_iterations;
  int64_t dense_max_vertices;
  double diverged_residual;
  /* Nonlinear constraint rows after the barycenter's (RS/src/constraints/mod.rs Constraint::Area,
   * ::Volume over the whole mesh): rows[i] 0 none, 1 the total area, 2 the enclosed volume; their
   * targets (NaN: the value at session creation).  They enter the direction through the Schur
   * complement of their rows (RS/src/solve/schur.rs:101 SchurComplement::with_direction: one more
   * GMRES solve per row) and the projection through Newton rounds (schur.rs:495 project):
   * newton_fixed 1 takes exactly newton_count rounds (NewtonSteps::Fixed), 0 repeats while a kept
   * row is above 1e-8 |target| (NewtonSteps::UntilTolerance{max: newton_count}); each round is
   * kept only if it lowers max |Φ_r| / tol_r. */
  int32_t rows[2];
  double row_targets[2];
  int32_t newton_fixed;
  int32_t newton_count;
  /* Continuous collision detection in the line search (LineSearchOptions::ccd, RS/src/step/
   * session.rs:917-956): with ccd 1 every step finds the faces that intersect at its start (their
   * vertices widened by one ring are exempt, IntersectionBand) and, for a finite τ₀ > 0, caps τ₀ at
   * ⅔ τ_max of ddgrs_ccd_first_contact along −x (ccd_options.reference_length NaN: the session's
   * ℓ̄₀). */
  int32_t ccd;
  ddgrs_ccd_options ccd_options;
} ddgrs_step_options;

/* A step's report (RS/src/step/session.rs:146 StepReport, the fields this path produces).
 * accepted 0 is StepStatus::Stuck: the unmoved positions are projected.  energy_after is F at the
 * projected positions; direction_norm ‖x‖ and direction_max ‖x‖_∞ of the direction. */
typedef struct ddgrs_step_report {
  double energy_before, energy_trial, gradient_norm, sobolev_norm_sq, tau, tau0;
  int32_t accepted, trials;
  ddgrs_gmres_report gmres;
  int64_t pairs_exact, nodes_lumped;
  double seconds_trees, seconds_energy, seconds_metric, seconds_precond, seconds_solve, seconds_line_search,
      seconds_projection, seconds_total;
  /* The learned preconditioner (no counterpart in the reference): the session's mode (0 plain,
   * 1 plain solve recording training pairs, 2 (I + Q) P⁻¹); learned_fallback 1 when that solve
   * stagnated and the plain solve's result was used (then `gmres` is the plain solve's report),
   * 2 when Q was not tried because the two solves before fell back (the session then stays plain
   * until set_learned is called again); the iterations of the (I + Q) P⁻¹ solve. */
  int32_t learned_mode, learned_fallback;
  int64_t learned_iterations;
  double seconds_learned; /* features and the network's forward pass */
  double energy_after, direction_norm, direction_max;
  int64_t halvings;
  /* The recovery chain: 0 not needed, 1 the GMRES retry solved the step (then `gmres` is its
   * report), 2 the dense LU did (gmres.relative_residual is its explicit residual). */
  int32_t fallback_stage;
  /* The nonlinear rows: Newton rounds kept by the projection (StepReport::newton_steps) and the
   * smallest kept equilibrated Schur pivot (LinearReport::schur_pivot; 1 without rows). */
  int32_t newton_rounds;
  double min_pivot;
  /* The CCD (StepReport::tau_ccd, ::ccd): has_ccd 1 when it ran (ccd on and τ₀ finite and > 0),
   * then tau_ccd = ccd.tau_max (tau0 above is after the cap); seconds_ccd the band and the CCD. */
  int32_t has_ccd;
  double tau_ccd;
  ddgrs_ccd_report ccd;
  double seconds_ccd;
} ddgrs_step_report;

/* ddgrs_saddle_gmres_history for the last step's direction solve (the last GMRES run of the
 * recovery chain when it escalated; empty after a dense stage). */
DDGRS_API int ddgrs_session_gmres_history(ddgrs_session* h, double* cycle_residuals, int64_t cap_cycles,
                                          double* history, int64_t cap_history, int64_t* counts2, ddgrs_error* err);

/* The escalations of the last step as text (LinearReport::fallbacks, RS/src/solve/mod.rs), one a
 * line, NUL-terminated within `capacity` bytes; *count = the number of escalations. */
DDGRS_API int ddgrs_session_fallbacks(ddgrs_session* h, char* text, int64_t capacity, int64_t* count, ddgrs_error* err);

/* Φ of the constraint rows (the barycenter's x, y, z, then the nonlinear rows) at the last step's
 * moved positions before the projection and after it (schur.rs:495 Projection::phi); *nrows = 3 + the
 * nonlinear rows, or 0 when the session has none (its projection is the barycenter translation
 * alone).  At most `capacity` values of each are written. */
DDGRS_API int ddgrs_session_constraints(ddgrs_session* h, double* phi_before, double* phi, int64_t capacity,
                                        int64_t* nrows, ddgrs_error* err);

/* The Schur complement of k nonlinear rows on the host (RS/src/solve/schur.rs:63 SchurComplement:
 * from_solutions :142, factor :185, direction_from :318, correction :341), the step's own code: from
 * the rows c (k × nd, row-major) and the columns z = A′⁻¹ C_Nᵀ (k × nd, position blocks), the
 * direction x = y − Z N⁻¹ C_N y, the correction h = Z N⁻¹ Φ_N (phi: k values), S = −C_N Z (k × k,
 * as computed) and the smallest kept equilibrated pivot.  Needs no device. */
DDGRS_API int ddgrs_schur(int64_t k, int64_t nd, const double* c, const double* z, const double* y, const double* phi,
                          double* x, double* h, double* s, double* min_pivot, ddgrs_error* err);

/* The trials of the last step in order (StepReport::trials, RS/src/step/session.rs:136 Trial): τ,
 * F (NaN when not finite) and the outcome (TrialOutcome: 0 accepted, 1 Armijo failed, 2 non-finite);
 * at most `capacity` are written, *count is the number of trials. */
DDGRS_API int ddgrs_session_trials(ddgrs_session* h, int64_t capacity, double* tau, double* energy, int32_t* outcome,
                                   int64_t* count, ddgrs_error* err);

/* Fills the reference defaults ((6, 12), Barnes–Hut θ 0.5, restart 50, 1e-6, 200, c 1e-2). */
DDGRS_API void ddgrs_step_options_default(ddgrs_step_options* o);
/* A session on the surface (which it moves): ℓ̄₀ and the barycenter target at the current
 * positions, and the mesh's L′ symbolic analysis (elimination order, fronts) uploaded to the
 * device, so that the first step does no per-mesh work.  Replaces Session::new
 * (RS/src/step/session.rs:371). */
DDGRS_API int ddgrs_session_create(ddgrs_surface* s, const ddgrs_step_options* o, ddgrs_session** out,
                                   ddgrs_error* err);
DDGRS_API void ddgrs_session_destroy(ddgrs_session* h);
/* One descent step; the surface holds the new positions.  Replaces Session::step
 * (RS/src/step/session.rs:724). */
DDGRS_API int ddgrs_session_step(ddgrs_session* h, ddgrs_step_report* report, ddgrs_error* err);
/* [ℓ̄₀, last τ, barycenter target (3)] (session.rs:647 mean_edge_length0). */
DDGRS_API int ddgrs_session_state(ddgrs_session* h, double* out5, ddgrs_error* err);
/* Sets [ℓ̄₀, last τ, barycenter target (3)]: a session created on a remeshed surface continues the
 * run (session.rs:1193 keeps ℓ̄₀, the line search's last τ and the targets across a remeshing). */
DDGRS_API int ddgrs_session_set_state(ddgrs_session* h, const double* in5, ddgrs_error* err);
/* The targets of the area / volume rows (in the order of ddgrs_step_options.rows; NaN for none):
 * the given ones or the values at creation. */
DDGRS_API int ddgrs_session_row_targets(ddgrs_session* h, double* out2, ddgrs_error* err);

/* ---- the learned right preconditioner --------------------------------------------------------
 * No counterpart in the reference, whose solver has the plain preconditioner only.  A mesh
 * network (per-vertex features: position, unit normal, vertex area, mean curvature; four
 * message-passing layers of width 64) maps the geometry to per-vertex 3×3 blocks B_v and per-edge
 * 3×3 blocks G_uv = ½ (g_u + g_v); the step's preconditioner becomes z = (I + Q) P⁻¹ r with
 *   (Q x)_v = s (B_v x_v + Σ_{u ∈ N(v)} G_uv x_u)
 * on the vertex rows (the constraint rows unchanged), a linear map, so GMRES stays a Krylov method
 * with a fixed right preconditioner.  Trained on pairs (r, P⁻¹ r, y = A′⁻¹ r): with the scalar
 * c = ⟨y, P⁻¹ r⟩ / ‖P⁻¹ r‖² (a scalar does not change GMRES's Krylov spaces) and x = c P⁻¹ r, the
 * loss is ‖x + Q x − y‖² / ‖y − x‖² (1 at Q = 0, where the output layer starts).  Vectors are
 * saddle vectors of 3·nv + 3 (FP64, host), positions 3·nv. */
/* A network for the surface's mesh (parameters from `seed`) with room for `max_pairs` pairs. */
DDGRS_API int ddgrs_learned_create(ddgrs_surface* s, uint64_t seed, int64_t max_pairs, ddgrs_learned** out,
                                   ddgrs_error* err);
DDGRS_API void ddgrs_learned_destroy(ddgrs_learned* l);
/* out4 = [parameters, pairs, Adam steps, features per vertex]; *scale = s (0 before training). */
DDGRS_API int ddgrs_learned_info(ddgrs_learned* l, int64_t* out4, double* scale, ddgrs_error* err);
/* Copies the FP32 parameters out (`get`) and then in (`set`); either may be null. */
DDGRS_API int ddgrs_learned_params(ddgrs_learned* l, float* get, const float* set, ddgrs_error* err);
/* Sets s (else the first training sets it to the pairs' median ‖y − x‖ / ‖x‖). */
DDGRS_API int ddgrs_learned_set_scale(ddgrs_learned* l, double scale, ddgrs_error* err);
/* The Krylov pairs a recording step (session mode 1) adds after its own (g, A′⁻¹ g): 0 to 8
 * (default 3); 0 records one (residual, plain-preconditioned solve) pair per step. */
DDGRS_API int ddgrs_learned_set_krylov_pairs(ddgrs_learned* l, int32_t krylov_pairs, ddgrs_error* err);
DDGRS_API int ddgrs_learned_add_pair(ddgrs_learned* l, const double* positions, const double* r, const double* pr,
                                     const double* y, ddgrs_error* err);
/* The features (nv × 8, FP32) and the network's blocks: B_v (nv × 9, row-major), then E_uv (one row
 * of 9 per undirected edge {u < v}, edges in ascending (u, v) order). */
DDGRS_API int ddgrs_learned_features(ddgrs_learned* l, const double* positions, float* feat, ddgrs_error* err);
DDGRS_API int ddgrs_learned_blocks(ddgrs_learned* l, const double* positions, float* blocks, ddgrs_error* err);
/* The loss of one pair and (grad non-null) its gradient with respect to the parameters. */
DDGRS_API int ddgrs_learned_loss_grad(ddgrs_learned* l, int64_t pair, double* loss, float* grad, ddgrs_error* err);
/* Adam (β 0.9 / 0.999) over the pairs in order, `batch` pairs per update; the parameters with the
 * lowest mean loss over all pairs (checked before training, once per epoch of ⌊pairs / batch⌋
 * updates and at the end) are kept, so the loss never ends above its starting value; loss2 = the
 * mean loss over all pairs before and of the kept parameters. */
DDGRS_API int ddgrs_learned_train(ddgrs_learned* l, int64_t iterations, int64_t batch, double lr, double* loss2,
                                  ddgrs_error* err);
/* zs = Q rs at the given positions (the learned map alone, applied to the given vectors). */
DDGRS_API int ddgrs_learned_apply_batch(ddgrs_learned* l, const double* positions, int64_t nb, const double* rs,
                                        double* zs, ddgrs_error* err);
/* Attaches `l` (mesh of the session's surface) to the session: mode 1 solves with the plain
 * preconditioner and adds each step's pairs to `l` (r = the step's right-hand side g with y = the
 * direction, then ddgrs_learned_set_krylov_pairs Krylov pairs r_k = A′ z_{k−1}, y = z_{k−1},
 * z₀ = P⁻¹ g, z_k = P⁻¹ r_k);
 * mode 2 solves with (I + Q) P⁻¹ and, when that solve has not reached the tolerance within twice the
 * last converged solve's iterations plus 5 (one restart cycle before any converged solve; at
 * most the iteration cap), again with the plain one;
 * mode 0 (or l null) detaches.  `l` must stay alive while attached. */
DDGRS_API int ddgrs_session_set_learned(ddgrs_session* h, ddgrs_learned* l, int32_t mode, ddgrs_error* err);

/* ---- remeshing (RS/src/remesh/) ------------------------------------------------------------- */

typedef struct ddgrs_remeshed ddgrs_remeshed;

/* RemeshOptions (RS/src/remesh/mod.rs:96) and the collapse guard: split / collapse thresholds as
 * multiples of L0, flip/smooth rounds, smoothing step rho, flip cap per vertex, order (0 flip then
 * smooth, the default; 1 smooth then flip), which parts run, max_growth (INFINITY disables it);
 * guard_gap > 0 checks every collapse with the CCD gap rule at that absolute gap (the session's
 * CcdCollapseGuard, RS/src/ccd.rs:1448; 0: no guard). */
typedef struct ddgrs_remesh_options {
  double split, collapse;
  int32_t rounds;
  double rho;
  int64_t flip_cap_per_vertex;
  int32_t order;
  int32_t split_on, collapse_on, flip_on, smooth_on;
  double max_growth;
  double guard_gap;
} ddgrs_remesh_options;

/* RemeshReport (RS/src/remesh/mod.rs:187).  refused_*[k]: refusals of kind k in the order of
 * RefusalKind::ALL (RS/src/remesh/mesh.rs:75: NoSuchEdge, BothPinned, LinkCondition, Foldover,
 * RepeatedFace, NonManifoldResult, BoundaryEdge, NotDelaunay, ExistingEdge, Guard, DegenerateFace);
 * smoothing_iterations: passes of the parallel guarded smoothing over all rounds (no counterpart). */
typedef struct ddgrs_remesh_report {
  int64_t splits, collapses, flips, splits_deferred;
  int64_t refused_splits[11], refused_collapses[11], refused_flips[11];
  int64_t smoothing_moves, smoothing_halved, smoothing_refused, smoothing_skipped;
  int64_t flip_cap_hits;
  int64_t n_vertices, n_faces;
  int64_t smoothing_iterations;
  double seconds;
} ddgrs_remesh_report;

DDGRS_API void ddgrs_remesh_options_default(ddgrs_remesh_options* o);
/* One remeshing pass of the mesh (nv vertices, nf faces, positions 3·nv) on the device: splits of
 * the long edges, collapses of the short ones shortest first, Delaunay flips and tangential
 * smoothing, the result numbered as the reference numbers it.  pinned: npinned vertex ids (never
 * moved or removed); band: nv flags exempting pairs of the collapse guard (IntersectionBand,
 * ddgrs_ccd_band) or NULL.  Replaces remesh (RS/src/remesh/mod.rs:235) with the session's guard.
 * The result (read with ddgrs_remeshed_*) must be destroyed. */
DDGRS_API int ddgrs_remesh(ddgrs_ctx* ctx, int64_t nv, int64_t nf, const uint32_t* faces, const double* positions,
                           int64_t npinned, const uint32_t* pinned, double l0, const ddgrs_remesh_options* o,
                           const uint8_t* band, ddgrs_remeshed** out, ddgrs_remesh_report* report, ddgrs_error* err);
/* One EditMesh operation (op 0 split, 1 collapse, 2 flip; RS/src/remesh/mesh.rs:592, :679, :888)
 * with the optional Delaunay / foldover checks, then EditMesh::finish (:991).  *refusal: 0 done,
 * else 1 + the RefusalKind index; *reordered: the flip's FlipOutcome::reordered. */
DDGRS_API int ddgrs_remesh_op(ddgrs_ctx* ctx, int64_t nv, int64_t nf, const uint32_t* faces, const double* positions,
                              int64_t npinned, const uint32_t* pinned, int32_t op, uint32_t a, uint32_t b,
                              int32_t delaunay, int32_t foldover, int32_t* refusal, int32_t* reordered,
                              ddgrs_remeshed** out, ddgrs_error* err);
/* sizes4 = [vertices, faces, vertex-map entries, pins] (Remeshed, RS/src/remesh/mesh.rs:210). */
DDGRS_API int ddgrs_remeshed_sizes(const ddgrs_remeshed* r, int64_t* sizes4, ddgrs_error* err);
/* positions 3·nv, faces 3·nf, face_parents nf, the vertex map (VertexMap, RS/src/remesh/
 * transport.rs:15) as map_offsets nv + 1, map_sources / map_weights per entry, and the new ids of
 * the input pins in the caller's order. */
DDGRS_API int ddgrs_remeshed_read(const ddgrs_remeshed* r, double* positions, uint32_t* faces, uint32_t* face_parents,
                                  uint32_t* map_offsets, uint32_t* map_sources, double* map_weights, uint32_t* pinned,
                                  ddgrs_error* err);
DDGRS_API void ddgrs_remeshed_destroy(ddgrs_remeshed* r);

#ifdef __cplusplus
}
#endif
#endif /* DDGRS_H */
