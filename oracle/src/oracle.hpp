// FP64 oracle — TEST INFRASTRUCTURE ONLY.  Declarations shared by the oracle's translation units.
#pragma once
#include <atomic>

#include "core.hpp"

namespace orc {

struct FacePartials {
  double energy = 0.0;
  std::vector<double> d_area;
  std::vector<V3> d_bary, d_normal;
};

FacePartials exact_face_partials(const FaceGeometry& g, const Exponents& e, int threads);
double exact_energy_value(const FaceGeometry& g, const Exponents& e, int threads);
bool exact_energy_and_differential(const Mesh& m, const Positions& p, const Exponents& e, int threads,
                                   double* energy, std::vector<V3>* diff, FacePartials* partials,
                                   Error* err);

// ---------------------------------------------------------------------------------------------
// Axis-aligned boxes (crates/ddg-repulsive-curves/src/spatial/aabb.rs)

struct Aabb {
  V3 lo{INFINITY, INFINITY, INFINITY};
  V3 hi{-INFINITY, -INFINITY, -INFINITY};
  bool is_empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
  void grow(V3 p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin_(lo[k], p[k]);
      hi[k] = fmax_(hi[k], p[k]);
    }
  }
  Aabb unite(const Aabb& o) const {
    Aabb r = *this;
    for (int k = 0; k < 3; ++k) {
      r.lo[k] = fmin_(r.lo[k], o.lo[k]);
      r.hi[k] = fmax_(r.hi[k], o.hi[k]);
    }
    return r;
  }
  // aabb.rs:73 `diameter_sq`
  double diameter_sq() const {
    if (is_empty()) return 0.0;
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
      double w = hi[k] - lo[k];
      s += w * w;
    }
    return s;
  }
};

// spatial/mod.rs:113 `box_distance_sq`
inline double box_distance_sq(const Aabb& a, const Aabb& b) {
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double d = fmax_(fmax_(a.lo[k], b.lo[k]) - fmin_(a.hi[k], b.hi[k]), 0.0);
    d2 += d * d;
  }
  return d2;
}

// spatial/mod.rs:125 `farthest_corner_distance`
inline double farthest_corner_distance(const Aabb& b, V3 x) {
  double r2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double d = fmax_(b.hi[k] - x[k], x[k] - b.lo[k]);
    r2 += d * d;
  }
  return std::sqrt(r2);
}

// ---------------------------------------------------------------------------------------------
// BVH (crates/ddg-repulsive-curves/src/spatial/bvh.rs)

struct BvhNode {
  Aabb bbox;
  double mass = 0.0;
  V3 moment{0.0, 0.0, 0.0};
  uint32_t start = 0, end = 0, right = 0;
  bool is_leaf() const { return right == 0; }
  V3 centroid() const { return {moment[0] / mass, moment[1] / mass, moment[2] / mass}; }
};

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> order;
  size_t leaf_size = 8;
};

// bvh.rs:126 `build_with` (min_fraction 0.25) followed by :221 `refit`.
Bvh bvh_build(const std::vector<V3>& pts, const std::vector<Aabb>& bnds, const std::vector<double>& mass,
              size_t leaf_size, double min_fraction = 0.25);
void bvh_refit(Bvh& b, const std::vector<V3>& pts, const std::vector<Aabb>& bnds,
               const std::vector<double>& mass);

// ---------------------------------------------------------------------------------------------
// Face tree (crates/ddg-repulsive-surfaces/src/spatial/mod.rs)

struct Cluster {
  double area = 0.0;
  V3 centroid{0, 0, 0};
  std::array<double, 6> projector{};
  Aabb bbox;
  double radius = 0.0;
  // spatial/mod.rs:101 `projector_quadratic_form`
  double pq(V3 v) const {
    const auto& p = projector;
    return p[0] * v[0] * v[0] + p[3] * v[1] * v[1] + p[5] * v[2] * v[2] +
           2.0 * (p[1] * v[0] * v[1] + p[2] * v[0] * v[2] + p[4] * v[1] * v[2]);
  }
};

struct FaceTree {
  Bvh bvh;
  std::vector<Cluster> clusters;
  std::vector<uint32_t> leaves, leaf_ordinal, leaf_of_face;
};

std::vector<Aabb> face_bounds(const Mesh& m, const Positions& p);
FaceTree face_tree_build(const Mesh& m, const Positions& p, const FaceGeometry& g, size_t leaf_size);
void face_tree_refit(FaceTree& t, const Mesh& m, const Positions& p, const FaceGeometry& g);

// ---------------------------------------------------------------------------------------------
// Barnes–Hut (crates/ddg-repulsive-surfaces/src/energy/barnes_hut.rs)

struct InteractionLists {
  std::vector<uint64_t> lumped_off, exact_off, partner_off, lumped_t_off, exact_t_off;
  std::vector<uint32_t> lumped, exact, lumped_t, exact_t;
  uint64_t pairs_exact = 0;
};

InteractionLists bh_build_lists(const FaceTree& t, double theta, int threads);

struct BhTree {
  double theta = 0.5;
  size_t leaf_size = 8;
  Mesh mesh;
  Positions positions;
  FaceGeometry geom;
  FaceTree tree;
  InteractionLists lists;
};

bool bh_build(const Mesh& m, const Positions& p, double theta, size_t leaf_size, int threads, BhTree* out,
              Error* err);
bool bh_refit(BhTree& bh, const Positions& p, Error* err);
double bh_energy(const BhTree& bh, const Exponents& e, int threads);
FacePartials bh_face_partials(const BhTree& bh, const Exponents& e, int threads);
bool bh_energy_and_differential(const BhTree& bh, const Exponents& e, int threads, double* energy,
                                std::vector<V3>* diff, Error* err);

// ---------------------------------------------------------------------------------------------
// Block cluster tree and hierarchical products (spatial/block_tree.rs, metric/hierarchical.rs)

struct ClusterKernelFns {
  std::function<bool(size_t, size_t)> admissible;
  std::function<void(size_t, size_t, double*)> far;   // N values
  std::function<void(size_t, size_t, double*)> near;  // N values
  std::function<double(size_t)> weight;
};

struct BctStats {
  uint64_t admissible_blocks = 0, inadmissible_blocks = 0, admissible_pairs = 0, near_entries = 0, nodes = 0,
           depth = 0;
};

struct BlockTree {
  int N = 1;
  std::vector<uint32_t> start, end, right, order, position, leaves, chain;
  std::vector<uint64_t> chain_ptr;
  std::vector<double> weight;
  std::vector<uint64_t> far_ptr;
  std::vector<uint32_t> far_col;
  std::vector<double> far_coef;
  std::vector<uint64_t> near_ptr, near_off;
  std::vector<uint32_t> near_col;
  std::vector<double> near_val;
  uint64_t near_total = 0;
  BctStats stats;
};

BlockTree block_tree_build(const Bvh& bvh, const ClusterKernelFns& K, int N, size_t leaf_pairs, bool fill);
std::vector<double> block_tree_multiply_permuted(const BlockTree& t, const std::vector<int>& map,
                                                 const std::vector<double>& psi, int threads);
std::vector<double> block_tree_multiply(const BlockTree& t, const std::vector<int>& map, const std::vector<double>& psi,
                                        int threads);
std::array<V3, 3> df_coefficients(const std::array<V3, 3>& p);

struct FaceMetric {
  Mesh mesh;
  FaceGeometry geom;
  FaceTree tree;
  std::vector<std::array<V3, 3>> df;
  BlockTree bct;
  std::vector<double> ha, ha_perm;  // 2 per face
  std::vector<uint64_t> inc_ptr;
  std::vector<std::pair<uint32_t, uint8_t>> inc;
};

bool face_metric_build(const Mesh& m, const Positions& p, const Exponents& e, double chi, size_t leaf_size,
                       size_t leaf_pairs, int threads, FaceMetric* out, Error* err);
void face_metric_apply(const FaceMetric& fm, double w_hi, double w_lo, size_t nb, const double* xs, double* ys,
                       int threads);

struct VertexFractional {
  Bvh bvh;
  BlockTree bct;
  std::vector<double> mass, ha;
};

bool vertex_fractional_build(const Positions& p, const std::vector<double>& mass, double tau, double chi,
                             size_t leaf_size, size_t leaf_pairs, int threads, VertexFractional* out, Error* err);
void vertex_fractional_apply(const VertexFractional& vf, size_t nb, const double* xs, double* ys, int threads);

// ---------------------------------------------------------------------------------------------
// Saddle solver and descent step (solve/{precond,gmres,iterative}.rs, step/session.rs)

struct Triplet {
  uint32_t r, c;
  double v;
};
struct Csr {
  size_t nrows = 0, ncols = 0;
  std::vector<size_t> ptr;
  std::vector<uint32_t> idx;
  std::vector<double> val;
};
Csr csr_from_triplets(size_t nrows, size_t ncols, const std::vector<Triplet>& t);
std::vector<Triplet> cotan_triplets(const Mesh& m, const Positions& p);
double mean_edge_length(const Mesh& m, const Positions& p);
std::vector<uint32_t> nested_dissection(const Csr& k, const Positions& p);

struct Ldlt {
  size_t n = 0;
  std::vector<uint32_t> perm;  // perm[new] = old
  std::vector<int64_t> parent;
  std::vector<uint64_t> lp;
  std::vector<uint32_t> li;
  std::vector<double> lx, d;
};
bool ldlt_factor(const Csr& a, const std::vector<uint32_t>& perm, Ldlt* f, Error* err);
void ldlt_solve(const Ldlt& f, const double* b, double* x);
double fractional_symbol_ratio(double s);

// precond.rs:466 PreconditionerOptions: form (0 Augmented, 1 Congruence), scale (0 Symbol, 1 Unit),
// refine (one step of iterative refinement in every L′ solve of an application).
struct SaddleOptions {
  double chi = 0.5;
  size_t leaf_size = 8, leaf_pairs = 64;
  double w_tp = 1.0, c_lo = 1.0;
  double eps_rel = 1e-10;
  int form = 0, scale = 0;
  bool refine = false;
};

struct Saddle {
  size_t nv = 0;
  double w_tp = 1.0, c_lo = 1.0, eps = 0.0, frac_scale = 1.0;
  int form = 0;
  bool refine = false;
  FaceMetric metric;
  std::vector<double> mass, crow;
  Csr k, lprime_matrix;
  Ldlt lprime;
  VertexFractional frac;
};
bool saddle_build(const Mesh& m, const Positions& p, const Exponents& e, const SaddleOptions& o, int threads,
                  Saddle* out, Error* err);
void saddle_apply(const Saddle& sd, const double* x, double* y, int threads);
void precond_apply(const Saddle& sd, const double* r, double* z, int threads);

struct GmresOptions {
  size_t restart = 50;
  double tolerance = 1e-6;
  size_t max_iterations = 200;
};
struct GmresReport {
  size_t iterations = 0, cycles = 0, operator_applications = 0, preconditioner_applications = 0;
  bool converged = false;
  double relative_residual = 0.0;
  std::vector<double> cycle_residuals, history;
};
GmresReport gmres_right(const std::function<void(const double*, double*)>& op,
                        const std::function<void(const double*, double*)>& prec, const std::vector<double>& b,
                        const GmresOptions& o, std::vector<double>* x);

// The line search's first step (step/options.rs StartRule; session.rs:1604 start_tau).
enum StartKind { START_GROW_MAXDISP = 0, START_ABSOLUTE = 1, START_MAXDISP = 2, START_REFERENCE = 3 };
// Nonlinear constraint rows (constraints/mod.rs RowKind::Area, ::Volume) of one component.
enum { ROW_AREA = 1, ROW_VOLUME = 2 };
double total_area(const Mesh& m, const Positions& p);
double signed_volume(const Mesh& m, const Positions& p);
std::vector<double> area_gradient(const Mesh& m, const Positions& p);    // 3 nv
std::vector<double> volume_gradient(const Mesh& m, const Positions& p);  // 3 nv
double row_value(const Mesh& m, const Positions& p, int kind);
std::vector<double> row_gradient(const Mesh& m, const Positions& p, int kind);
double row_tolerance(double target);  // 1e-8 |target|, or 1e-8 for a zero target

// The Schur complement of the nonlinear rows (schur.rs SchurComplement): C_N (k dense rows of
// 3 nv), Z = A′⁻¹ C_Nᵀ (position blocks), S = −C_N Z, the equilibration and the pivoted Cholesky
// factor of the kept rows.
struct Schur {
  std::vector<std::vector<double>> c, z;
  std::vector<double> s, d, l;
  std::vector<size_t> kept;  // pivot order
  double min_pivot = 1.0;
};
void schur_build(Schur* S, std::vector<std::vector<double>> c, std::vector<std::vector<double>> z);
std::vector<double> schur_solve_n(const Schur& S, const std::vector<double>& v);
std::vector<double> schur_combine(const Schur& S, const std::vector<double>& c);
std::vector<double> schur_apply_c(const Schur& S, const double* v);
std::vector<double> schur_direction(const Schur& S, const std::vector<double>& y);   // y − Z N⁻¹ C_N y
std::vector<double> schur_correction(const Schur& S, const std::vector<double>& phi);  // Z N⁻¹ Φ_N

// ---- continuous collision detection options and report (ccd.cpp) ----
struct CcdOptions {
  double gap = 0.05;                     // relative to the reference length
  double reference_length = std::nan("");  // NaN: the mean edge length of the start positions
  uint32_t max_iterations = 100;
  double close_factor = 0.5;
  size_t leaf_size = 8;
};
// kind 0: vertex a against triangle b; kind 1: edges a < b (ids of the sorted unique edge list)
struct CcdContact {
  double tau = 0.0;
  int kind = 0;
  uint32_t a = 0, b = 0;
};
struct CcdReport {
  double tau_max = 0.0, gap_abs = 0.0;
  bool has_contact = false;
  CcdContact contact;
  uint64_t iterations = 0, pairs_tested = 0, close_pairs = 0, touching_pairs = 0, exempt_pairs = 0;
  uint32_t windows = 0;
};

struct SessionOptions {
  bool energy_exact = false;  // EnergyEval::Exact instead of Barnes–Hut
  double theta = 0.5;
  SaddleOptions saddle;
  GmresOptions gmres;
  int start_kind = START_GROW_MAXDISP;
  double start_value = 1.0;   // τ (absolute) or the lengths of MaxDisplacement
  double armijo_c = 1e-2;     // ArmijoRule::Standard{c}
  double tau_min_rel = 1e-12;
  long max_halvings = -1;     // < 0: no cap
  // nonlinear rows after the barycenter's (ROW_AREA, ROW_VOLUME), their targets (NaN: the value
  // at creation) and the projection's Newton rounds (schur.rs NewtonSteps: Fixed(n) or
  // UntilTolerance{max: n})
  std::vector<int> rows;
  std::vector<double> targets;
  bool newton_fixed = false;
  int newton_count = 5;
  // the CCD cap of τ₀ (session.rs:917-956), reference_length NaN: the session's ℓ̄₀
  bool ccd = false;
  CcdOptions ccd_opts;
};
struct Session {
  Mesh mesh;
  Positions positions;
  Exponents exps;
  SessionOptions opts;
  double l0 = 0.0, last_tau = 0.0;
  V3 barycenter_target{0, 0, 0};
  std::vector<double> row_targets;  // of the nonlinear rows
  // the last step's constraint rows (as in StepReport)
  std::vector<double> last_phi_before, last_phi;
  bool last_has_ccd = false;
  CcdReport last_ccd;
  int last_newton_rounds = 0;
  double last_min_pivot = 1.0;
};
// TrialOutcome (session.rs): 0 accepted, 1 Armijo failed, 2 non-finite.
struct Trial {
  double tau = 0.0, energy = 0.0;
  int outcome = 0;
};
struct StepReport {
  double energy0 = 0.0, energy1 = 0.0, energy_after = 0.0, slope = 0.0, tau = 0.0, tau0 = 0.0;
  double gradient_norm = 0.0, direction_norm = 0.0, direction_max = 0.0;
  bool accepted = false;
  size_t trials = 0, halvings = 0;
  std::vector<Trial> trial;
  GmresReport gmres;
  // wall seconds of the phases: trees and F, dF; the solver's build; GMRES; line search; projection
  double seconds[5] = {0, 0, 0, 0, 0};
  // the constraint rows: Φ (barycenter x, y, z, then the nonlinear rows) at the moved positions
  // before the projection and after it, the Newton rounds kept, the smallest kept Schur pivot
  std::vector<double> phi_before, phi;
  int newton_rounds = 0;
  double min_pivot = 1.0;
  // the CCD (when it ran: ccd on, τ₀ finite and positive); tau0 above is after its cap
  bool has_ccd = false;
  CcdReport ccd;
};
bool session_create(const Mesh& m, const Positions& p, const Exponents& e, const SessionOptions& o, Session* s,
                    Error* err);
bool session_step(Session& s, int threads, StepReport* rep, Error* err);

// ---- continuous collision detection (ccd.cpp; crates/ddg-repulsive-surfaces/src/ccd.rs) ----
double point_segment_distance(V3 p, V3 a, V3 b);
double segment_segment_distance(V3 p1, V3 q1, V3 p2, V3 q2);
double point_triangle_distance(V3 p, V3 a, V3 b, V3 c);
bool segment_meets_triangle(V3 p, V3 q, const std::array<V3, 3>& tri);
bool triangles_intersect(const std::array<V3, 3>& t1, const std::array<V3, 3>& t2);
std::vector<std::array<uint32_t, 2>> mesh_edges(const Mesh& m);
std::vector<std::array<uint32_t, 2>> intersecting_pairs(const Mesh& m, const Positions& p);
std::vector<uint8_t> intersection_band(const Mesh& m, const Positions& p, size_t* n_pairs, size_t* n_vertices);
double min_distance(const Mesh& m, const Positions& p);
bool ccd_validate(const CcdOptions& o, Error* err);
// band: per-vertex flags or nullptr (an empty band is no band)
bool first_contact(const Mesh& m, const Positions& start, const Positions& vel, double tau_end, const CcdOptions& o,
                   const std::vector<uint8_t>* band, CcdReport* rep, Error* err);

// ---- remeshing (remesh.cpp; crates/ddg-repulsive-surfaces/src/remesh/) ------------------------
// Refusal kinds in the order of mesh.rs:75 `RefusalKind::ALL` (1-based; 0 = done).
enum RemeshRefusal : int {
  REMESH_OK = 0,
  REMESH_NO_SUCH_EDGE = 1,
  REMESH_BOTH_PINNED = 2,
  REMESH_LINK_CONDITION = 3,
  REMESH_FOLDOVER = 4,
  REMESH_REPEATED_FACE = 5,
  REMESH_NON_MANIFOLD_RESULT = 6,
  REMESH_BOUNDARY_EDGE = 7,
  REMESH_NOT_DELAUNAY = 8,
  REMESH_EXISTING_EDGE = 9,
  REMESH_GUARD = 10,
  REMESH_DEGENERATE_FACE = 11,
};
// mod.rs:96 `RemeshOptions` (order 0: flip then smooth, 1: smooth then flip)
struct RemeshOptions {
  double split = 1.5, collapse = 0.5;
  int rounds = 5;
  double rho = 0.5;
  size_t flip_cap_per_vertex = 100;
  int order = 0;
  bool en_split = true, en_collapse = true, en_flip = true, en_smooth = true;
  double max_growth = 2.0;
};
// mod.rs:187 `RemeshReport` (+ the collapse guard's counters)
struct RemeshReport {
  uint64_t splits = 0, collapses = 0, flips = 0;
  uint64_t refused_splits[11] = {}, splits_deferred = 0, refused_collapses[11] = {}, refused_flips[11] = {};
  uint64_t smoothing_moves = 0, smoothing_halved = 0, smoothing_refused = 0, smoothing_skipped = 0;
  uint32_t flip_cap_hits = 0;
  size_t n_vertices = 0, n_faces = 0;
  uint64_t guard_refused = 0, guard_tested = 0, guard_exempt = 0;
};
// mesh.rs:210 `Remeshed` with the vertex map in CSR form (transport.rs:15)
struct Remeshed {
  Positions positions;
  std::vector<std::array<uint32_t, 3>> faces;
  std::vector<uint32_t> face_parents, map_offsets, map_sources, pinned;
  std::vector<double> map_weights;
  size_t n_old = 0;
};
struct EditMesh;
struct CollapseGuard;
bool edit_mesh_new(const Mesh& s, const Positions& p, const std::vector<uint32_t>& pins, EditMesh** out, Error* err);
void edit_mesh_free(EditMesh* m);
// op 0 split, 1 collapse, 2 flip; returns a RemeshRefusal
int edit_mesh_op(EditMesh& m, int op, uint32_t a, uint32_t b, bool delaunay, bool foldover, bool* reordered);
bool edit_mesh_finish(const EditMesh& m, Remeshed* out, Error* err);
// guard_gap > 0: the CCD collapse guard with that absolute gap and the optional band
bool remesh(const Mesh& s, const Positions& p, const std::vector<uint32_t>& pins, double l0, const RemeshOptions& o,
            double guard_gap, const std::vector<uint8_t>* band, Remeshed* out, RemeshReport* rep, Error* err);
double remesh_angle_between(V3 u, V3 v);
V3 remesh_tri_normal(const std::array<V3, 3>& p);
bool remesh_folds(V3 before, V3 after);
bool remesh_violates_delaunay(V3 p, V3 q, V3 c, V3 d);
V3 remesh_circumcenter(const std::array<V3, 3>& p);
V3 remesh_barycenter(const std::array<V3, 3>& p);
bool remesh_smoothing_displacement(V3 x, double rho, const std::vector<std::pair<double, V3>>& faces, V3* u);
double remesh_triangle_distance(const std::array<V3, 3>& t1, const std::array<V3, 3>& t2);

}  // namespace orc
