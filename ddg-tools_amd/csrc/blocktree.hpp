// The block cluster tree on the device and the hierarchical products built on it:
// B̃, B̃⁰ over face pairs (RS/src/metric/hierarchical.rs:283 FaceMetricTree) and the vertex-pair
// fractional operator L̃^V_{2−s} (hierarchical.rs:1003 VertexFractionalTree).
#pragma once
#include "kernels.hpp"
#include "trees.hpp"

namespace ddgrs {

// Rows of a near-field panel (a leaf's rows are split into panels of at most this many).
constexpr uint32_t kPanelRows = 4;
// Columns of a panel segment, the unit of work of the near-field kernels (a multiple of 64).  A few
// leaves span large parts of a surface (the BVH keeps long thin leaves along mesh seams) and are
// near tens of thousands of columns; split, their panels spread over many waves.
constexpr uint32_t kSegCols = 512;
// The segment length: kSegCols, or DDGRS_SEGMENT_COLUMNS (a positive multiple of 64, at most
// kSegCols) when set.
uint32_t panel_segment_cols();

// Host part of a BlockTreeDev: the structure, panels and weights (built without the device).
struct BlockTreePlan {
  int nk = 1;
  int n = 0, nn = 0, nl = 0, np = 0;
  uint64_t near_total = 0;
  std::vector<uint32_t> order, position, start, end, right, parent, leaves;
  std::vector<uint32_t> col_ptr, cols, panel_leaf, panel_row0, panel_rows, far_ptr, far_col, far_row;
  std::vector<uint64_t> panel_val;
  // segments: per segment its panel and first column; per panel its first segment (np + 1)
  uint32_t seg_cols = kSegCols;
  std::vector<uint32_t> seg_panel, seg_col0, seg_ptr;
  std::vector<double> weight, inv_mass;
};
BlockTreePlan plan_block_tree(const Topology& t, const BlockStructure& s, int nk, std::vector<double> weight_bvh,
                              const std::vector<double>& node_mass);

// Counts of a block structure (RS/src/spatial/block_tree.rs:340-346): blocks, admissible pairs,
// near entries (diagonal blocks without their diagonal), stored near-block entries, levels.
struct BlockStats {
  uint64_t far_blocks = 0, near_blocks = 0, admissible_pairs = 0, near_entries = 0, near_total = 0;
  uint32_t depth = 0;
};
BlockStats block_stats(const BlockStructure& s);

// Per block pair of a level: admissible, near and child pairs it produces (and their prefix sums).
struct BlockCounts {
  uint32_t adm, nearb, kids;
};
// Device scratch of a structure build (kept between builds).
struct BlockBuildScratch {
  DevBuf<uint32_t> pa, pb, qa, qb, adm_a, adm_b, near_a, near_b, near_row, near_col, near_ptr;
  DevBuf<uint32_t> rcount, kcount, pcount, range_ptr, pan_ptr, range_end, range_end_s, leaf_first, leaf_end, cursor;
  DevBuf<uint32_t> scount, seg_base, range_len, range_off;
  DevBuf<uint64_t> entries, val_base, range_key, range_key_s;
  DevBuf<double> lo, hi, radius;
  DevBuf<uint8_t> cls, temp;
  DevBuf<BlockCounts> cnt, off;
  DevBuf<unsigned long long> sums;
  // Device-driven levels: per-block counts and offsets, the level state, and list capacities from
  // earlier builds (0 before the first).
  DevBuf<BlockCounts> bsum, boff;
  DevBuf<uint32_t> lstate, level_m;
  std::vector<uint32_t> prev_m;  // pairs per level of the previous build (sizes the level grids)
  uint64_t cap_pairs = 0, cap_adm = 0, cap_near = 0;
  uint32_t last_depth = 0;
};

// Device form of a block structure over a BVH, for products in difference form
//   z_S = Σ_{T near S} v_ST (u_S − u_T) + Σ_{far (A, B), S ∈ A} c_AB (u_S − ū_B),
// v_ST = H_ST a_T (near values × the column weight), c_AB = κ_AB A_B (far coefficient × the column
// cluster's weight), ū_B the weighted mean of u over B.  This equals the reference's
// H̃a·u − H̃(a u) (hierarchical.rs:696 product) term by term, but no large term cancels: every
// FP32 rounding is relative to a difference of u, not to u.
//
// Near field: one panel per leaf (split into panels of at most kPanelRows rows): the leaf's rows against
// every column that a near block of one of its ancestors covers, stored dense [row][column] with NK
// FP32 values per entry.  Far field: the admissible blocks per row node.  The far part is pushed to
// the leaves exactly: for S in leaf L,
//   Σ_{A ∋ S} Σ_B c_AB (u_S − ū_B) = ᾱ_L (u_S − ū_L) + δ_L,
//   ᾱ_L = Σ_{A ⊇ L} α_A, α_A = Σ_B c_AB,  δ_L = Σ_{A ⊇ L} [α_A (ū_L − ū_A) + γ_A],
//   γ_A = Σ_B c_AB (ū_A − ū_B)   (FP64, from FP64 node means).
struct BlockTreeDev {
  int nk = 1;
  int n = 0, nn = 0, nl = 0, np = 0;
  uint64_t near_total = 0;  // stored near entries (Σ over panels of rows × columns)
  uint32_t n_far = 0, n_near = 0;
  DevBuf<uint32_t> order, position, start, end, right, parent, leaves;
  DevBuf<uint32_t> panel_leaf, panel_row0, panel_rows, col_ptr, cols;
  DevBuf<uint64_t> panel_val;  // per panel: first entry of its values
  // Panel segments (at most seg_cols columns each), one wave each in the near-field kernels.  A
  // panel of several segments is finished by the last of its segments to arrive, which sums the
  // segments' partial sums in segment order (FP64).
  uint32_t seg_cols = kSegCols;
  int nseg = 0;
  DevBuf<uint32_t> seg_panel, seg_col0, seg_ptr;
  DevBuf<uint32_t> seg_arrive;  // per panel, zero between products
  DevBuf<float> partial;        // per segment: kPanelRows × C near sums
  DevBuf<double> partial64;     // the same for the extended product
  DevBuf<float> val;           // NK per entry, then NK zeros (read by masked lanes)
  DevBuf<uint32_t> far_ptr, far_col, far_row;
  DevBuf<float> far_coef;      // NK per far entry: κ_AB A_B
  DevBuf<double> weight;       // per BVH position: a (FP64)
  DevBuf<double> inv_mass;     // per node: 1 / A
  DevBuf<double> alpha_node;   // per node, NK
  DevBuf<double> alpha_leaf;   // per leaf ordinal, NK
  DevBuf<uint32_t> counters;   // per node, zero between products (upward pass)
  DevBuf<double> sigma, mean, gamma, delta;  // product scratch (C per node / leaf)
  BlockBuildScratch scratch;

  // Structure, panels and weights; near values and far coefficients are filled by the caller's
  // kernels, then finish_far() computes α.
  void upload(ddgrs_ctx* ctx, const BlockTreePlan& p);
  // The same structure, panels and weights built on the device (bitwise equal to
  // upload(plan_block_tree(...)) of build_block_structure's lists): level-synchronous
  // classification of the block pairs, stable grouping by row node, panels from the near lists.
  // `elem_weight`: per element in the tree's input numbering (device pointer).
  // `dt`: the tree and clusters on the device (copied there instead of uploaded), or null.
  BlockStats build_structure(ddgrs_ctx* ctx, const Topology& t, const Clusters& c, double chi, size_t leaf_pairs,
                             int nk, const double* elem_weight, const DeviceTree* dt = nullptr);
  void finish_far(ddgrs_ctx* ctx);
  // A host-planned structure's near lists, for lists() (upload() carries the panels only).
  void set_near_lists(ddgrs_ctx* ctx, const BlockStructure& s);
  // The tree and the grouped block lists of the last build, on the host: order (n), start, end,
  // right (nn), far_ptr and near_ptr (nn + 1), far_col (n_far) and near_col (n_near); per row node
  // the column nodes in creation order (RS/src/spatial/block_tree.rs:198 group_by_row).
  void lists(ddgrs_ctx* ctx, std::vector<uint32_t>* order_h, std::vector<uint32_t>* start_h,
             std::vector<uint32_t>* end_h, std::vector<uint32_t>* right_h, std::vector<uint32_t>* far_ptr_h,
             std::vector<uint32_t>* far_col_h, std::vector<uint32_t>* near_ptr_h,
             std::vector<uint32_t>* near_col_h);
  // Room for the near values and the zero entry after them.
  void reserve_values(ddgrs_ctx* ctx);
  // z = w_k (H̃a·u − H̃(a u)) per BVH position and channel, channels [0, H) on kernel 0 with w0,
  // [H, C) on kernel NK − 1 with w1.  u and z are hi/lo FP32 pairs: 2 C floats per position
  // (C hi, then C lo).
  // extended: the near field's differences and sums in FP64 (the values stay FP32), for solves that
  // need a residual below the FP32 product's rounding.
  template <int C, int H>
  void multiply(ddgrs_ctx* ctx, const float* u_hl, double w0, double w1, float* z_hl, bool extended = false);
};

// Host part of a FaceMetricDev on a given face tree.
struct FaceMetricPlan {
  Topology topo;
  BlockStructure bs;
  BlockTreePlan bt;
  std::vector<double> centroid, projector, node_mass;
  std::vector<uint32_t> inc_ptr, inc;  // vertex → (BVH position of the face, corner), faces ascending
};
FaceMetricPlan plan_face_metric(const std::vector<uint32_t>& faces, size_t nv, const Topology& t, const Clusters& c,
                                double chi, size_t leaf_pairs, const std::vector<double>& face_area);

// B̃, B̃⁰ of one configuration (hierarchical.rs:283).
struct FaceMetricDev {
  Topology topo;
  BlockStats stats;
  BlockTreeDev bt;
  int nv = 0, nf = 0;
  DevBuf<float> df;         // 9 per BVH position: g_c(S) for c = 0..2
  DevBuf<double> area_p;    // per BVH position
  DevBuf<uint32_t> vert_p;  // 3 per BVH position
  DevBuf<float> m_hl, z_hl; // product scratch: 12 channels, hi/lo
  DevBuf<uint32_t> inc_ptr, inc;  // vertex → (face, corner), with faces' BVH positions
  DevBuf<double> node_cen, node_proj, node_mass;  // the build's cluster data (far coefficients)
  double w_hi = 1.0, w_lo = 1.0;
  bool extended = false;  // products with the extended near field (BlockTreeDev::multiply)

  // Builds on a face tree of the surface's current positions with separation χ and leaf_pairs
  // (structure and panels on the device).
  void build(ddgrs_surface* s, const Topology& t, const Clusters& c, double chi, size_t leaf_pairs, double s_exp);
  // The same on ctx's stream (the surface's device data must be current there).
  void build(ddgrs_ctx* ctx, ddgrs_surface* s, const Topology& t, const Clusters& c, double chi, size_t leaf_pairs,
             double s_exp, const DeviceTree* dt = nullptr);
  // The same from a host plan.
  void build(ddgrs_surface* s, const FaceMetricPlan& p, double s_exp);
  // y = (w_hi B̃ + w_lo B̃⁰) x for one interleaved 3 nv FP64 vector (device pointers).
  void apply(ddgrs_ctx* ctx, const double* x, double* y);
};

// Host part of a VertexFractionalDev: vertex BVH (points f_v, masses a_v), block structure, panels.
struct VertexFractionalPlan {
  Topology topo;
  Clusters clusters;
  BlockStructure bs;
  BlockTreePlan bt;
  std::vector<double> pos, mass;
  double tau = 0.0;
};
// `topo`: the vertex BVH (points f_v, leaf size as given) of the same positions.
VertexFractionalPlan plan_vertex_fractional(const std::vector<double>& pos, const std::vector<double>& mass,
                                            double tau, double chi, size_t leaf_pairs, Topology topo);

// L̃^V_τ over vertex pairs (hierarchical.rs:1003).
struct VertexFractionalDev {
  Topology topo;
  BlockStats stats;
  BlockTreeDev bt;
  int nv = 0;
  DevBuf<double> mass;      // per vertex (FP64)
  DevBuf<float> u_hl, z_hl; // product scratch: 3 channels, hi/lo
  DevBuf<double> node_cen, node_mass, pos_dev;  // the build's cluster data and positions
  bool extended = false;    // products with the extended near field (BlockTreeDev::multiply)
  void build(ddgrs_ctx* ctx, const VertexFractionalPlan& p);
  // The same with the structure and panels built on the device; `topo` the vertex BVH of `pos`.
  void build(ddgrs_ctx* ctx, const std::vector<double>& pos, const std::vector<double>& mass, double tau, double chi,
             size_t leaf_pairs, Topology topo);
  // y = L̃^V x (interleaved 3 nv FP64, device pointers).
  void apply(ddgrs_ctx* ctx, const double* x, double* y);
};

}  // namespace ddgrs
