// The reference preconditioner on the device (RS/src/solve/precond.rs): L′ = [[K, c₁ᵀ],[c₁, 0]]
// per axis with K = W + εM and the global barycenter row c₁ = a/Σa, factored once per step on the
// host (sparse LDLᵀ in a nested-dissection order), solved on the device in FP64; the augmented
// form z = L′⁻¹ v + y, y = L′⁻¹ r, v = s L̃^V y − K y (precond.rs:1026 apply_fine).
#pragma once
#include <memory>

#include "blocktree.hpp"

namespace ddgrs {

struct HostCsr {
  size_t n = 0;
  std::vector<uint32_t> ptr, idx;
  std::vector<double> val;
};

// The factor P A Pᵀ = L D Lᵀ (L unit lower, stored by rows and by columns, strict part).
struct Factor {
  size_t n = 0;
  std::vector<uint32_t> perm;                 // perm[new] = old
  std::vector<uint32_t> rp, rj, cp, ci;       // rows of L (columns < row), columns of L (rows > col)
  std::vector<double> rx, cx, d;
};

// Factor of the symmetric matrix `a` (both triangles stored) in the elimination order `order`
// (up-looking LDLᵀ: T. Davis, "Algorithm 849", ACM TOMS 31(4), 2005).  Throws SingularSaddle on a
// zero or non-finite pivot.
Factor factor_ldlt(const HostCsr& a, const std::vector<uint32_t>& order);

struct NdBlocks;
// The same factor by the multifrontal method over the separator tree of `nd` (order nd.order, the
// multiplier row last): one dense front per block (the block's rows and the later rows coupled to
// them), children's update matrices added in, blocked right-looking LDLᵀ of the block's columns;
// the blocks of one height are independent.  Exact zeros are not stored.
Factor factor_multifrontal(const HostCsr& a, const NdBlocks& nd);

// A nested-dissection order and its separator tree: blocks are contiguous ranges of the order
// (leaf domains and separators), a separator after the two domains it splits; parent[b] is the
// separator that splits the domain b came from (UINT32_MAX for the root).
struct NdBlocks {
  std::vector<uint32_t> order;
  std::vector<uint32_t> start, end, parent;
};
// Nested dissection of the graph of `k` by coordinate bisection of `pos` (domains of at most 64
// vertices are leaves).
NdBlocks nested_dissection(const HostCsr& k, const std::vector<double>& pos);

// Device solves of L′ by blocks of the separator tree.  With the blocks of one height of the tree
// independent, a solve is one launch per height and direction: a block J computes
// r_J = b_J − L_{J,desc} x_desc (forward; its descendants are done) and x_J = L_JJ⁻¹ r_J with the
// dense inverse of its diagonal block; backward likewise with the ancestors and L_JJ⁻ᵀ.
// Host plan of the block solves: blocks (the separator tree's, then the multiplier row as the
// root), their heights, the off-block split points of every row and column of L, the dense
// inverses of the diagonal blocks (row-major), and the work items (block, first row) per height.
// A height whose blocks all fit one workgroup chunk (≤ 64 rows) is solved by one fused launch
// per direction (items = blocks); a height with a larger block is split into a residual launch
// and a dense launch, one wave per row each (rows of the height in split_rows).  The top height
// is the multiplier row alone: its forward residual spans every column and is summed by
// kLongParts workgroups, then reduced in a fixed order.
struct LprimePlan {
  std::vector<uint32_t> bstart, bend, bheight, rows, rsplit, csplit, item_block, item_row0, level_off;
  std::vector<uint32_t> block_of, split_rows, split_off;  // split_off[h]: rows of height h (empty if fused)
  std::vector<uint8_t> fused;
  std::vector<uint64_t> inv_off;
  std::vector<double> inv, inv_t;  // per block: M = L_JJ⁻¹ and Mᵀ, row-major
  uint32_t max_rows = 0;
};
// Throws if an off-block entry of L joins two blocks of the same height (the blocks of one height
// must be independent).
LprimePlan make_lprime_plan(const Factor& f, const NdBlocks& nd);
// The same without the inverses (only the pattern of `f` is read; inv_off is set).
LprimePlan make_lprime_schedule(const Factor& f, const NdBlocks& nd);
// The dense inverses (and transposes) of a schedule from the rows of `f`.
void host_block_inverses(const Factor& f, LprimePlan& p);

struct LprimeDev {
  size_t n = 0, nv = 0;
  std::vector<uint32_t> level_off;              // work items of each height: [level_off[h], level_off[h+1])
  std::vector<uint32_t> split_off;
  std::vector<uint8_t> fused;
  uint32_t max_rows = 0;
  DevBuf<uint32_t> perm, rp, rj, cp, ci, rsplit, csplit;
  DevBuf<double> rx, cx, d;
  DevBuf<uint32_t> item_block, item_row0, blk_start, blk_rows, block_of, split_rows;
  DevBuf<uint64_t> blk_inv;                     // offset of each block's dense inverse (row-major k × k)
  DevBuf<double> inv, inv_t;
  DevBuf<double> xb, xf, rbuf, parts;           // n × 3 work arrays (permuted numbering), partial sums
  double inverse_seconds = 0.0;
  // Block schedule and dense inverses of the blocks' diagonal parts of L.
  void upload(ddgrs_ctx* ctx, const Factor& f, const LprimePlan& p);
  // The schedule and the pattern of L only; rx, cx, d, inv and inv_t are sized for a device
  // factorization to fill.
  void setup(ddgrs_ctx* ctx, const Factor& pattern, const LprimePlan& p);
  // x = L′⁻¹ r for a full saddle vector (3 nv + 3, interleaved positions then multipliers).
  void solve(ddgrs_ctx* ctx, const double* r, double* x);
};

// Host part of the preconditioner at given positions: K (cotan + εM), the barycenter row, L′'s
// factor and block plan and the fractional scale.
struct PrecondPlan {
  size_t nv = 0;
  double eps = 0.0, frac_scale = 1.0;
  std::vector<double> crow;
  HostCsr k;
  Factor f;
  LprimePlan lp;
};
PrecondPlan plan_preconditioner(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                                const std::vector<double>& mass, double s_exp, double w_tp, double eps_rel);
// K = W + ε M with W the cotan stiffness (RS/src/geometry.rs:221), assembled row by row in the
// order of the reference's triplets (faces ascending, corners ascending); tr W into *trw.
HostCsr cotan_stiffness(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                        const std::vector<double>& mass, double eps_rel, double* eps);

struct LprimeSymbolic;
struct LprimeFactorDev;

enum { kFormAugmented = 0, kFormCongruence = 1 };

struct Preconditioner {
  size_t nv = 0, n = 0;
  double eps = 0.0, frac_scale = 1.0;
  LprimeDev lp;
  VertexFractionalDev vf;
  DevBuf<uint32_t> kp, kj;  // K (CSR, symmetric) of a host plan
  DevBuf<double> kx;
  DevBuf<double> y, v;      // work vectors (n)
  std::vector<double> crow; // barycenter row a_v / Σ a of a host plan
  DevBuf<double> crow_d;
  std::unique_ptr<LprimeFactorDev> fac;  // the device factorization (set up once per mesh)
  // What apply() reads: K and the barycenter row, from the host plan or the device factorization.
  const uint32_t *k_ptr = nullptr, *k_idx = nullptr;
  const double *k_val = nullptr, *crow_dev = nullptr;
  // PreconditionerOptions (precond.rs:466): the form (kFormAugmented: z = L′⁻¹[(s L^V − K) y; 0] + y,
  // kFormCongruence: z = L′⁻¹[L^V y; 0], with y = L′⁻¹ r), the fractional factor's scale (the planar
  // symbol 1/(w_TP 4 I(σ) I(2−s)), or 1 with unit_scale) and one step of iterative refinement in
  // every L′ solve (refine).
  int form = kFormAugmented;
  bool unit_scale = false, refine = false;
  DevBuf<double> res, dx, part;  // refinement work
  double scale() const { return unit_scale ? 1.0 : frac_scale; }

  Preconditioner();
  ~Preconditioner();
  // Builds K = W + ε M, L′ and L̃^V_{2−s} at the surface's current positions (L′ factored on the
  // device unless DDGRS_LPRIME_ON_HOST=1 asks for the host factorization).
  void build(ddgrs_surface* s, double s_exp, double w_tp, double eps_rel, double chi, size_t leaf_size, size_t leaf_pairs);
  void build(ddgrs_ctx* ctx, const PrecondPlan& p);
  // Device factorization: the symbolic structure (once per mesh), then per configuration K, L′'s
  // factor and block inverses (enqueued; check_factor() after the next synchronization).
  void setup_device(ddgrs_ctx* ctx, const LprimeSymbolic& sym);
  void factor_device(ddgrs_ctx* ctx, ddgrs_surface* s, double s_exp, double w_tp, double eps_rel);
  void check_factor(ddgrs_ctx* ctx);
  double eps_value(ddgrs_ctx* ctx);
  // L̃^V_{2−s} on the vertex BVH `vertex_tree` of `pos` (structure and panels on the device).
  void build_fractional(ddgrs_ctx* ctx, const std::vector<double>& pos, const std::vector<double>& mass, double tau,
                        double chi, size_t leaf_pairs, Topology vertex_tree);
  // x = L′⁻¹ r (device pointers, length 3 nv + 3).
  void lprime_solve(ddgrs_ctx* ctx, const double* r, double* x);
  // The same, followed by one step of iterative refinement when `refine` is set.
  void lprime_solve_refined(ddgrs_ctx* ctx, const double* r, double* x);
  // z = P⁻¹ r (device pointers, length 3 nv + 3).
  void apply(ddgrs_ctx* ctx, const double* r, double* z);
};

}  // namespace ddgrs
