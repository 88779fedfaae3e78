// L′ factored on the device.  The sparsity of L′ = [[K, c], [cᵀ, 0]] is fixed by the mesh's
// connectivity, so the nested-dissection order, the separator tree, the fronts of the multifrontal
// method and every index map are computed once (LprimeSymbolic, host) and reused while the positions
// change; each step then computes K's values, assembles the fronts, factors them height by height
// and forms L's rows and columns and the dense block inverses that the block solves use, all on the
// device (LprimeFactorDev).  The numbers are those of factor_multifrontal up to rounding (the
// same eliminations; products summed in another order).
#pragma once
#include "precond.hpp"

namespace ddgrs {

// Columns of a panel (blocked elimination of the large fronts).
constexpr uint32_t kPanelCols = 32;
// Rows and columns of an update tile.
constexpr uint32_t kTile = 64;
// Rows below a panel's diagonal block solved by one workgroup (one row per thread).
constexpr uint32_t kPanelItemRows = 128;

struct FrontItem {  // rows [r0, r1) of a child's update matrix, added into its parent's front
  uint32_t child, r0, r1;
};
struct FrontTile {  // trailing-update tile of a large front: rows from i0, columns from j0
  uint32_t block, i0, j0;
};
struct FrontPanel {  // a panel's rows [c1 + r0, c1 + r0 + kPanelItemRows) below its diagonal block
  uint32_t block, r0;
};

struct LprimeSymbolic {
  size_t nv = 0, n = 0;
  HostCsr kpat;                   // K's pattern: each vertex and its neighbours, ascending (no values)
  std::vector<uint32_t> kdiag;    // per vertex: index of the diagonal entry in kpat
  NdBlocks nd;
  Factor pattern;                 // perm, rows and columns of L (every front entry; no values)
  LprimePlan plan;                // block solve schedule (no inverses)
  // blocks of the separator tree, then the multiplier row (the root)
  uint32_t nb = 0, H = 0, max_front = 0, max_rows = 0;
  std::vector<uint32_t> bstart, bk, bf, bparent, bheight, child_ptr, child;
  std::vector<uint64_t> foff;     // per block: first entry of its front (f × f row-major; lower part used)
  uint64_t front_total = 0;
  std::vector<uint64_t> amap_k;   // per K entry: front position of its lower-triangle copy, or kNoPos
  std::vector<uint64_t> amap_c;   // per vertex: front position of (multiplier row, vertex)
  std::vector<uint32_t> ea_off, ea_map;  // per block: its update rows' positions in the parent's front
  std::vector<uint64_t> gmap_r, gmap_c, dpos;  // front positions of L's rows, L's columns and D
  // Fronts that fit one workgroup's LDS are factored there (one launch per height); the others by
  // panels of kPanelCols columns and tiled trailing updates.
  std::vector<uint32_t> lds_blocks, lds_off;   // per height
  std::vector<uint32_t> lds_rows;              // per height: the largest such front
  std::vector<FrontItem> ea_items;             // large fronts' extend-add work
  std::vector<uint32_t> ea_item_off;           // per (height, child slot): H × max_children + 1
  uint32_t max_children = 0;
  std::vector<FrontPanel> panels;
  std::vector<uint32_t> tile_off, panel_off, step_c0, step_off;  // step_off: per height
  std::vector<FrontTile> tiles;
  static constexpr uint64_t kNoPos = ~0ull;
};
// `lds_bytes`: the LDS a workgroup may use (fronts whose lower triangle and one column fit are
// factored in LDS).
LprimeSymbolic lprime_symbolic(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                               size_t lds_bytes);

struct LprimeFactorDev {
  size_t nv = 0, n = 0;
  uint32_t H = 0, max_children = 0, max_rows = 0;
  uint64_t front_total = 0;
  std::vector<uint32_t> lds_off, lds_rows, ea_item_off, step_off, step_c0, panel_off, tile_off;
  DevBuf<uint32_t> kp, kj, kdiag, bstart, bk, bf, bparent, child_ptr, child, ea_off, ea_map, lds_blocks;
  DevBuf<FrontPanel> panels;
  DevBuf<uint64_t> foff, amap_k, amap_c, gmap_r, gmap_c, dpos;
  DevBuf<FrontItem> ea_items;
  DevBuf<FrontTile> tiles;
  DevBuf<double> front, kx, mass, crow, diagv, sums;  // sums: tr W, Σ a, ε
  DevBuf<uint32_t> bad;
  uint32_t* bad_host = nullptr;  // pinned: the first failing pivot of the last factorization
  size_t nnz_l = 0, nnz_k = 0;
  LprimeFactorDev() = default;
  LprimeFactorDev(const LprimeFactorDev&) = delete;
  LprimeFactorDev& operator=(const LprimeFactorDev&) = delete;
  ~LprimeFactorDev();
  // Uploads the symbolic structure and sets up `lp`'s schedule and pattern.
  void setup(ddgrs_ctx* ctx, const LprimeSymbolic& sym, LprimeDev& lp);
  // Enqueues K = W + εM at the surface's positions (K's values into kx, the barycenter row into
  // crow), the factorization of L′ and L's rows, columns, D and block inverses into `lp`.
  void factor(ddgrs_surface* s, double eps_rel, LprimeDev& lp);
  // After a synchronization: throws SingularSaddle if a pivot of the last factorization was zero
  // or not finite.
  void check(ddgrs_ctx* ctx);
  double eps(ddgrs_ctx* ctx);
};

}  // namespace ddgrs
