// L′ factored on the device.  The sparsity of L′ = [[K, c], [cᵀ, 0]] is fixed by the mesh's
// connectivity, so the nested-dissection order, the separator tree, the fronts of the multifrontal
// method and every index map are computed once (LprimeSymbolic, host) and reused while the positions
// change; each step then computes K's values, assembles the fronts, factors them height by height
// and forms the dense inverses of the diagonal blocks, all on the device (LprimeFactorDev), which
// also solves with the factored fronts.  The numbers are those of
// factor_multifrontal up to rounding (the same eliminations; products summed in another order).
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
// Solves of heights with a large front: rows per work item (four per wave).
constexpr uint32_t kSolveRows = 16;
// LDS of a small-front solve workgroup (its matrices are staged there).
constexpr size_t kSolveSmallLds = 64 * 1024;
// Subtree solves of the lowest heights: waves per workgroup (one front per wave at a time) and the
// largest LDS slice of a wave, in doubles (w or x_S and y_J or t: 3 f + 3 k).
constexpr uint32_t kSubWaves = 8;
constexpr uint32_t kSubWaveDoubles = 1024;
constexpr uint32_t kSubLevels = 3;
struct SolveItem {  // rows [r0, r0 + kSolveRows) of a block (or of its later front rows)
  uint32_t block, r0;
};
// Tasks of the one-launch solve: a small block's forward or backward, or one SolveItem of a large
// block's forward y_J / u_J or backward t / x_J phase.
enum SolveTaskKind : uint32_t { kTaskFwdSmall, kTaskFwdY, kTaskFwdU, kTaskBwdSmall, kTaskBwdT, kTaskBwdX };
struct SolveTask {
  uint32_t kind, block, r0;
};
struct TaskCounts {  // per block: forward tasks (of them y tasks), backward tasks (of them t tasks)
  uint32_t nf, ny, nb, nt;
};

struct LprimeSymbolic {
  size_t nv = 0, n = 0;
  HostCsr kpat;                   // K's pattern: each vertex and its neighbours, ascending (no values)
  std::vector<uint32_t> kdiag;    // per vertex: index of the diagonal entry in kpat
  NdBlocks nd;
  std::vector<uint32_t> perm;     // the elimination order (perm[new] = old), the multiplier last
  std::vector<uint32_t> block_of; // per row of the order: its block
  // blocks of the separator tree, then the multiplier row (the root)
  uint32_t nb = 0, H = 0, max_front = 0, max_rows = 0;
  std::vector<uint32_t> bstart, bk, bf, bparent, bheight, child_ptr, child;
  std::vector<uint64_t> foff;     // per block: first entry of its front (f × f row-major; lower part used)
  uint64_t front_total = 0;
  std::vector<uint64_t> amap_k;   // per K entry: front position of its lower-triangle copy, or kNoPos
  std::vector<uint64_t> amap_c;   // per vertex: front position of (multiplier row, vertex)
  std::vector<uint32_t> ea_off, ea_map;  // per block: its update rows' positions in the parent's front
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
  // solves: the blocks with a front, by height; each block's later front rows (global indices,
  // at ea_off); the offsets of the dense inverses of the diagonal blocks (k × k row-major)
  std::vector<uint32_t> hblocks, hoff, srows;
  std::vector<uint64_t> moff;
  std::vector<uint8_t> hsplit;                 // per height: solved by the large-front path
  std::vector<uint32_t> hlds;                  // per height: LDS of the small-front path
  std::vector<SolveItem> krows, urows;         // large-front heights' work items
  std::vector<uint32_t> krows_off, urows_off;  // per height
  std::vector<uint64_t> ltoff;                 // per block: its L_SJᵀ (large-front heights only)
  std::vector<FrontTile> lt_tiles;             // transpose tiles (block, first later row, first column)
  // Heights [0, sub_levels) are solved by one launch per phase, one workgroup per subtree of those
  // heights (rooted at a block whose parent is higher, or at the root), level by level inside it.
  // sub_fronts: each subtree's blocks, level-major; the range of subtree s at level h is
  // [sub_lvl_off[s·sub_levels + h], sub_lvl_off[s·sub_levels + h + 1]).
  uint32_t sub_levels = 0, sub_wave_doubles = 0;
  std::vector<uint32_t> sub_fronts, sub_lvl_off;
  static constexpr uint64_t kNoPos = ~0ull;
};
// `lds_bytes`: the LDS a workgroup may use (fronts whose lower triangle and one column fit are
// factored in LDS).
LprimeSymbolic lprime_symbolic(const std::vector<uint32_t>& faces, size_t nv, const std::vector<double>& pos,
                               size_t lds_bytes);

// The one-launch solve's tasks in ticket order (forward by height, a large block's y tasks before
// its u tasks; backward from the root down, t tasks before x tasks) and the per-block task counts.
void solve_tasks(const LprimeSymbolic& S, std::vector<SolveTask>& tasks, std::vector<TaskCounts>& counts);

struct LprimeFactorDev {
  size_t nv = 0, n = 0;
  uint32_t H = 0, max_children = 0, max_rows = 0, max_front = 0;
  uint64_t front_total = 0;
  std::vector<uint32_t> lds_off, lds_rows, ea_item_off, step_off, step_c0, panel_off, tile_off, hoff;
  std::vector<uint32_t> krows_off, urows_off, hlds;
  std::vector<uint8_t> hsplit;
  DevBuf<uint32_t> kp, kj, kdiag, bstart, bk, bf, bparent, child_ptr, child, ea_off, ea_map, lds_blocks;
  DevBuf<FrontPanel> panels;
  DevBuf<uint32_t> perm, block_of, hblocks, srows;
  DevBuf<uint64_t> moff;
  DevBuf<double> minv, minv_t, ubuf, yb, xb, tb;  // block inverses M_J, M_Jᵀ; update vectors, permuted vectors
  DevBuf<SolveItem> krows, urows;
  DevBuf<uint64_t> ltoff;
  DevBuf<double> lt;
  DevBuf<double> pj;  // large fronts: P_J = L_SJ M_J ((f − k) × k), at ltoff
  DevBuf<FrontTile> lt_tiles;
  DevBuf<uint64_t> foff, amap_k, amap_c;
  DevBuf<FrontItem> ea_items;
  DevBuf<FrontTile> tiles;
  DevBuf<double> front, kx, mass, crow, diagv, sums;  // sums: tr W, Σ a, ε
  DevBuf<uint32_t> bad;
  uint32_t* bad_host = nullptr;  // pinned: the first failing pivot of the last factorization
  // one-launch solve: tasks in ticket order, per-block task counts, counters (forward and backward
  // tasks done per block, then the ticket), the failure flag
  DevBuf<SolveTask> tasks;
  DevBuf<TaskCounts> task_counts;
  DevBuf<uint32_t> task_done, solve_failed;
  DevBuf<uint32_t> sub_fronts, sub_lvl_off;
  uint32_t sub_levels = 0, sub_wave_doubles = 0, nsub = 0;
  uint32_t nb = 0, ntasks = 0;
  size_t one_lds = 0;   // its LDS per workgroup
  bool one_ok = false;  // the one-launch solve fits the device (else the per-height launches)
  size_t nnz_k = 0;
  uint32_t n_lt_tiles = 0;
  LprimeFactorDev() = default;
  LprimeFactorDev(const LprimeFactorDev&) = delete;
  LprimeFactorDev& operator=(const LprimeFactorDev&) = delete;
  ~LprimeFactorDev();
  // Uploads the symbolic structure.
  void setup(ddgrs_ctx* ctx, const LprimeSymbolic& sym);
  // Enqueues on ctx's stream K = W + εM at the surface's positions (K's values into kx, the
  // barycenter row into crow), the factorization of L′ (in the fronts) and the dense inverses of
  // its diagonal blocks.
  void factor(ddgrs_ctx* ctx, ddgrs_surface* s, double eps_rel);
  // x = L′⁻¹ r for a full saddle vector (3 nv + 3, interleaved positions then multipliers), the
  // three axes together: forward by height (each front's right-hand side from its block and its
  // children's update vectors, its block solved by the dense inverse, its update vector formed),
  // D⁻¹, backward by height (each block from its later front rows); the lowest heights in one
  // launch per phase (a workgroup per subtree), the others one launch per height and phase.  DDGRS_LPRIME_SOLVE_ONE_LAUNCH=1 asks for a single launch whose workgroups take the
  // blocks' tasks in dependency order (the same arithmetic per task; not yet validated on the
  // device, so off by default).
  void solve(ddgrs_ctx* ctx, const double* r, double* x);
  // After a synchronization: throws SingularSaddle if a pivot of the last factorization was zero
  // or not finite.
  void check(ddgrs_ctx* ctx);
  double eps(ddgrs_ctx* ctx);
};

}  // namespace ddgrs
