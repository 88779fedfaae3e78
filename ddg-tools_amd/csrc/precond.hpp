// The reference preconditioner on the device (RS/src/solve/precond.rs): L′ = [[K, c₁ᵀ],[c₁, 0]]
// per axis with K = W + εM and the global barycenter row c₁ = a/Σa, factored once per step on the
// host (sparse LDLᵀ in a nested-dissection order), solved on the device in FP64; the augmented
// form z = L′⁻¹ v + y, y = L′⁻¹ r, v = s L̃^V y − K y (precond.rs:1026 apply_fine).
#pragma once
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
  std::vector<uint32_t> parent;               // elimination tree, UINT32_MAX for a root
  std::vector<uint32_t> level;                // height of each node above the leaves
  uint32_t n_levels = 0;
};

// Factor of the symmetric matrix `a` (both triangles stored) in the elimination order `order`
// (up-looking LDLᵀ: T. Davis, "Algorithm 849", ACM TOMS 31(4), 2005).  Throws SingularSaddle on a
// zero or non-finite pivot.
Factor factor_ldlt(const HostCsr& a, const std::vector<uint32_t>& order);
// Nested dissection of the graph of `k` by coordinate bisection of `pos` (domains before their
// separators).
std::vector<uint32_t> nested_dissection_order(const HostCsr& k, const std::vector<double>& pos);

struct LprimeDev {
  size_t n = 0, nv = 0;
  Factor f;
  DevBuf<uint32_t> perm, rp, rj, cp, ci, wide_rows, narrow_rows;
  DevBuf<double> rx, cx, d;
  std::vector<uint32_t> wide_off;    // levels [0, n_wide): row ranges in wide_rows (levels bottom-up)
  std::vector<uint32_t> narrow_off;  // levels n_wide..: row ranges in narrow_rows
  DevBuf<uint32_t> narrow_off_d;
  DevBuf<double> xb;                 // n × 3 work array (permuted numbering)
  void upload(ddgrs_ctx* ctx);
  // x = L′⁻¹ r for a full saddle vector (3 nv + 3, interleaved positions then multipliers).
  void solve(ddgrs_ctx* ctx, const double* r, double* x);
};

struct Preconditioner {
  size_t nv = 0, n = 0;
  double eps = 0.0, frac_scale = 1.0;
  LprimeDev lp;
  VertexFractionalDev vf;
  DevBuf<uint32_t> kp, kj;  // K (CSR, symmetric)
  DevBuf<double> kx;
  DevBuf<double> y, v;      // work vectors (n)
  std::vector<double> crow; // barycenter row a_v / Σ a
  DevBuf<double> crow_d;
  double factor_seconds = 0.0;

  // Builds K = W + ε M, L′ and L̃^V_{2−s} at the surface's current positions.
  void build(ddgrs_surface* s, double s_exp, double w_tp, double eps_rel, double chi, size_t leaf_size, size_t leaf_pairs);
  // z = P⁻¹ r (device pointers, length 3 nv + 3).
  void apply(ddgrs_ctx* ctx, const double* r, double* z);
};

}  // namespace ddgrs
