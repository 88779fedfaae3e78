// The block cluster tree on the device and the hierarchical products built on it:
// B̃, B̃⁰ over face pairs (RS/src/metric/hierarchical.rs:283 FaceMetricTree) and the vertex-pair
// fractional operator L̃^V_{2−s} (hierarchical.rs:1003 VertexFractionalTree).
#pragma once
#include "kernels.hpp"
#include "trees.hpp"

namespace ddgrs {

// Device copy of a block structure over a BVH with NK kernels (values FP32, NK per entry).
struct BlockTreeDev {
  int nk = 1;
  int n = 0, nn = 0, nl = 0;
  uint64_t near_total = 0;
  DevBuf<uint32_t> order, position, start, end, leaves, chain_ptr, chain;
  DevBuf<uint32_t> far_ptr, far_col, far_row, near_ptr, near_col, near_row, needed;
  DevBuf<uint64_t> near_off;
  DevBuf<float> far_coef, near_val;  // nk per entry
  int n_needed = 0;
  uint32_t n_far = 0, n_near_blocks = 0;
  // scratch of a product
  DevBuf<float> up, farv, hx;

  void upload(ddgrs_ctx* ctx, const Topology& t, const BlockStructure& s, int nk);
  // hx = H̃ ψ (both in BVH order, C channels per element; channels [0, H) on kernel 0, [H, C) on
  // kernel nk − 1).  Results stay in `hx`.
  template <int C, int H>
  void multiply(ddgrs_ctx* ctx, const float* psi);
};

// B̃, B̃⁰ of one configuration (hierarchical.rs:283).
struct FaceMetricDev {
  Topology topo;
  BlockStructure bs;
  BlockTreeDev bt;
  int nv = 0, nf = 0;
  DevBuf<float> df;        // 9 per BVH position: g_c(S) for c = 0..2
  DevBuf<float> area_p;    // per BVH position
  DevBuf<uint32_t> vert_p; // 3 per BVH position
  DevBuf<float> ha;        // 2 per BVH position (H̃a for hi, lo)
  DevBuf<float> m, psi, z; // product scratch, 12 channels
  DevBuf<uint32_t> inc_ptr, inc;  // vertex → (face, corner), with faces' BVH positions
  double w_hi = 1.0, w_lo = 1.0;

  // Builds on the face tree of `bh` (same positions) with separation χ and leaf_pairs.
  void build(ddgrs_surface* s, const Topology& t, const Clusters& c, double chi, size_t leaf_pairs, double s_exp);
  // y = (w_hi B̃ + w_lo B̃⁰) x for one interleaved 3 nv FP64 vector (device pointers).
  void apply(ddgrs_ctx* ctx, const double* x, double* y);
};

// L̃^V_τ over vertex pairs (hierarchical.rs:1003).
struct VertexFractionalDev {
  Topology topo;
  Clusters clusters;
  BlockStructure bs;
  BlockTreeDev bt;
  int nv = 0;
  DevBuf<float> mass_p, ha;  // per BVH position
  DevBuf<float> psi;
  DevBuf<double> mass;       // per vertex (FP64)
  void build(ddgrs_ctx* ctx, const std::vector<double>& pos, const std::vector<double>& mass, double tau, double chi,
             size_t leaf_size, size_t leaf_pairs);
  // y = L̃^V x (interleaved 3 nv FP64, device pointers).
  void apply(ddgrs_ctx* ctx, const double* x, double* y);
};

}  // namespace ddgrs
