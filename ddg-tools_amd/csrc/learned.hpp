// The learned right preconditioner layered on the plain one: a mesh network maps per-vertex
// geometry features to per-vertex 3×3 blocks B_v and per-edge 3×3 blocks E_uv, and
//   (Q x)_v = s (B_v x_v + Σ_{u ∈ N(v)} E_uv x_u)
// corrects the plain preconditioner's output: z = (I + Q) P⁻¹ r, so GMRES stays a Krylov method
// with a linear preconditioner.  Q is trained on pairs (r, y = A′⁻¹ r) from plain solves along
// descent trajectories: with x = c P⁻¹ r, c = ⟨y, P⁻¹ r⟩ / ‖P⁻¹ r‖² (a scalar leaves the Krylov
// spaces unchanged), it reduces ‖x + Q x − y‖² / ‖y − x‖²; it starts at zero (the output layer's
// weights are zero), where the combined preconditioner is the plain one.
//
// Network: h₀ = relu(F W₀ + b₀) (8 features → 64); four message-passing layers
// h_{l+1} = relu(h_l W_s + m_l W_n + b), m_l = the mean of h_l over the vertex's neighbours; the
// vertex head B_v = h₄ W_v + b_v (64 → 9); the edge head per undirected edge {u, v}:
// t = relu((h_u + h_v) W_e1 + b_e1) (64 → 64), E_uv = t W_e2 + b_e2 (64 → 9), symmetric in u and v,
// so Q is symmetric when every B_v is.  The products with 64 output columns run on BF16 matrix
// cores with FP32 accumulation, the 9-output heads on the vector units in FP32.  Training: forward,
// backward and Adam as kernels on the device; gradients are checked against a NumPy restatement.
#pragma once
#include <vector>

#include "kernels.hpp"

namespace ddgrs {

constexpr int kLnFeat = 8, kLnWidth = 64, kLnLayers = 4, kLnOut = 9;

// Offsets of the parameters in one FP32 array.
struct LearnedLayout {
  size_t w0, b0, w[kLnLayers], b[kLnLayers], wv, bv, we1, be1, we2, be2, total;  // w[l]: [W_s; W_n] (128 × 64)
};
LearnedLayout learned_layout();

struct LearnedModel {
  int nv = 0;
  std::vector<uint32_t> faces;          // the mesh (3 per face)
  std::vector<uint32_t> nbr_ptr_h, nbr_h;  // neighbours (mesh edges), ascending
  DevBuf<uint32_t> nbr_ptr, nbr;
  // undirected edges {u < v} in ascending order, and per neighbour entry its edge
  int ne = 0;
  std::vector<uint32_t> edge_h, emap_h;  // 2 per edge; 1 per neighbour entry
  DevBuf<uint32_t> edge, emap;
  DevBuf<float> e_in, e_pre, e_act, e_out, e_dout, e_d;  // edge head: ne × 64 (e_out, e_dout: ne × 9)
  DevBuf<float> params, grad, adam_m, adam_v, gpart, best_params;
  DevBuf<float> feat, act, pre, mean, out, dout, dh, dtmp, qd;  // activations and their gradients
  DevBuf<double> lsum;                   // loss partial sums
  // training pairs (features, x = c P⁻¹ r, e = y − x: FP32 per vertex) and ‖e‖²
  DevBuf<float> pair_feat, pair_r, pair_e;
  std::vector<double> pair_e2, pair_ratio;
  int npairs = 0, cap_pairs = 0;
  float scale = 0.f;  // s (set from the pairs at the first training)
  int krylov_pairs = 3;  // pairs (A′ z, z) a recording step adds after its (g, A′⁻¹ g)
  int64_t adam_t = 0;
  bool blocks_ready = false;  // `out` holds the blocks of the last forward

  void init(ddgrs_ctx* ctx, const std::vector<uint32_t>& faces, int nv, uint64_t seed);
  // Features at positions (3 nv FP64, host) into `dst` (device, nv × 8).
  void features(ddgrs_ctx* ctx, const double* pos, float* dst);
  // Blocks at the given features (device): B_v into `out` (nv × 9), E_uv into `e_out` (ne × 9).
  void forward(ddgrs_ctx* ctx, const float* f);
  void add_pair(ddgrs_ctx* ctx, const double* pos, const double* r, const double* pr, const double* y);
  // Loss of pair p at the current parameters (read: returned, after a synchronization; else 0)
  // and (add_grad) its gradient added into `grad`.
  double loss_and_grad(ddgrs_ctx* ctx, int p, bool add_grad, bool read = true);
  // Adam over the pairs (minibatches of `batch`, `iters` updates), keeping the parameters of the
  // lowest mean loss over all pairs (checked before training, once per epoch and at the end);
  // returns the mean loss over all pairs before and of the kept parameters.
  void train(ddgrs_ctx* ctx, int iters, int batch, double lr, double* before, double* after);
  // z += Q x for a full saddle vector x (3 nv + 3, device; the blocks of the last forward).
  void apply(ddgrs_ctx* ctx, const double* r, double* z);
};

// The model of a C-ABI handle.
LearnedModel* learned_model(ddgrs_learned* l);

}  // namespace ddgrs
