// The saddle operator A′ and restarted right-preconditioned GMRES on the device.
#pragma once
#include <functional>

#include "precond.hpp"

namespace ddgrs {

// A′ [x; λ] = [w_TP (B̃ + c_lo B̃⁰) x + C_Qᵀ λ; C_Q x] with the global barycenter rows
// (RS/src/solve/iterative.rs:284 HierarchicalOperator::apply_batch).
struct SaddleOp {
  FaceMetricDev* fm = nullptr;
  const double* crow = nullptr;  // device, nv
  size_t nv = 0;
  DevBuf<double> part;           // reduction partials
  void apply(ddgrs_ctx* ctx, const double* x, double* y);
};

struct GmresSettings {
  size_t restart = 50;
  double tolerance = 1e-6;
  size_t max_iterations = 200;
};

struct GmresResult {
  size_t iterations = 0, cycles = 0, operator_applications = 0, preconditioner_applications = 0;
  bool converged = false;
  bool extended = false;  // on_stall switched the products
  double relative_residual = 0.0;
  std::vector<double> history, cycle_residuals;
};

// FP64 vector operations with fixed reduction trees (run-to-run reproducible).
struct DeviceVecs {
  size_t n = 0;
  DevBuf<double> partial, scalars;
  DevBuf<double> Vb, w, t, r, small;  // GMRES workspace: basis, work vectors, Hessenberg and rotations
  void init(size_t n);
  // out[k] = V_k · w for k < count (V rows of length n, stride n); out on the device
  void dots(ddgrs_ctx* ctx, const double* V, size_t count, const double* w, double* out);
  // w −= Σ_k h[k] V_k
  void subtract(ddgrs_ctx* ctx, const double* V, size_t count, const double* h, double* w);
  double norm_host(ddgrs_ctx* ctx, const double* x);
};

// Right-preconditioned restarted GMRES (RS/src/solve/gmres.rs:311-583: x₀ = 0, explicit residual
// at every cycle start, convergence on the explicit residual).  The Arnoldi orthogonalisation is
// classical Gram–Schmidt applied twice (two fixed-order GEMV passes per pass instead of j dependent
// reductions), the Givens rotations and the Hessenberg back substitution run on the device; all in
// FP64.  Vectors of length n are device pointers.
//
// on_stall (optional) is called once when the explicit residual stays above the tolerance although
// the Arnoldi estimate of the cycle before reached it, or a whole cycle reduced it less than 4×
// within 100× of the tolerance: the operator's rounding is the floor.  If it returns true (it has switched op and prec to more
// exact products), the solve continues from the current x with a new explicit residual.
GmresResult gmres_device(ddgrs_ctx* ctx, size_t n, const std::function<void(const double*, double*)>& op,
                         const std::function<void(const double*, double*)>& prec, const double* b, double* x,
                         const GmresSettings& o, DeviceVecs& vec,
                         const std::function<bool()>& on_stall = std::function<bool()>());

}  // namespace ddgrs
