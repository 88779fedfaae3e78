// The device Hs descent session (RS/src/step/session.rs): positions resident in HBM, one step =
// energy and differential, the saddle solve for the direction, the Armijo line search and the
// barycenter projection.
#pragma once
#include <future>
#include <string>
#include <vector>

#include <cmath>

#include "bh.hpp"
#include "ccd.hpp"
#include "constraints.hpp"
#include "gmres.hpp"
#include "learned.hpp"
#include "lprime_gpu.hpp"

namespace ddgrs {

struct StepOptions {
  double alpha = 6.0, beta = 12.0;
  bool energy_exact = false;     // all pairs instead of Barnes–Hut
  double theta = 0.5;            // Barnes–Hut opening parameter
  size_t leaf_size = 8;          // face and vertex BVH leaves
  double chi = 0.5;              // block cluster tree separation
  size_t leaf_pairs = 64;        // dense block threshold
  double w_tp = 1.0, c_lo = 1.0; // metric weights of B̃ and B̃⁰
  double eps_rel = 1e-10;        // ε of K = W + ε M relative to trW / trM
  int pc_form = 0;               // preconditioner form: 0 augmented, 1 congruence
  bool pc_unit_scale = false;    // L̃^V unscaled instead of the planar symbol's scale
  bool pc_refine = false;        // one step of iterative refinement per L′ solve
  size_t restart = 50;
  double tolerance = 1e-6;
  size_t max_iterations = 200;
  double armijo_c = 1e-2;
  double tau_min_rel = 1e-12;    // τ_min = tau_min_rel · ℓ̄₀ / ‖x‖_∞
  long max_halvings = -1;        // no cap when negative
  int start_kind = DDGRS_START_GROW_PREVIOUS;
  double start_value = 1.0;      // τ (absolute) or the lengths of MaxDisplacement
  bool robust = true;            // the recovery chain of a diverged solve
  long retry_max_iterations = 0; // its GMRES retry's cap (0: max(4 max_iterations, 800))
  int64_t dense_max_vertices = 6000;  // its dense LU's largest mesh
  double diverged_residual = 1e-1;    // an unconverged solve above it has diverged
  // nonlinear rows after the barycenter's (kRowArea, kRowVolume), targets (NaN: at creation), Newton
  int rows[2] = {0, 0};
  double row_targets[2] = {NAN, NAN};
  bool newton_fixed = false;
  int newton_count = 5;
  bool ccd = false;               // the CCD cap of τ₀ (with the exempt band)
  CcdOpts ccd_opts;
};

// TrialOutcome (RS/src/step/session.rs): 0 accepted, 1 Armijo failed, 2 non-finite.
struct TrialRecord {
  double tau, energy;
  int32_t outcome;
};

using StepReport = ddgrs_step_report;


struct Session {
  ddgrs_surface* s = nullptr;
  StepOptions opts;
  Exps exps{};
  double l0 = 0.0, last_tau = 0.0;
  double target[3] = {0.0, 0.0, 0.0};
  BhEval bh;
  GpuBvhBuilder face_bvh, vert_bvh;
  FaceMetricDev fm;
  Preconditioner pc;
  CcdDev ccd;              // the CCD cap of the line search (edges set at init when it is on)
  std::future<LprimeSymbolic> lp_symbolic;  // L′'s symbolic analysis, started at init
  bool lp_ready = false;
  ddgrs_ctx* aux = nullptr;   // second stream: vertex BVH, L̃^V and L′'s factorization of a step
  ddgrs_ctx* aux2 = nullptr;  // third stream: the face metric's block tree of a step
  SaddleOp op;
  DeviceVecs vec;
  GmresResult last_gmres;
  DevBuf<double> g, sol, p0, partials, rows, scal, mwork, xpart;
  // the learned preconditioner: mode 1 records pairs, 2 solves with (I + Q) P⁻¹ (plain on a miss)
  LearnedModel* learned = nullptr;
  int learned_mode = 0;
  DevBuf<double> pg, kr, kz;  // P⁻¹ g and the Krylov vectors of recorded pairs
  std::vector<TrialRecord> trials;  // the last step's line search
  std::vector<std::string> fallbacks;  // the last step's escalations (LinearReport::fallbacks)
  // the nonlinear constraint rows, their targets, the last step's Schur complement and Φ
  std::vector<int> nl_rows;
  std::vector<double> nl_targets;
  SchurRows schur;
  std::vector<double> last_phi_before, last_phi;
  DevBuf<double> zrhs, zsol;
  size_t last_iterations = 0;  // iterations of the last converged solve
  int learned_misses = 0;        // consecutive (I + Q) P⁻¹ solves that fell back to the plain one
  static constexpr int kLearnedMissLimit = 2;  // after that many, the session solves plain

  Session() = default;
  Session(const Session&) = delete;
  Session& operator=(const Session&) = delete;
  ~Session();
  // ℓ̄₀ and the barycenter target at the surface's current positions (session.rs:412 new).
  void init(ddgrs_surface* s, const StepOptions& o);
  void step(StepReport* rep);

 private:
  double read_scalar(const double* d);
  double energy_value();
  void project();
  // the projection with nonlinear rows: the barycenter translation and the Newton rounds (host)
  void project_constrained(StepReport* rep);
};

}  // namespace ddgrs
