// One Hs descent step on the device (RS/src/step/session.rs:749 step_inner, remeshing and CCD off):
// energy and differential (Barnes–Hut on frozen lists, or exact all pairs), the hierarchical
// saddle solver of the step (metric tree on the energy's face tree, iterative.rs:487-560), the
// direction by right-preconditioned GMRES, the Armijo line search with the reference's start rules
// and halving cap (session.rs:906, :958-1069, :1604), and the projection onto the global
// barycenter constraint with F at the projected positions.  Positions and every vector stay on the device; the host reads scalars
// (energies, the Armijo test, norms) and builds the tree structures.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>

#include "bh.hpp"
#include "capi_util.hpp"
#include "gmres.hpp"
#include "robust.hpp"
#include "session.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

// positions = p0 − τ x (x: the position block of the direction)
__global__ __launch_bounds__(256) void trial_kernel(size_t n, const double* __restrict__ p0, const double* __restrict__ x,
                                                    double tau, double* __restrict__ p) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = p0[i] - tau * x[i];
}

// a_v = Σ_{S∋v} a_S / 3 (faces ascending, RS/src/geometry.rs:194) and a_v f_v
__global__ __launch_bounds__(256) void vertex_mass_kernel(int64_t nv, const uint32_t* __restrict__ inc_ptr,
                                                          const uint32_t* __restrict__ inc, const double* __restrict__ area,
                                                          const double* __restrict__ pos, double* __restrict__ out) {
  int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  double a = 0.0;
  for (uint32_t q = inc_ptr[v]; q < inc_ptr[v + 1]; ++q) a += area[inc[q] >> 2] / 3.0;
  out[v] = a;
  for (int k = 0; k < 3; ++k) out[(size_t)nv * (1 + k) + v] = a * pos[3 * v + k];
}

__global__ __launch_bounds__(256) void translate_kernel(int64_t nv, double* __restrict__ pos, double sx, double sy,
                                                        double sz) {
  int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  pos[3 * v] += sx;
  pos[3 * v + 1] += sy;
  pos[3 * v + 2] += sz;
}

// Per block of 256 vertices: max over v of |x_v| (Euclidean), max of |x_i| (order-free maxima) and
// Σ x_i² (a fixed tree; the host adds the blocks in order).
__global__ __launch_bounds__(256) void xmax_kernel(int64_t nv, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double s0[256], s1[256], s2[256];
  int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double d = 0.0, m = 0.0, q = 0.0;
  if (v < nv) {
    double a = x[3 * v], b = x[3 * v + 1], c = x[3 * v + 2];
    q = a * a + b * b + c * c;
    d = sqrt(q);
    m = fmax(fabs(a), fmax(fabs(b), fabs(c)));
  }
  s0[threadIdx.x] = d;
  s1[threadIdx.x] = m;
  s2[threadIdx.x] = q;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
      s2[threadIdx.x] = s2[threadIdx.x] + s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s0[0];
    part[3 * blockIdx.x + 1] = s1[0];
    part[3 * blockIdx.x + 2] = s2[0];
  }
}

namespace {
double seconds_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}
}  // namespace

Session::~Session() {
  if (lp_symbolic.valid()) lp_symbolic.wait();
  if (aux) ctx_release(aux);
  if (aux2) ctx_release(aux2);
}

void Session::init(ddgrs_surface* s_, const StepOptions& o) {
  s = s_;
  if (!aux) aux = ctx_create_aux(s->ctx);
  if (!aux2) aux2 = ctx_create_aux(s->ctx);
  opts = o;
  exps = make_exps(o.alpha, o.beta);
  const size_t nv = (size_t)s->nv;
  std::vector<double> pos, mass;
  vertex_areas_host(s, &pos, &mass);
  // mean edge length ℓ̄₀ (RS/src/geometry.rs mean_edge_length) over the unique edges
  std::vector<std::pair<uint32_t, uint32_t>> e;
  for (int64_t f = 0; f < s->nf; ++f)
    for (int c = 0; c < 3; ++c) {
      uint32_t a = s->faces_h[3 * f + c], b = s->faces_h[3 * f + (c + 1) % 3];
      e.push_back({std::min(a, b), std::max(a, b)});
    }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  double sum = 0.0;
  for (auto [a, b] : e) {
    double d[3] = {pos[3 * a] - pos[3 * b], pos[3 * a + 1] - pos[3 * b + 1], pos[3 * a + 2] - pos[3 * b + 2]};
    sum += std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  }
  l0 = sum / (double)e.size();
  double tot = 0.0, bc[3] = {0.0, 0.0, 0.0};
  for (size_t v = 0; v < nv; ++v) {
    tot += mass[v];
    for (int k = 0; k < 3; ++k) bc[k] += mass[v] * pos[3 * v + k];
  }
  for (int k = 0; k < 3; ++k) target[k] = bc[k] / tot;
  last_tau = 0.0;
  if (opts.ccd) ccd.set_mesh(s->ctx, s->faces_h, (uint32_t)nv);
  // the nonlinear rows and their targets (the values at creation unless given)
  nl_rows.clear();
  nl_targets.clear();
  const HostMesh hm{s->faces_h.data(), s->nf, s->nv};
  for (int r = 0; r < 2; ++r)
    if (o.rows[r] == kRowArea || o.rows[r] == kRowVolume) {
      nl_rows.push_back(o.rows[r]);
      nl_targets.push_back(std::isnan(o.row_targets[r]) ? row_value(hm, pos.data(), o.rows[r]) : o.row_targets[r]);
    }
  last_phi_before.clear();
  last_phi.clear();
  // the mesh's L′ structure (fixed while only the positions change) on another thread while the
  // step's buffers are set up, then uploaded: per-mesh work, done once before the first step
  lp_ready = false;
  const size_t lds = s->ctx->prop.sharedMemPerBlock;
  lp_symbolic = std::async(std::launch::async, [this, nv, lds, p = std::move(pos)] {
    return lprime_symbolic(s->faces_h, nv, p, lds);
  });
  const size_t n = 3 * nv + 3;
  vec.init(n);
  g.reserve(n);
  sol.reserve(n);
  p0.reserve(3 * nv);
  partials.reserve(7 * (size_t)s->nf);
  rows.reserve((size_t)s->nf);
  scal.reserve(16);
  mwork.reserve(4 * nv);
  xpart.reserve(3 * (size_t)blocks_for((int64_t)nv, 256));
  HIPCHECK(hipSetDevice(aux->device));
  pc.setup_device(aux, lp_symbolic.get());
  lp_ready = true;
}

double Session::read_scalar(const double* d) {
  double* pin = (double*)s->ctx->staging(8);
  download(s->ctx, pin, d, 1);
  s->ctx->sync();
  return pin[0];
}

double Session::energy_value() {
  ddgrs_ctx* ctx = s->ctx;
  if (opts.energy_exact) {
    exact_pairs(ctx, s->nf, s->packs, s->area.ptr, exps, partials.ptr, rows.ptr);
    sum_f64(ctx, s->nf, rows.ptr, scal.ptr);
  } else {
    bh.energy(ctx, exps, scal.ptr);
  }
  return read_scalar(scal.ptr);
}

void Session::project() {
  ddgrs_ctx* ctx = s->ctx;
  const int64_t nv = s->nv;
  launch(ctx, "vertex_mass", vertex_mass_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)s->inc_ptr.ptr,
         (const uint32_t*)s->inc.ptr, (const double*)s->area.ptr, (const double*)s->positions.ptr, mwork.ptr);
  // Σ a_v and Σ a_v f_v as fixed-order sums
  for (int k = 0; k < 4; ++k) sum_f64(ctx, nv, mwork.ptr + (size_t)k * nv, scal.ptr + 4 + k);
  double* pin = (double*)ctx->staging(32);
  download(ctx, pin, (const double*)(scal.ptr + 4), 4);
  ctx->sync();
  double tot = pin[0];
  double shift[3];
  for (int k = 0; k < 3; ++k) shift[k] = target[k] - pin[1 + k] / tot;
  launch(ctx, "translate", translate_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, s->positions.ptr, shift[0],
         shift[1], shift[2]);
  compute_face_geometry(s, s->packs, false);
}

// schur.rs:495 project with nonlinear rows, on the host in FP64: the positional projection
// (constraints/mod.rs:657: translate by the barycenter's offset from its target, areas at the
// positions), then Newton rounds f ← f − Z N⁻¹ Φ_N, each followed by the positional projection and
// kept only if it lowers max_r |Φ_r| / tol_r over the kept rows; Φ before and after is kept.
void Session::project_constrained(StepReport* rep) {
  ddgrs_ctx* ctx = s->ctx;
  const int64_t nv = s->nv;
  const size_t nd = 3 * (size_t)nv, kn = nl_rows.size();
  const HostMesh hm{s->faces_h.data(), s->nf, nv};
  std::vector<double> pos(nd);
  download(ctx, pos.data(), (const double*)s->positions.ptr, nd);
  ctx->sync();
  auto values = [&](const std::vector<double>& q) {
    std::vector<double> phi(3 + kn);
    double b[3];
    area_barycenter(hm, q.data(), b);
    for (int k = 0; k < 3; ++k) phi[k] = b[k] - target[k];
    for (size_t r = 0; r < kn; ++r) phi[3 + r] = row_value(hm, q.data(), nl_rows[r]) - nl_targets[r];
    return phi;
  };
  auto positional = [&](std::vector<double>& q) {
    double b[3];
    area_barycenter(hm, q.data(), b);
    for (int64_t v = 0; v < nv; ++v)
      for (int k = 0; k < 3; ++k) q[3 * v + k] -= b[k] - target[k];
  };
  last_phi_before = values(pos);
  positional(pos);
  std::vector<double> phi = values(pos);
  int rounds = 0;
  std::vector<size_t> kept = schur.kept;
  std::sort(kept.begin(), kept.end());
  if (!kept.empty()) {
    auto measure = [&](const std::vector<double>& ph) {
      double m = 0.0;
      for (size_t i : kept) {
        const double v = std::fabs(ph[3 + i]) / row_tolerance(nl_targets[i]);
        m = (std::isnan(v) || std::isnan(m)) ? NAN : std::max(m, v);
      }
      return m;
    };
    double m = measure(phi);
    while (rounds < opts.newton_count) {
      if (!opts.newton_fixed && m <= 1.0) break;
      const std::vector<double> h = schur.correction(std::vector<double>(phi.begin() + 3, phi.end()));
      std::vector<double> trial = pos;
      for (size_t i = 0; i < nd; ++i) trial[i] -= h[i];
      positional(trial);
      const std::vector<double> pt = values(trial);
      const double mt = measure(pt);
      if (!(mt < m)) break;
      pos = std::move(trial);
      phi = pt;
      m = mt;
      ++rounds;
    }
  }
  last_phi = phi;
  rep->newton_rounds = rounds;
  upload(ctx, s->positions, pos.data(), nd);
  compute_face_geometry(s, s->packs, false);
}

// The near-field panels' column counts on stderr (count, mean, largest, and the 64-column chunks
// of the longest panels), for reading the panel kernels' balance.
static void trace_panels(ddgrs_ctx* ctx, const BlockTreeDev& bt) {
  std::vector<uint32_t> col_ptr(bt.nl + 1), leaf(bt.np), rows(bt.np);
  download(ctx, col_ptr.data(), bt.col_ptr.ptr, col_ptr.size());
  download(ctx, leaf.data(), bt.panel_leaf.ptr, leaf.size());
  download(ctx, rows.data(), bt.panel_rows.ptr, rows.size());
  std::vector<uint32_t> k(bt.np);
  double sum = 0.0, rsum = 0.0;
  for (int p = 0; p < bt.np; ++p) {
    k[p] = col_ptr[leaf[p] + 1] - col_ptr[leaf[p]];
    sum += k[p];
    rsum += rows[p];
  }
  std::sort(k.begin(), k.end());
  if (k.empty()) return;
  auto at = [&](double q) { return k[std::min(k.size() - 1, (size_t)(q * k.size()))]; };
  std::fprintf(stderr, "panels: %d leaves %d panels, rows/panel %.2f, columns mean %.1f p50 %u p90 %u p99 %u max %u\n",
               bt.nl, bt.np, rsum / bt.np, sum / bt.np, at(0.5), at(0.9), at(0.99), k.back());
}

void Session::step(StepReport* rep) {
  auto t_step = std::chrono::steady_clock::now();
  ddgrs_ctx* ctx = s->ctx;
  const int64_t nv = s->nv, nf = s->nf;
  const size_t nd = 3 * (size_t)nv, n = nd + 3;
  std::memset(rep, 0, sizeof(*rep));
  // 1. trees on the device; L′'s factorization is enqueued before the host builds the Barnes–Hut
  // lists, so the two overlap
  auto t = std::chrono::steady_clock::now();
  // DDGRS_TRACE_STEP=1: wall times of the tree stage's parts on stderr
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  auto t_part = t;
  auto mark = [&](const char* what) {
    if (!trace) return;
    std::fprintf(stderr, "step: %-18s %8.2f ms\n", what, seconds_since(t_part) * 1e3);
    t_part = std::chrono::steady_clock::now();
  };
  compute_face_geometry(s, s->packs, true);
  const HostGeometry geo = download_geometry(s);
  const double s_exp = (exps.beta64 - 2.0) / exps.alpha64;
  mark("geometry");
  // the vertex side on the second stream and another thread: vertex BVH, L′'s factorization,
  // L̃^V (positions and areas are final on the device: the geometry download synchronized)
  aux->profiling = ctx->profiling;
  auto side = std::async(std::launch::async, [&, s_exp] {
    HIPCHECK(hipSetDevice(aux->device));
    Topology vtopo = vert_bvh.build(aux, (size_t)nv, s->positions.ptr, s->positions.ptr, s->positions.ptr, opts.leaf_size);
    if (!lp_ready) {
      pc.setup_device(aux, lp_symbolic.get());
      lp_ready = true;
    }
    pc.form = opts.pc_form;
    pc.unit_scale = opts.pc_unit_scale;
    pc.refine = opts.pc_refine;
    pc.factor_device(aux, s, s_exp, opts.w_tp, opts.eps_rel);
    pc.build_fractional(aux, geo.pos, geo.vmass, 2.0 - s_exp, opts.chi, opts.leaf_pairs, std::move(vtopo));
    aux->sync();
  });
  // the face tree is built into the Barnes–Hut evaluator's members, which the step's other users
  // read (nothing rebuilds it before the next step)
  build_face_tree(s, opts.leaf_size, face_bvh, &bh.topo, &bh.clusters);
  const Topology& topo = bh.topo;
  const Clusters& clusters = bh.clusters;
  const DeviceTree dtree = device_tree(face_bvh, topo);
  mark("face tree");
  // the face metric's block tree on the third stream and another thread (reads the face tree and
  // the surface's geometry, which stay unchanged until the line search)
  fm.w_hi = opts.w_tp;
  fm.w_lo = opts.w_tp * opts.c_lo;
  aux2->profiling = ctx->profiling;
  auto metric = std::async(std::launch::async, [&, s_exp] {
    HIPCHECK(hipSetDevice(aux2->device));
    const auto t0 = std::chrono::steady_clock::now();
    fm.build(aux2, s, topo, clusters, opts.chi, opts.leaf_pairs, s_exp, &dtree);
    aux2->sync();
    return seconds_since(t0);
  });
  if (!opts.energy_exact) bh.build_from_members(s, opts.theta, opts.leaf_size);
  mark("Barnes-Hut lists");
  rep->seconds_trees = seconds_since(t);
  t = std::chrono::steady_clock::now();
  if (opts.energy_exact) {
    exact_pairs(ctx, nf, s->packs, s->area.ptr, exps, partials.ptr, rows.ptr);
    sum_f64(ctx, nf, rows.ptr, scal.ptr);
    rep->pairs_exact = (int64_t)nf * (nf - 1);
  } else {
    bh.gradient(s, exps, scal.ptr, partials.ptr);
    rep->pairs_exact = (int64_t)bh.pairs_exact;
    rep->nodes_lumped = (int64_t)bh.n_lumped;
  }
  pull_back(s, partials.ptr, g.ptr);
  HIPCHECK(hipMemsetAsync(g.ptr + nd, 0, 3 * sizeof(double), ctx->stream));
  const double f0 = read_scalar(scal.ptr);
  const double gnorm = vec.norm_host(ctx, g.ptr);
  rep->seconds_energy = seconds_since(t);
  rep->energy_before = f0;
  rep->gradient_norm = gnorm;
  if (!std::isfinite(f0) || !std::isfinite(gnorm)) {
    Error e(DDGRS_NON_FINITE_ENERGY, "step", "the energy or its differential is not finite");
    throw e;
  }
  // the learned preconditioner's blocks at this step's positions, while the other streams finish
  const bool use_learned = learned && learned_mode == 2 && learned_misses < kLearnedMissLimit;
  if (use_learned) {
    const auto tl = std::chrono::steady_clock::now();
    learned->features(ctx, geo.pos.data(), learned->feat.ptr);
    learned->forward(ctx, learned->feat.ptr);
    rep->seconds_learned = seconds_since(tl);
  }
  // 2. the solver of A′ and the direction x = A′⁻¹ [g; 0]
  t = std::chrono::steady_clock::now();
  const double metric_build = metric.get();  // the third stream is synchronized
  if (trace) std::fprintf(stderr, "step: %-18s %8.2f ms (wait %.2f ms)\n", "metric build", metric_build * 1e3,
                          seconds_since(t) * 1e3);
  if (ctx->profiling) ctx_merge_times(ctx, aux2);
  if (trace) trace_panels(ctx, fm.bt);
  rep->seconds_metric = seconds_since(t);
  t = std::chrono::steady_clock::now();
  side.get();  // the second stream is synchronized
  pc.check_factor(ctx);
  if (ctx->profiling) ctx_merge_times(ctx, aux);
  rep->seconds_precond = seconds_since(t);
  op.fm = &fm;
  op.crow = pc.crow_dev;
  op.nv = (size_t)nv;
  t = std::chrono::steady_clock::now();
  GmresSettings gs{opts.restart, opts.tolerance, opts.max_iterations};
  auto A = [&](const double* x, double* y) { op.apply(ctx, x, y); };
  auto P = [&](const double* r, double* z) { pc.apply(ctx, r, z); };
  // FP32 products first; the extended near field once their rounding is the residual's floor
  fm.extended = pc.vf.extended = false;
  auto stall = [&]() {
    if (fm.extended) return false;
    fm.extended = pc.vf.extended = true;
    return true;
  };
  const int lmode = learned ? learned_mode : 0;
  rep->learned_mode = lmode;
  GmresResult gr;
  bool solved = false;
  if (lmode == 2 && !use_learned) {
    rep->learned_fallback = 2;  // Q off for the rest of the session after consecutive misses
  } else if (lmode == 2) {
    // (I + Q) P⁻¹ with the blocks formed above; on a miss the plain solve below from x₀ = 0
    kz.reserve(n);
    auto PQ = [&](const double* r, double* z) {
      pc.apply(ctx, r, z);
      HIPCHECK(hipMemcpyAsync(kz.ptr, z, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      learned->apply(ctx, kz.ptr, z);
    };
    // stagnation bound: twice the last converged solve's iterations (+5; one restart cycle before
    // any) before the plain fallback
    GmresSettings gl = gs;
    gl.max_iterations = std::min(gs.max_iterations, last_iterations > 0 ? 2 * last_iterations + 5 : gs.restart);
    gr = gmres_device(ctx, n, A, PQ, g.ptr, sol.ptr, gl, vec, stall);
    fm.extended = pc.vf.extended = false;
    rep->learned_iterations = (int64_t)gr.iterations;
    solved = gr.converged;
    rep->learned_fallback = solved ? 0 : 1;
    learned_misses = solved ? 0 : learned_misses + 1;
  }
  if (!solved) gr = gmres_device(ctx, n, A, P, g.ptr, sol.ptr, gs, vec, stall);
  fm.extended = pc.vf.extended = false;
  // the recovery chain of a diverged solve (RS/src/step/direction.rs:277 escalate): right GMRES
  // with restart 60 and at least 800 iterations, then a dense LU of the exact metric's saddle
  // system for meshes of at most dense_max_vertices vertices
  fallbacks.clear();
  if (!gr.converged && !(gr.relative_residual <= opts.diverged_residual) && opts.robust) {
    const double r0 = gr.relative_residual;
    char line[256];
    GmresSettings g2{60, opts.tolerance,
                     opts.retry_max_iterations > 0 ? (size_t)opts.retry_max_iterations
                                                   : std::max<size_t>(4 * opts.max_iterations, 800)};
    GmresResult gr2 = gmres_device(ctx, n, A, P, g.ptr, sol.ptr, g2, vec, stall);
    fm.extended = pc.vf.extended = false;
    if (gr2.converged || gr2.relative_residual <= opts.diverged_residual) {
      std::snprintf(line, sizeof line,
                    "gmres diverged (relative residual %.3e): retried with right GMRES, augmented preconditioner on "
                    "vertex pairs, restart 60, at most %zu iterations (final residual %.3e)",
                    r0, g2.max_iterations, gr2.relative_residual);
      fallbacks.push_back(line);
      gr = gr2;
      rep->fallback_stage = 1;
    } else if (nv <= opts.dense_max_vertices) {
      const double rr = dense_saddle_solve(ctx, s, topo, clusters, s_exp, fm.w_hi, fm.w_lo, pc.crow_dev, g.ptr, sol.ptr);
      std::snprintf(line, sizeof line,
                    "gmres diverged (relative residual %.3e): dense LU of the %lld x %lld metric (residual %.3e)", r0,
                    (long long)nv, (long long)nv, rr);
      fallbacks.push_back(line);
      gr = GmresResult();
      gr.converged = rr <= opts.tolerance;
      gr.relative_residual = rr;
      rep->fallback_stage = 2;
    } else {
      gr = gr2;
      gr.relative_residual = r0;  // SolverDiverged reports the first solve's residual
    }
  }
  if (gr.converged) last_iterations = gr.iterations;
  if (lmode == 1 && gr.converged) {
    // the pair (g, P⁻¹ g, A′⁻¹ g) at this step's positions, then pairs of the vectors the plain
    // solve's Krylov space is made of: with z₀ = P⁻¹ g, r_k = A′ z_{k−1} has A′⁻¹ r_k = z_{k−1}
    // exactly, and z_k = P⁻¹ r_k
    pg.reserve(n);
    kr.reserve(n);
    kz.reserve(n);
    pc.apply(ctx, g.ptr, pg.ptr);
    std::vector<double> hr(n), hp(n), hy(n);
    auto add = [&](const double* r, const double* pr, const double* y) {
      if (learned->npairs >= learned->cap_pairs) return;
      download(ctx, hr.data(), r, n);
      download(ctx, hp.data(), pr, n);
      download(ctx, hy.data(), y, n);
      ctx->sync();
      learned->add_pair(ctx, geo.pos.data(), hr.data(), hp.data(), hy.data());
    };
    add(g.ptr, pg.ptr, sol.ptr);
    double *z = pg.ptr, *r = kr.ptr, *pz = kz.ptr;
    for (int k = 0; k < learned->krylov_pairs && learned->npairs < learned->cap_pairs; ++k) {
      op.apply(ctx, z, r);
      pc.apply(ctx, r, pz);
      add(r, pz, z);
      std::swap(z, pz);
    }
  }
  rep->seconds_solve = seconds_since(t);
  rep->gmres.iterations = (int64_t)gr.iterations;
  rep->gmres.cycles = (int64_t)gr.cycles;
  rep->gmres.operator_applications = (int64_t)gr.operator_applications;
  rep->gmres.preconditioner_applications = (int64_t)gr.preconditioner_applications;
  rep->gmres.converged = gr.converged ? 1 : 0;
  rep->gmres.extended = gr.extended ? 1 : 0;
  rep->gmres.relative_residual = gr.relative_residual;
  last_gmres = gr;
  if (!gr.converged && !(gr.relative_residual <= opts.diverged_residual)) {
    Error e(DDGRS_SOLVER_DIVERGED, "gmres", "true relative residual above 0.1 after the iteration cap");
    e.value = gr.relative_residual;
    throw e;
  }
  // the nonlinear rows (schur.rs:101 with_direction): z_r = A′⁻¹ [C_N row r; 0] with the step's
  // operator and preconditioner, S = −C_N Z and its factor, then x = y − Z N⁻¹ C_N y (host, FP64)
  rep->min_pivot = 1.0;
  if (!nl_rows.empty()) {
    const HostMesh hm{s->faces_h.data(), nf, nv};
    std::vector<std::vector<double>> crow, zcol;
    std::vector<double> hb(n, 0.0), hz(n);
    for (int kind : nl_rows) {
      crow.push_back(row_gradient(hm, geo.pos.data(), kind));
      std::copy(crow.back().begin(), crow.back().end(), hb.begin());
      upload(ctx, zrhs, hb.data(), n);
      zsol.reserve(n);
      GmresResult gz = gmres_device(ctx, n, A, P, zrhs.ptr, zsol.ptr, gs, vec, stall);
      fm.extended = pc.vf.extended = false;
      if (!gz.converged && !(gz.relative_residual <= opts.diverged_residual)) {
        Error e(DDGRS_SOLVER_DIVERGED, "gmres", "a Schur column's solve: true relative residual above 0.1");
        e.value = gz.relative_residual;
        throw e;
      }
      download(ctx, hz.data(), (const double*)zsol.ptr, n);
      ctx->sync();
      zcol.emplace_back(hz.begin(), hz.begin() + (long)nd);
    }
    schur.build(std::move(crow), std::move(zcol));
    std::vector<double> y(nd);
    download(ctx, y.data(), (const double*)sol.ptr, nd);
    ctx->sync();
    const std::vector<double> x = schur.direction(y);
    HIPCHECK(hipMemcpyAsync(sol.ptr, x.data(), nd * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
    rep->min_pivot = schur.min_pivot;
  }
  // ⟨g, x⟩, ‖x‖, ‖x‖_∞ and the largest vertex displacement
  vec.dots(ctx, g.ptr, 1, sol.ptr, scal.ptr + 1);
  const unsigned nbv = blocks_for(nv, 256);
  launch(ctx, "xmax", xmax_kernel, dim3(nbv), dim3(256), 0, nv, (const double*)sol.ptr, xpart.ptr);
  std::vector<double> xp(3 * (size_t)nbv);
  download(ctx, xp.data(), (const double*)xpart.ptr, 3 * (size_t)nbv);
  const double slope = read_scalar(scal.ptr + 1);
  double maxd = 0.0, xinf = 0.0, xn2 = 0.0;
  for (unsigned b = 0; b < nbv; ++b) {
    maxd = std::fmax(maxd, xp[3 * b]);
    xinf = std::fmax(xinf, xp[3 * b + 1]);
    xn2 += xp[3 * b + 2];
  }
  const double xnorm = std::sqrt(xn2);
  rep->sobolev_norm_sq = slope;
  rep->direction_norm = xnorm;
  rep->direction_max = xinf;
  // 3. line search (Armijo, halving; the start rule of session.rs:1604 start_tau)
  t = std::chrono::steady_clock::now();
  double tau0;
  switch (opts.start_kind) {
    case DDGRS_START_ABSOLUTE: tau0 = opts.start_value; break;
    case DDGRS_START_MAX_DISPLACEMENT: tau0 = opts.start_value * l0 / maxd; break;
    case DDGRS_START_REFERENCE: tau0 = 2.0 / xnorm; break;
    default: tau0 = last_tau > 0.0 ? 2.0 * last_tau : opts.start_value * l0 / maxd; break;
  }
  const double tau_min = opts.tau_min_rel * l0 / xinf;
  // the CCD cap (session.rs:917-956): the band of the faces that intersect at the start, then
  // τ₀ ← min(τ₀, ⅔ τ_max) along −x when a contact is found
  if (opts.ccd) {
    const auto tc = std::chrono::steady_clock::now();
    const size_t np = ccd.band_flags(ctx, s->faces.ptr, s->positions.ptr);
    if (std::isfinite(tau0) && tau0 > 0.0) {
      CcdOpts co = opts.ccd_opts;
      if (std::isnan(co.reference_length)) co.reference_length = l0;
      const CcdResult cr = ccd.first_contact(ctx, s->faces.ptr, s->positions.ptr, geo.pos.data(), sol.ptr, -1.0, tau0,
                                             co, np ? ccd.band.ptr : nullptr);
      if (cr.has_contact) tau0 = std::fmin(tau0, 2.0 / 3.0 * cr.tau_max);
      rep->has_ccd = 1;
      rep->tau_ccd = cr.tau_max;
      rep->ccd.tau_max = cr.tau_max;
      rep->ccd.has_contact = cr.has_contact ? 1 : 0;
      rep->ccd.contact_kind = cr.kind;
      rep->ccd.contact_a = cr.a;
      rep->ccd.contact_b = cr.b;
      rep->ccd.gap_abs = cr.gap_abs;
      rep->ccd.iterations = (int64_t)cr.iterations;
      rep->ccd.pairs_tested = (int64_t)cr.pairs_tested;
      rep->ccd.close_pairs = (int64_t)cr.close_pairs;
      rep->ccd.touching_pairs = (int64_t)cr.touching_pairs;
      rep->ccd.exempt_pairs = (int64_t)cr.exempt_pairs;
      rep->ccd.windows = (int32_t)cr.windows;
    }
    rep->seconds_ccd = seconds_since(tc);
    rep->ccd.seconds = rep->seconds_ccd;
  }
  rep->tau0 = tau0;
  HIPCHECK(hipMemcpyAsync(p0.ptr, s->positions.ptr, nd * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  bool accepted = false;
  double tau = tau0;
  trials.clear();
  if (slope > 0.0 && std::isfinite(tau0) && tau0 > 0.0) {
    for (;;) {
      if (!(tau >= tau_min) || (opts.max_halvings >= 0 && rep->halvings > opts.max_halvings)) break;
      launch(ctx, "trial", trial_kernel, dim3(blocks_for((int64_t)nd, 256)), dim3(256), 0, nd, (const double*)p0.ptr,
             (const double*)sol.ptr, tau, s->positions.ptr);
      compute_face_geometry(s, s->packs, false);
      if (!opts.energy_exact) bh.refresh(s);
      const double ft = energy_value();
      rep->trials++;
      const bool ok = std::isfinite(ft) && ft <= f0 - opts.armijo_c * tau * slope;
      trials.push_back({tau, std::isfinite(ft) ? ft : NAN, !std::isfinite(ft) ? 2 : ok ? 0 : 1});
      if (ok) {
        accepted = true;
        rep->energy_trial = ft;
        break;
      }
      tau *= 0.5;
      rep->halvings++;
    }
  }
  if (!accepted) {
    HIPCHECK(hipMemcpyAsync(s->positions.ptr, p0.ptr, nd * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    compute_face_geometry(s, s->packs, false);
    rep->energy_trial = f0;
  }
  rep->seconds_line_search = seconds_since(t);
  // 4. projection onto the barycenter constraint, then F at the projected positions
  // (session.rs:1505 `after`)
  t = std::chrono::steady_clock::now();
  if (nl_rows.empty())
    project();
  else
    project_constrained(rep);
  if (!opts.energy_exact) bh.refresh(s);
  rep->energy_after = energy_value();
  rep->seconds_projection = seconds_since(t);
  rep->accepted = accepted ? 1 : 0;
  rep->tau = accepted ? tau : 0.0;
  if (accepted) last_tau = tau;
  rep->seconds_total = seconds_since(t_step);
}

}  // namespace ddgrs
