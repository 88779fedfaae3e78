// Restarted right-preconditioned GMRES and the saddle operator on the device (see gmres.hpp).
#include <cmath>

#include "gmres.hpp"

namespace ddgrs {

#pragma clang fp contract(off)

constexpr int kDotChunk = 2048;  // elements per block of a reduction (fixed: results independent of the device)

__device__ inline double block_sum256(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// partial[b·count + k] = Σ_{i ∈ chunk b} V_k[i] w[i]
__global__ __launch_bounds__(256) void dots_partial_kernel(size_t n, const double* __restrict__ V, int count,
                                                           const double* __restrict__ w, double* __restrict__ partial) {
  __shared__ double sh[256];
  const size_t i0 = (size_t)blockIdx.x * kDotChunk;
  double wl[kDotChunk / 256];
#pragma unroll
  for (int e = 0; e < kDotChunk / 256; ++e) {
    size_t i = i0 + (size_t)e * 256 + threadIdx.x;
    wl[e] = i < n ? w[i] : 0.0;
  }
  for (int k = 0; k < count; ++k) {
    const double* v = V + (size_t)k * n;
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < kDotChunk / 256; ++e) {
      size_t i = i0 + (size_t)e * 256 + threadIdx.x;
      if (i < n) s += v[i] * wl[e];
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * count + k] = s;
  }
}

__global__ __launch_bounds__(64) void dots_final_kernel(int nblocks, int count, const double* __restrict__ partial,
                                                        double* __restrict__ out) {
  int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= count) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * count + k];
  out[k] = s;
}

__global__ __launch_bounds__(256) void subtract_kernel(size_t n, const double* __restrict__ V, int count,
                                                       const double* __restrict__ h, double* __restrict__ w) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int k = 0; k < count; ++k) s -= h[k] * V[(size_t)k * n + i];
  w[i] = s;
}

void DeviceVecs::init(size_t n_) {
  n = n_;
  size_t nb = (n + kDotChunk - 1) / kDotChunk;
  partial.reserve(nb * 64);
  scalars.reserve(256);
}

void DeviceVecs::dots(ddgrs_ctx* ctx, const double* V, size_t count, const double* w, double* out) {
  unsigned nb = (unsigned)((n + kDotChunk - 1) / kDotChunk);
  partial.reserve((size_t)nb * count);
  launch(ctx, "gm_dots", dots_partial_kernel, dim3(nb), dim3(256), 0, n, V, (int)count, w, partial.ptr);
  launch(ctx, "gm_dots_final", dots_final_kernel, dim3(blocks_for((int64_t)count, 64)), dim3(64), 0, (int)nb,
         (int)count, (const double*)partial.ptr, out);
}

void DeviceVecs::subtract(ddgrs_ctx* ctx, const double* V, size_t count, const double* h, double* w) {
  launch(ctx, "gm_subtract", subtract_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, n, V, (int)count, h, w);
}

double DeviceVecs::norm_host(ddgrs_ctx* ctx, const double* x) {
  dots(ctx, x, 1, x, scalars.ptr);
  double* pin = (double*)ctx->staging(8);
  download(ctx, pin, (const double*)scalars.ptr, 1);
  ctx->sync();
  return std::sqrt(pin[0]);
}

// ---- GMRES kernels ---------------------------------------------------------------------------

struct GmresState {
  double *H, *cs, *sn, *g, *h, *h2, *y, *out;  // H: column j holds m+1 entries
  int m1;                                       // m + 1
};

// The new Hessenberg column h (+ the second Gram–Schmidt pass h2) with ‖w‖: previous rotations,
// the new rotation, g, and the Arnoldi estimate |g_{j+1}| / ‖b‖ (gmres.rs:466-487).
__global__ void givens_kernel(GmresState st, int j, const double* __restrict__ hn_sq, double bnorm) {
  double* col = st.H + (size_t)j * st.m1;
  for (int i = 0; i <= j; ++i) col[i] = st.h[i] + st.h2[i];
  double hn = sqrt(hn_sq[0]);
  col[j + 1] = hn;
  for (int i = 0; i < j; ++i) {
    double t = st.cs[i] * col[i] + st.sn[i] * col[i + 1];
    col[i + 1] = -st.sn[i] * col[i] + st.cs[i] * col[i + 1];
    col[i] = t;
  }
  double denom = hypot(col[j], col[j + 1]);
  if (denom == 0.0) {
    st.cs[j] = 1.0;
    st.sn[j] = 0.0;
  } else {
    st.cs[j] = col[j] / denom;
    st.sn[j] = col[j + 1] / denom;
  }
  col[j] = denom;
  col[j + 1] = 0.0;
  st.g[j + 1] = -st.sn[j] * st.g[j];
  st.g[j] *= st.cs[j];
  st.out[0] = fabs(st.g[j + 1]) / bnorm;
  st.out[1] = hn;
}

// Back substitution H y = g (k × k upper triangular, gmres.rs:502-512).
__global__ void backsub_kernel(GmresState st, int k) {
  for (int i = k - 1; i >= 0; --i) {
    double s = st.g[i];
    for (int l = i + 1; l < k; ++l) s -= st.H[(size_t)l * st.m1 + i] * st.y[l];
    double hii = st.H[(size_t)i * st.m1 + i];
    st.y[i] = hii != 0.0 ? s / hii : 0.0;
  }
}

__global__ __launch_bounds__(256) void combine_kernel(size_t n, const double* __restrict__ V, int k,
                                                      const double* __restrict__ y, double* __restrict__ u) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int l = 0; l < k; ++l) s += y[l] * V[(size_t)l * n + i];
  u[i] = s;
}

__global__ __launch_bounds__(256) void scale_kernel(size_t n, const double* __restrict__ w, const double* __restrict__ nrm2,
                                                    int take_sqrt, double* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = take_sqrt ? sqrt(nrm2[0]) : nrm2[0];
  out[i] = w[i] / s;
}

__global__ __launch_bounds__(256) void residual_kernel(size_t n, const double* __restrict__ b, const double* __restrict__ t,
                                                       double* __restrict__ r) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) r[i] = b[i] - t[i];
}

__global__ __launch_bounds__(256) void add_kernel(size_t n, const double* __restrict__ t, double* __restrict__ x) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] += t[i];
}

__global__ void init_g_kernel(double* g, int m1, const double* nrm2) {
  for (int i = 0; i < m1; ++i) g[i] = 0.0;
  g[0] = sqrt(nrm2[0]);
}

GmresResult gmres_device(ddgrs_ctx* ctx, size_t n, const std::function<void(const double*, double*)>& op,
                         const std::function<void(const double*, double*)>& prec, const double* b, double* x,
                         const GmresSettings& o, DeviceVecs& vec, const std::function<bool()>& on_stall) {
  GmresResult rep;
  bool est_reached = false;  // the last cycle's Arnoldi estimate reached the tolerance
  double prev_rel = 0.0;
  const size_t m = std::min(o.restart, std::max<size_t>(n, 1));
  const unsigned nb = blocks_for((int64_t)n, 256);
  HIPCHECK(hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
  const double bnorm = vec.norm_host(ctx, b);
  if (bnorm == 0.0) {
    rep.converged = true;
    return rep;
  }
  DevBuf<double>&Vb = vec.Vb, &w = vec.w, &t = vec.t, &r = vec.r, &small = vec.small;
  Vb.reserve((m + 1) * n);
  w.reserve(n);
  t.reserve(n);
  r.reserve(n);
  const int m1 = (int)m + 1;
  small.reserve((size_t)m1 * m + 6 * (size_t)m1 + 8);
  GmresState st;
  st.m1 = m1;
  st.H = small.ptr;
  st.cs = st.H + (size_t)m1 * m;
  st.sn = st.cs + m1;
  st.g = st.sn + m1;
  st.h = st.g + m1;
  st.h2 = st.h + m1;
  st.y = st.h2 + m1;
  st.out = st.y + m1;  // 2 scalars + ‖w‖²
  double* nrm2 = st.out + 2;
  double* pin = (double*)ctx->staging(16);
  for (;;) {
    // explicit residual r = b − A x (x = 0 in the first cycle)
    if (rep.cycles > 0) {
      op(x, t.ptr);
      rep.operator_applications++;
      launch(ctx, "gm_residual", residual_kernel, dim3(nb), dim3(256), 0, n, b, (const double*)t.ptr, r.ptr);
    } else {
      HIPCHECK(hipMemcpyAsync(r.ptr, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    vec.dots(ctx, r.ptr, 1, r.ptr, nrm2);
    download(ctx, pin, (const double*)nrm2, 1);
    ctx->sync();
    const double beta = std::sqrt(pin[0]);
    const double rel = beta / bnorm;
    rep.relative_residual = rel;
    rep.cycle_residuals.push_back(rel);
    if (!std::isfinite(rel)) break;
    if (rel <= o.tolerance) {
      rep.converged = true;
      break;
    }
    const bool stalled = est_reached || (rel > 0.25 * prev_rel && rel < 100.0 * o.tolerance);
    if (rep.cycles > 0 && !rep.extended && on_stall && stalled && on_stall()) {
      rep.extended = true;
      est_reached = false;
      continue;  // the explicit residual again, with the switched products
    }
    prev_rel = rel;
    if (rep.iterations >= o.max_iterations) break;
    rep.cycles++;
    launch(ctx, "gm_scale", scale_kernel, dim3(nb), dim3(256), 0, n, (const double*)r.ptr, (const double*)nrm2, 1, Vb.ptr);
    launch(ctx, "gm_init", init_g_kernel, dim3(1), dim3(1), 0, st.g, m1, (const double*)nrm2);
    size_t j = 0, k_end = 0;
    for (;;) {
      double* vj = Vb.ptr + j * n;
      prec(vj, t.ptr);
      op(t.ptr, w.ptr);
      rep.operator_applications++;
      rep.preconditioner_applications++;
      // classical Gram–Schmidt, twice
      vec.dots(ctx, Vb.ptr, j + 1, w.ptr, st.h);
      vec.subtract(ctx, Vb.ptr, j + 1, st.h, w.ptr);
      vec.dots(ctx, Vb.ptr, j + 1, w.ptr, st.h2);
      vec.subtract(ctx, Vb.ptr, j + 1, st.h2, w.ptr);
      vec.dots(ctx, w.ptr, 1, w.ptr, nrm2);
      launch(ctx, "gm_givens", givens_kernel, dim3(1), dim3(1), 0, st, (int)j, (const double*)nrm2, bnorm);
      download(ctx, pin, (const double*)st.out, 2);
      ctx->sync();
      const double est = pin[0], hn = pin[1];
      rep.iterations++;
      rep.history.push_back(est);
      if (hn == 0.0 || !std::isfinite(hn) || est <= o.tolerance || rep.iterations >= o.max_iterations || j + 1 >= m) {
        k_end = j + 1;
        est_reached = est <= o.tolerance;
        break;
      }
      launch(ctx, "gm_scale", scale_kernel, dim3(nb), dim3(256), 0, n, (const double*)w.ptr, (const double*)nrm2, 1,
             Vb.ptr + (j + 1) * n);
      ++j;
    }
    launch(ctx, "gm_backsub", backsub_kernel, dim3(1), dim3(1), 0, st, (int)k_end);
    launch(ctx, "gm_combine", combine_kernel, dim3(nb), dim3(256), 0, n, (const double*)Vb.ptr, (int)k_end,
           (const double*)st.y, w.ptr);
    prec(w.ptr, t.ptr);
    rep.preconditioner_applications++;
    launch(ctx, "gm_add", add_kernel, dim3(nb), dim3(256), 0, n, (const double*)t.ptr, x);
  }
  return rep;
}

// ---- saddle operator -------------------------------------------------------------------------

// Per block partial sums of C_Q x (3 axes); y[3v+a] += c_v λ_a.
__global__ __launch_bounds__(256) void cq_kernel(uint32_t nv, const double* __restrict__ crow, const double* __restrict__ x,
                                                 double* __restrict__ y, double* __restrict__ partial) {
  __shared__ double sh[256];
  const size_t nd = 3 * (size_t)nv;
  uint32_t v = blockIdx.x * 256 + threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  if (v < nv) {
    double c = crow[v];
    for (int a = 0; a < 3; ++a) {
      s[a] = c * x[3 * (size_t)v + a];
      y[3 * (size_t)v + a] += c * x[nd + a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    double t = block_sum256(s[a], sh);
    if (threadIdx.x == 0) partial[3 * (size_t)blockIdx.x + a] = t;
  }
}

__global__ void cq_final_kernel(uint32_t nblocks, uint32_t nv, const double* __restrict__ partial, double* __restrict__ y) {
  int a = threadIdx.x;
  if (a >= 3) return;
  double s = 0.0;
  for (uint32_t b = 0; b < nblocks; ++b) s += partial[3 * (size_t)b + a];
  y[3 * (size_t)nv + a] = s;
}

void SaddleOp::apply(ddgrs_ctx* ctx, const double* x, double* y) {
  fm->apply(ctx, x, y);
  const unsigned nb = blocks_for((int64_t)nv, 256);
  part.reserve(3 * (size_t)nb);
  launch(ctx, "cq", cq_kernel, dim3(nb), dim3(256), 0, (uint32_t)nv, crow, x, y, part.ptr);
  launch(ctx, "cq_final", cq_final_kernel, dim3(1), dim3(64), 0, (uint32_t)nb, (uint32_t)nv, (const double*)part.ptr, y);
}

}  // namespace ddgrs
