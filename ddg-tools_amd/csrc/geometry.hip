// Per-face geometry, the FP32 face packs read by the pair kernels, the pull-back of per-face
// partials to the vertices, and deterministic reductions.
//
// Face quantities follow RS/src/geometry.rs:33 `face_quantities` and the pull-back follows
// :87 `face_pull_back` / :172 `pull_back` (faces added in ascending order per vertex).  These O(F)
// epilogues run in FP64 without contraction (the surface state is FP64), so for equal inputs they
// round exactly as the reference does.
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace ddgrs {

__device__ inline void load3(const double* p, uint32_t v, double o[3]) {
  o[0] = p[3 * v];
  o[1] = p[3 * v + 1];
  o[2] = p[3 * v + 2];
}

// One thread per face: a_S, N_S, X_S (FP64) and the FP32 packs
//   xhi = (X hi, a), xlo = (X - hi, 0), nrm = (N, 0)
// and the validity of the face (surface.rs:526-538: area > 0, finite area and normal).
__global__ __launch_bounds__(256) void face_geometry_kernel(int64_t nf, const uint32_t* __restrict__ faces,
                                                            const double* __restrict__ pos, double* __restrict__ area,
                                                            double* __restrict__ normal, double* __restrict__ bary,
                                                            float4* __restrict__ xhi, float4* __restrict__ xlo,
                                                            float4* __restrict__ nrm, int* __restrict__ first_bad) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double p0[3], p1[3], p2[3];
  load3(pos, faces[3 * f], p0);
  load3(pos, faces[3 * f + 1], p1);
  load3(pos, faces[3 * f + 2], p2);
  double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double third = 1.0 / 3.0;
  double x[3], nn[3];
  for (int k = 0; k < 3; ++k) x[k] = ((p0[k] + p1[k]) + p2[k]) * third;
  double inv = 1.0 / len;
  for (int k = 0; k < 3; ++k) nn[k] = n[k] * inv;
  double a = 0.5 * len;
  area[f] = a;
  for (int k = 0; k < 3; ++k) {
    normal[3 * f + k] = nn[k];
    bary[3 * f + k] = x[k];
  }
  float hx = (float)x[0], hy = (float)x[1], hz = (float)x[2];
  xhi[f] = make_float4(hx, hy, hz, (float)a);
  xlo[f] = make_float4((float)(x[0] - (double)hx), (float)(x[1] - (double)hy), (float)(x[2] - (double)hz), 0.f);
  nrm[f] = make_float4((float)nn[0], (float)nn[1], (float)nn[2], 0.f);
  bool ok = a > 0.0 && isfinite(a) && isfinite(nn[0]) && isfinite(nn[1]) && isfinite(nn[2]);
  if (!ok && first_bad) atomicMin(first_bad, (int)f);
}

// First vertex with a non-finite coordinate (surface.rs:520-525); integer min is order free.
__global__ __launch_bounds__(256) void finite_positions_kernel(int64_t nv, const double* __restrict__ pos,
                                                               int* __restrict__ first_bad) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  if (!(isfinite(pos[3 * v]) && isfinite(pos[3 * v + 1]) && isfinite(pos[3 * v + 2]))) atomicMin(first_bad, (int)v);
}

// One thread per vertex: ∂F/∂f_v = Σ_{S∋v, ascending} face_pull_back(S)[corner of v].
// partials: 7 per face (∂/∂a, ∂/∂X, ∂/∂N), face order.
__global__ __launch_bounds__(256) void pull_back_kernel(int64_t nv, const uint32_t* __restrict__ faces,
                                                        const double* __restrict__ pos,
                                                        const uint32_t* __restrict__ inc_ptr,
                                                        const uint32_t* __restrict__ inc,
                                                        const double* __restrict__ partials,
                                                        double* __restrict__ out) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t q = inc_ptr[v]; q < inc_ptr[v + 1]; ++q) {
    uint32_t e = inc[q];
    uint32_t s = e >> 2, c = e & 3u;
    double p[3][3];
    for (int j = 0; j < 3; ++j) load3(pos, faces[3 * s + j], p[j]);
    double e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    double e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double il = 1.0 / len;
    double nrm[3] = {n[0] * il, n[1] * il, n[2] * il};
    // opposite edge of corner c: [p2-p1, p0-p2, p1-p0]
    int i0 = (c + 1) % 3, i1 = (c + 2) % 3;
    double eo[3] = {p[i1][0] - p[i0][0], p[i1][1] - p[i0][1], p[i1][2] - p[i0][2]};
    const double* fp = partials + 7 * (int64_t)s;
    double d_area = fp[0];
    double third = 1.0 / 3.0;
    double xb[3] = {fp[1] * third, fp[2] * third, fp[3] * third};
    double inv_2a = 1.0 / len;
    double cr[3] = {nrm[1] * eo[2] - nrm[2] * eo[1], nrm[2] * eo[0] - nrm[0] * eo[2], nrm[0] * eo[1] - nrm[1] * eo[0]};
    double sa = 0.5 * d_area;
    double ga[3] = {cr[0] * sa, cr[1] * sa, cr[2] * sa};
    double ce[3] = {eo[1] * nrm[2] - eo[2] * nrm[1], eo[2] * nrm[0] - eo[0] * nrm[2], eo[0] * nrm[1] - eo[1] * nrm[0]};
    double sn = (ce[0] * fp[4] + ce[1] * fp[5] + ce[2] * fp[6]) * inv_2a;
    for (int k = 0; k < 3; ++k) acc[k] = acc[k] + ((ga[k] + xb[k]) + nrm[k] * sn);
  }
  for (int k = 0; k < 3; ++k) out[3 * v + k] = acc[k];
}

// Deterministic FP64 sum of n values in one block (fixed strided partition + fixed tree).
__global__ __launch_bounds__(1024) void sum_f64_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double sh[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ __launch_bounds__(256) void fill_int_kernel(int64_t n, int* __restrict__ x, int value) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = value;
}

// ---- host side -------------------------------------------------------------------------------

void FacePacks::ensure(int64_t nf) {
  xhi.reserve(nf);
  xlo.reserve(nf);
  nrm.reserve(nf);
}

void compute_face_geometry(ddgrs_surface* s, FacePacks& packs, bool validate) {
  ddgrs_ctx* ctx = s->ctx;
  s->area.reserve(s->nf);
  s->normal.reserve(3 * s->nf);
  s->bary.reserve(3 * s->nf);
  packs.ensure(s->nf);
  Scratch& sc = scratch_of(ctx);
  int* flags = sc.ints(2);
  if (validate) {
    launch(ctx, "fill_int", fill_int_kernel, dim3(1), dim3(256), 0, (int64_t)2, flags, INT32_MAX);
    launch(ctx, "finite_positions", finite_positions_kernel, dim3(blocks_for(s->nv, 256)), dim3(256), 0, s->nv,
           (const double*)s->positions.ptr, flags);
  }
  launch(ctx, "face_geometry", face_geometry_kernel, dim3(blocks_for(s->nf, 256)), dim3(256), 0, s->nf,
         (const uint32_t*)s->faces.ptr, (const double*)s->positions.ptr, s->area.ptr, s->normal.ptr, s->bary.ptr,
         packs.xhi.ptr, packs.xlo.ptr, packs.nrm.ptr, validate ? flags + 1 : (int*)nullptr);
  if (validate) {
    int h[2];
    int* pin = (int*)ctx->staging(8);
    download(ctx, pin, flags, 2);
    ctx->sync();
    h[0] = pin[0];
    h[1] = pin[1];
    if (h[0] != INT32_MAX) {
      Error e(DDGRS_NON_FINITE_POSITION, "validate_positions", "vertex " + std::to_string(h[0]) + " has a non-finite position");
      e.a = h[0];
      throw e;
    }
    if (h[1] != INT32_MAX) {
      Error e(DDGRS_DEGENERATE_FACE, "validate_positions", "face " + std::to_string(h[1]) + " has zero area");
      e.a = h[1];
      throw e;
    }
  }
}

void pull_back(ddgrs_surface* s, const double* partials, double* out) {
  launch(s->ctx, "pull_back", pull_back_kernel, dim3(blocks_for(s->nv, 256)), dim3(256), 0, s->nv,
         (const uint32_t*)s->faces.ptr, (const double*)s->positions.ptr, (const uint32_t*)s->inc_ptr.ptr,
         (const uint32_t*)s->inc.ptr, partials, out);
}

void sum_f64(ddgrs_ctx* ctx, int64_t n, const double* x, double* out) {
  launch(ctx, "sum_f64", sum_f64_kernel, dim3(1), dim3(1024), 0, n, x, out);
}

}  // namespace ddgrs
