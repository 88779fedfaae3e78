// Hierarchical products on the device (see blocktree.hpp).
//
// Follows RS/src/spatial/block_tree.rs:536 multiply_core (upward sums, far field per node,
// downward accumulation, rows = far part + near blocks of the chain root → leaf, diagonal skipped)
// and RS/src/metric/hierarchical.rs (FaceKernels :192-236, product :696: m = M x, ψ = a m,
// z = w (H̃a·m − H̃ψ), y = 2 Σ_{S∋v} a_S (Mᵀ row)·z_S; VertexFractionalTree::product :1313).
//
// Device form: the near field is stored dense per block in FP32 (NK values per entry, evaluated in
// FP64 with the reference's expressions, (min, max)-ordered so that mirrored blocks are bitwise
// transposes); far coefficients likewise.  The upward sum of a node is the sum of ψ over its
// contiguous BVH range (one wave per node that some far block reads), the downward accumulation is
// folded into the row kernel (a leaf sums the far values along its chain), and a row kernel runs
// one wave per leaf with lanes (row r, column lane j) sweeping the near blocks of the chain.
#include "blocktree.hpp"

namespace ddgrs {

__device__ __forceinline__ float wsum64(float x) {
  x += __shfl_xor(x, 32);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 1);
  return x;
}
__device__ __forceinline__ float jsum8(float x) {
  x += __shfl_xor(x, 1);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 4);
  return x;
}

template <int C>
__global__ __launch_bounds__(256) void bt_up_kernel(int n_needed, const uint32_t* __restrict__ needed,
                                                    const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                    const float* __restrict__ psi, float* __restrict__ up) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= n_needed) return;
  const uint32_t a = needed[w];
  float acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.f;
  for (uint32_t p = start[a] + lane; p < end[a]; p += 64)
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] += psi[(size_t)p * C + c];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = wsum64(acc[c]);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < C; ++c) up[(size_t)a * C + c] = acc[c];
}

template <int C, int H>
__global__ __launch_bounds__(256) void bt_far_kernel(int nn, const uint32_t* __restrict__ far_ptr,
                                                     const uint32_t* __restrict__ far_col, const float* __restrict__ coef,
                                                     int nk, const float* __restrict__ up, float* __restrict__ farv) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int a = t / C, c = t % C;
  if (a >= nn) return;
  const int k = c < H ? 0 : nk - 1;
  float acc = 0.f;
  for (uint32_t q = far_ptr[a]; q < far_ptr[a + 1]; ++q) acc += coef[(size_t)q * nk + k] * up[(size_t)far_col[q] * C + c];
  farv[(size_t)a * C + c] = acc;
}

template <int C, int H>
__global__ __launch_bounds__(256) void bt_rows_kernel(int nl, const uint32_t* __restrict__ leaves,
                                                      const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ chain_ptr,
                                                      const uint32_t* __restrict__ chain,
                                                      const uint32_t* __restrict__ near_ptr,
                                                      const uint32_t* __restrict__ near_col,
                                                      const uint64_t* __restrict__ near_off,
                                                      const float* __restrict__ near_val, int nk,
                                                      const float* __restrict__ farv, const float* __restrict__ psi,
                                                      float* __restrict__ hx) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= nl) return;
  const int r = lane >> 3, j = lane & 7;
  const uint32_t L = leaves[k];
  const uint32_t s = start[L];
  const int ns = (int)(end[L] - s);
  float acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.f;
  const uint32_t c0 = chain_ptr[k], c1 = chain_ptr[k + 1];
  for (uint32_t ci = c0; ci < c1; ++ci) {
    const uint32_t a = chain[ci];
    const uint32_t row = s - start[a] + r;
    for (uint32_t q = near_ptr[a]; q < near_ptr[a + 1]; ++q) {
      const uint32_t b = near_col[q];
      const uint32_t sb = start[b];
      const int cb = (int)(end[b] - sb);
      const uint64_t off = near_off[q] + (uint64_t)row * cb;
      if (r < ns) {
        for (int c = j; c < cb; c += 8) {
          if (a == b && (uint32_t)c == row) continue;
          const float* v = near_val + (off + c) * nk;
          const float v0 = v[0], v1 = v[nk - 1];
          const float* ps = psi + (size_t)(sb + c) * C;
#pragma unroll
          for (int ch = 0; ch < C; ++ch) acc[ch] += (ch < H ? v0 : v1) * ps[ch];
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = jsum8(acc[c]);
  if (j == 0 && r < ns) {
    float g[C];
#pragma unroll
    for (int c = 0; c < C; ++c) g[c] = 0.f;
    for (uint32_t ci = c0; ci < c1; ++ci) {
      const float* f = farv + (size_t)chain[ci] * C;
#pragma unroll
      for (int c = 0; c < C; ++c) g[c] += f[c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) hx[(size_t)(s + r) * C + c] = g[c] + acc[c];
  }
}

void BlockTreeDev::upload(ddgrs_ctx* ctx, const Topology& t, const BlockStructure& s, int nk_) {
  nk = nk_;
  n = (int)t.order.size();
  nn = (int)t.n_nodes();
  nl = (int)t.leaves.size();
  near_total = s.near_total;
  ddgrs::upload(ctx, order, t.order);
  std::vector<uint32_t> pos(n);
  for (int p = 0; p < n; ++p) pos[t.order[p]] = (uint32_t)p;
  ddgrs::upload(ctx, position, pos);
  ddgrs::upload(ctx, start, t.start);
  ddgrs::upload(ctx, end, t.end);
  ddgrs::upload(ctx, leaves, t.leaves);
  ddgrs::upload(ctx, chain_ptr, s.chain_ptr);
  ddgrs::upload(ctx, chain, s.chain);
  ddgrs::upload(ctx, far_ptr, s.far_ptr);
  ddgrs::upload(ctx, far_col, s.far_col);
  std::vector<uint32_t> far_row(s.far_col.size());
  for (int a = 0; a < nn; ++a)
    for (uint32_t q = s.far_ptr[a]; q < s.far_ptr[a + 1]; ++q) far_row[q] = (uint32_t)a;
  ddgrs::upload(ctx, this->far_row, far_row);
  ddgrs::upload(ctx, near_ptr, s.near_ptr);
  ddgrs::upload(ctx, near_col, s.near_col);
  ddgrs::upload(ctx, near_row, s.near_row);
  ddgrs::upload(ctx, near_off, s.near_off);
  std::vector<uint32_t> nd;
  for (int a = 0; a < nn; ++a)
    if (s.far_needed[a]) nd.push_back((uint32_t)a);
  n_needed = (int)nd.size();
  ddgrs::upload(ctx, needed, nd);
  n_far = (uint32_t)s.far_col.size();
  n_near_blocks = (uint32_t)s.near_col.size();
  far_coef.reserve((size_t)std::max<uint32_t>(n_far, 1) * nk);
  near_val.reserve((size_t)std::max<uint64_t>(near_total, 1) * nk);
}

// Kernel names carry the channel count (timing per product kind).
#define BT_NAME(base) (C == 12 ? base "_c12" : C == 3 ? base "_c3" : C == 2 ? base "_c2" : base "_c1")

template <int C, int H>
void BlockTreeDev::multiply(ddgrs_ctx* ctx, const float* psi) {
  up.reserve((size_t)nn * C);
  farv.reserve((size_t)nn * C);
  hx.reserve((size_t)n * C);
  launch(ctx, BT_NAME("bt_up"), bt_up_kernel<C>, dim3(blocks_for(n_needed, 4)), dim3(256), 0, n_needed,
         (const uint32_t*)needed.ptr, (const uint32_t*)start.ptr, (const uint32_t*)end.ptr, psi, up.ptr);
  launch(ctx, BT_NAME("bt_far"), bt_far_kernel<C, H>, dim3(blocks_for((int64_t)nn * C, 256)), dim3(256), 0, nn,
         (const uint32_t*)far_ptr.ptr, (const uint32_t*)far_col.ptr, (const float*)far_coef.ptr, nk,
         (const float*)up.ptr, farv.ptr);
  launch(ctx, BT_NAME("bt_rows"), bt_rows_kernel<C, H>, dim3(blocks_for(nl, 4)), dim3(256), 0, nl, (const uint32_t*)leaves.ptr,
         (const uint32_t*)start.ptr, (const uint32_t*)end.ptr, (const uint32_t*)chain_ptr.ptr,
         (const uint32_t*)chain.ptr, (const uint32_t*)near_ptr.ptr, (const uint32_t*)near_col.ptr,
         (const uint64_t*)near_off.ptr, (const float*)near_val.ptr, nk, (const float*)farv.ptr, psi, hx.ptr);
}

template void BlockTreeDev::multiply<12, 9>(ddgrs_ctx*, const float*);
template void BlockTreeDev::multiply<2, 1>(ddgrs_ctx*, const float*);
template void BlockTreeDev::multiply<3, 3>(ddgrs_ctx*, const float*);
template void BlockTreeDev::multiply<1, 1>(ddgrs_ctx*, const float*);

// ---- face metric ---------------------------------------------------------------------------

#pragma clang fp contract(off)

struct NodeGeo {
  const double *cen, *proj;  // 3, 6 per node
};

// Far coefficients [ρ^{−2s}, ½(vᵀP̄_I v + vᵀP̄_J v) ρ^{−4} ρ^{−2s}] at the centroids, (min, max)
// ordered (hierarchical.rs:185 hi_lo, :206 far).
__global__ __launch_bounds__(256) void fm_far_coef_kernel(uint32_t nfar, const uint32_t* __restrict__ far_row,
                                                          const uint32_t* __restrict__ far_col, NodeGeo g, double neg_s,
                                                          float* __restrict__ coef) {
  uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nfar) return;
  uint32_t a = far_row[q], b = far_col[q];
  if (a > b) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = g.cen[3 * a + k] - g.cen[3 * b + k];
  auto pq = [&](uint32_t n) {
    const double* p = g.proj + 6 * n;
    return p[0] * d[0] * d[0] + p[3] * d[1] * d[1] + p[5] * d[2] * d[2] +
           2.0 * (p[1] * d[0] * d[1] + p[2] * d[0] * d[2] + p[4] * d[1] * d[2]);
  };
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double hi = pow(r2, neg_s);
  double lo = 0.5 * (pq(a) + pq(b)) / (r2 * r2) * hi;
  coef[2 * (size_t)q] = (float)hi;
  coef[2 * (size_t)q + 1] = (float)lo;
}

// Near entries of one block per wave, lane = r·cb + c (≤ 64 entries per block).
__global__ __launch_bounds__(256) void fm_near_kernel(uint32_t nblocks, const uint32_t* __restrict__ near_row,
                                                      const uint32_t* __restrict__ near_col,
                                                      const uint64_t* __restrict__ near_off,
                                                      const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ order, const double* __restrict__ bary,
                                                      const double* __restrict__ nrm, double neg_s,
                                                      float* __restrict__ val) {
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nblocks) return;
  const uint32_t a = near_row[q], b = near_col[q];
  const uint32_t ra = end[a] - start[a], cb = end[b] - start[b];
  if (lane >= ra * cb) return;
  const uint32_t r = lane / cb, c = lane % cb;
  uint32_t i = order[start[a] + r], j = order[start[b] + c];
  float* out = val + 2 * (near_off[q] + lane);
  if (i == j) {
    out[0] = 0.f;
    out[1] = 0.f;
    return;
  }
  if (i > j) {
    uint32_t t = i;
    i = j;
    j = t;
  }
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = bary[3 * i + k] - bary[3 * j + k];
  double ns = nrm[3 * i] * d[0] + nrm[3 * i + 1] * d[1] + nrm[3 * i + 2] * d[2];
  double nt = nrm[3 * j] * d[0] + nrm[3 * j + 1] * d[1] + nrm[3 * j + 2] * d[2];
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double hi = pow(r2, neg_s);
  double lo = 0.5 * (ns * ns + nt * nt) / (r2 * r2) * hi;
  out[0] = (float)hi;
  out[1] = (float)lo;
}

// Per BVH position: D_f coefficients g_c = (n/|n|²) × e_opp(c) (RS/src/metric/mod.rs:120), area,
// vertices, and ψ = (a, a) for H̃a.
__global__ __launch_bounds__(256) void fm_face_kernel(int nf, const uint32_t* __restrict__ order,
                                                      const uint32_t* __restrict__ faces, const double* __restrict__ pos,
                                                      const double* __restrict__ area, float* __restrict__ df,
                                                      float* __restrict__ area_p, uint32_t* __restrict__ vert_p,
                                                      float* __restrict__ psi2) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nf) return;
  uint32_t s = order[p];
  double q[3][3];
  for (int c = 0; c < 3; ++c) {
    uint32_t v = faces[3 * s + c];
    vert_p[3 * p + c] = v;
    for (int k = 0; k < 3; ++k) q[c][k] = pos[3 * v + k];
  }
  double e1[3], e2[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = q[1][k] - q[0][k];
    e2[k] = q[2][k] - q[0][k];
  }
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double inv = 1.0 / (len * len);
  double w[3] = {n[0] * inv, n[1] * inv, n[2] * inv};
  for (int c = 0; c < 3; ++c) {
    int i0 = (c + 1) % 3, i1 = (c + 2) % 3;
    double e[3] = {q[i1][0] - q[i0][0], q[i1][1] - q[i0][1], q[i1][2] - q[i0][2]};
    double g[3] = {w[1] * e[2] - w[2] * e[1], w[2] * e[0] - w[0] * e[2], w[0] * e[1] - w[1] * e[0]};
    for (int k = 0; k < 3; ++k) df[9 * p + 3 * c + k] = (float)g[k];
  }
  float a = (float)area[s];
  area_p[p] = a;
  psi2[2 * p] = a;
  psi2[2 * p + 1] = a;
}

#pragma clang fp contract(fast)

// m = M x (9 gradient channels 3k + r, then 3 average channels), ψ = a m, per BVH position.
__global__ __launch_bounds__(256) void fm_m_kernel(int nf, const uint32_t* __restrict__ vert_p,
                                                   const float* __restrict__ df, const float* __restrict__ area_p,
                                                   const double* __restrict__ x, float* __restrict__ m,
                                                   float* __restrict__ psi) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nf) return;
  float xv[3][3];
  for (int c = 0; c < 3; ++c) {
    uint32_t v = vert_p[3 * p + c];
    for (int k = 0; k < 3; ++k) xv[c][k] = (float)x[3 * v + k];
  }
  const float* g = df + 9 * p;
  float out[12];
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) out[3 * k + r] = xv[0][k] * g[r] + xv[1][k] * g[3 + r] + xv[2][k] * g[6 + r];
  for (int k = 0; k < 3; ++k) out[9 + k] = (xv[0][k] + xv[1][k] + xv[2][k]) * (1.f / 3.f);
  float a = area_p[p];
  for (int c = 0; c < 12; ++c) {
    m[12 * (size_t)p + c] = out[c];
    psi[12 * (size_t)p + c] = a * out[c];
  }
}

// z = w (H̃a·m − H̃ψ) per BVH position and channel.
__global__ __launch_bounds__(256) void fm_z_kernel(int nf, const float* __restrict__ ha, const float* __restrict__ m,
                                                   const float* __restrict__ hx, float w_hi, float w_lo,
                                                   float* __restrict__ z) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nf * 12) return;
  int p = t / 12, c = t % 12;
  int kern = c < 9 ? 0 : 1;
  float w = c < 9 ? w_hi : w_lo;
  z[t] = w * (ha[2 * p + kern] * m[t] - hx[t]);
}

// y_v = 2 Σ_{S∋v} a_S (Mᵀ row of S at v)·z_S (faces ascending; FP64 accumulation).
__global__ __launch_bounds__(256) void fm_gather_kernel(int nv, const uint32_t* __restrict__ inc_ptr,
                                                        const uint32_t* __restrict__ inc, const float* __restrict__ df,
                                                        const float* __restrict__ area_p, const float* __restrict__ z,
                                                        double* __restrict__ y) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t q = inc_ptr[v]; q < inc_ptr[v + 1]; ++q) {
    uint32_t e = inc[q];
    uint32_t p = e >> 2, c = e & 3u;
    float a = area_p[p];
    const float* g = df + 9 * (size_t)p + 3 * c;
    const float* zs = z + 12 * (size_t)p;
    for (int k = 0; k < 3; ++k) {
      float t = a * (g[0] * zs[3 * k] + g[1] * zs[3 * k + 1] + g[2] * zs[3 * k + 2]) + a * (1.f / 3.f) * zs[9 + k];
      acc[k] += (double)t;
    }
  }
  for (int k = 0; k < 3; ++k) y[3 * (size_t)v + k] = 2.0 * acc[k];
}

void FaceMetricDev::build(ddgrs_surface* s, const Topology& t, const Clusters& cl, double chi, size_t leaf_pairs,
                          double s_exp) {
  ddgrs_ctx* ctx = s->ctx;
  topo = t;
  nv = (int)s->nv;
  nf = (int)s->nf;
  bs = build_block_structure(t, cl, chi, leaf_pairs);
  bt.upload(ctx, t, bs, 2);
  // node geometry (build-time centroids and projectors)
  DevBuf<double> cen, proj;
  ddgrs::upload(ctx, cen, cl.centroid);
  ddgrs::upload(ctx, proj, cl.projector);
  const double neg_s = -s_exp;
  launch(ctx, "fm_far_coef", fm_far_coef_kernel, dim3(blocks_for(bt.n_far, 256)), dim3(256), 0, bt.n_far,
         (const uint32_t*)bt.far_row.ptr, (const uint32_t*)bt.far_col.ptr, NodeGeo{cen.ptr, proj.ptr}, neg_s,
         bt.far_coef.ptr);
  launch(ctx, "fm_near", fm_near_kernel, dim3(blocks_for(bt.n_near_blocks, 4)), dim3(256), 0, bt.n_near_blocks,
         (const uint32_t*)bt.near_row.ptr, (const uint32_t*)bt.near_col.ptr, (const uint64_t*)bt.near_off.ptr,
         (const uint32_t*)bt.start.ptr, (const uint32_t*)bt.end.ptr, (const uint32_t*)bt.order.ptr,
         (const double*)s->bary.ptr, (const double*)s->normal.ptr, neg_s, bt.near_val.ptr);
  df.reserve(9 * (size_t)nf);
  area_p.reserve(nf);
  vert_p.reserve(3 * (size_t)nf);
  DevBuf<float> psi2(2 * (size_t)nf);
  launch(ctx, "fm_face", fm_face_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, (const uint32_t*)bt.order.ptr,
         (const uint32_t*)s->faces.ptr, (const double*)s->positions.ptr, (const double*)s->area.ptr, df.ptr, area_p.ptr,
         vert_p.ptr, psi2.ptr);
  bt.multiply<2, 1>(ctx, psi2.ptr);
  ha.reserve(2 * (size_t)nf);
  HIPCHECK(hipMemcpyAsync(ha.ptr, bt.hx.ptr, 2 * (size_t)nf * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  // vertex → (BVH position of the face, corner), faces ascending
  std::vector<uint32_t> position(nf);
  for (int p = 0; p < nf; ++p) position[t.order[p]] = (uint32_t)p;
  std::vector<uint32_t> ptr(nv + 1, 0), inc_h(3 * (size_t)nf);
  for (int f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) ptr[s->faces_h[3 * f + c] + 1]++;
  for (int v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
  std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) inc_h[fill[s->faces_h[3 * f + c]]++] = (position[f] << 2) | (uint32_t)c;
  ddgrs::upload(ctx, inc_ptr, ptr);
  ddgrs::upload(ctx, inc, inc_h);
  m.reserve(12 * (size_t)nf);
  psi.reserve(12 * (size_t)nf);
  z.reserve(12 * (size_t)nf);
  ctx->sync();  // the local buffers above are released on return
}

void FaceMetricDev::apply(ddgrs_ctx* ctx, const double* x, double* y) {
  launch(ctx, "fm_m", fm_m_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, (const uint32_t*)vert_p.ptr,
         (const float*)df.ptr, (const float*)area_p.ptr, x, m.ptr, psi.ptr);
  bt.multiply<12, 9>(ctx, psi.ptr);
  launch(ctx, "fm_z", fm_z_kernel, dim3(blocks_for((int64_t)nf * 12, 256)), dim3(256), 0, nf, (const float*)ha.ptr,
         (const float*)m.ptr, (const float*)bt.hx.ptr, (float)w_hi, (float)w_lo, z.ptr);
  launch(ctx, "fm_gather", fm_gather_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)inc_ptr.ptr,
         (const uint32_t*)inc.ptr, (const float*)df.ptr, (const float*)area_p.ptr, (const float*)z.ptr, y);
}

// ---- vertex-pair fractional operator ------------------------------------------------------

#pragma clang fp contract(off)

__global__ __launch_bounds__(256) void vf_far_coef_kernel(uint32_t nfar, const uint32_t* __restrict__ far_row,
                                                          const uint32_t* __restrict__ far_col,
                                                          const double* __restrict__ cen, double ex,
                                                          float* __restrict__ coef) {
  uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nfar) return;
  uint32_t a = far_row[q], b = far_col[q];
  if (a > b) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = cen[3 * a + k] - cen[3 * b + k];
  coef[q] = (float)pow(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], ex);
}

__global__ __launch_bounds__(256) void vf_near_kernel(uint32_t nblocks, const uint32_t* __restrict__ near_row,
                                                      const uint32_t* __restrict__ near_col,
                                                      const uint64_t* __restrict__ near_off,
                                                      const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ order, const double* __restrict__ pos,
                                                      double ex, float* __restrict__ val) {
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nblocks) return;
  const uint32_t a = near_row[q], b = near_col[q];
  const uint32_t ra = end[a] - start[a], cb = end[b] - start[b];
  if (lane >= ra * cb) return;
  const uint32_t r = lane / cb, c = lane % cb;
  uint32_t i = order[start[a] + r], j = order[start[b] + c];
  if (i == j) {
    val[near_off[q] + lane] = 0.f;
    return;
  }
  if (i > j) {
    uint32_t t = i;
    i = j;
    j = t;
  }
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = pos[3 * i + k] - pos[3 * j + k];
  val[near_off[q] + lane] = (float)pow(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], ex);
}

#pragma clang fp contract(fast)

__global__ __launch_bounds__(256) void vf_psi_kernel(int n, const uint32_t* __restrict__ order,
                                                     const float* __restrict__ mass_p, const double* __restrict__ x,
                                                     float* __restrict__ psi) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  uint32_t v = order[p];
  float a = mass_p[p];
  for (int k = 0; k < 3; ++k) psi[3 * (size_t)p + k] = a * (float)x[3 * (size_t)v + k];
}

__global__ __launch_bounds__(256) void vf_out_kernel(int n, const uint32_t* __restrict__ position,
                                                     const double* __restrict__ mass, const float* __restrict__ ha,
                                                     const float* __restrict__ hx, const double* __restrict__ x,
                                                     double* __restrict__ y) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  uint32_t p = position[v];
  double a = mass[v];
  float h = ha[p];
  for (int k = 0; k < 3; ++k) {
    float t = h * (float)x[3 * (size_t)v + k] - hx[3 * (size_t)p + k];
    y[3 * (size_t)v + k] = 2.0 * a * (double)t;
  }
}

void VertexFractionalDev::build(ddgrs_ctx* ctx, const std::vector<double>& pos, const std::vector<double>& m, double tau,
                                double chi, size_t leaf_size, size_t leaf_pairs) {
  nv = (int)m.size();
  // vertex BVH: point f_v, box the point, mass a_v (hierarchical.rs:1017-1060)
  topo = build_topology(nv, pos.data(), pos.data(), pos.data(), leaf_size);
  clusters = build_clusters(topo, pos.data(), pos.data(), pos.data(), m.data());
  bs = build_block_structure(topo, clusters, chi, leaf_pairs);
  bt.upload(ctx, topo, bs, 1);
  const double ex = -(tau + 1.0);
  DevBuf<double> cen, dpos;
  ddgrs::upload(ctx, cen, clusters.centroid);
  ddgrs::upload(ctx, dpos, pos);
  launch(ctx, "vf_far_coef", vf_far_coef_kernel, dim3(blocks_for(bt.n_far, 256)), dim3(256), 0, bt.n_far,
         (const uint32_t*)bt.far_row.ptr, (const uint32_t*)bt.far_col.ptr, (const double*)cen.ptr, ex, bt.far_coef.ptr);
  launch(ctx, "vf_near", vf_near_kernel, dim3(blocks_for(bt.n_near_blocks, 4)), dim3(256), 0, bt.n_near_blocks,
         (const uint32_t*)bt.near_row.ptr, (const uint32_t*)bt.near_col.ptr, (const uint64_t*)bt.near_off.ptr,
         (const uint32_t*)bt.start.ptr, (const uint32_t*)bt.end.ptr, (const uint32_t*)bt.order.ptr,
         (const double*)dpos.ptr, ex, bt.near_val.ptr);
  std::vector<float> mp(nv);
  for (int p = 0; p < nv; ++p) mp[p] = (float)m[topo.order[p]];
  ddgrs::upload(ctx, mass_p, mp);
  ddgrs::upload(ctx, mass, m);
  bt.multiply<1, 1>(ctx, mass_p.ptr);
  ha.reserve(nv);
  HIPCHECK(hipMemcpyAsync(ha.ptr, bt.hx.ptr, (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  psi.reserve(3 * (size_t)nv);
  ctx->sync();
}

void VertexFractionalDev::apply(ddgrs_ctx* ctx, const double* x, double* y) {
  launch(ctx, "vf_psi", vf_psi_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)bt.order.ptr,
         (const float*)mass_p.ptr, x, psi.ptr);
  bt.multiply<3, 3>(ctx, psi.ptr);
  launch(ctx, "vf_out", vf_out_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)bt.position.ptr,
         (const double*)mass.ptr, (const float*)ha.ptr, (const float*)bt.hx.ptr, x, y);
}

}  // namespace ddgrs
