// The learned right preconditioner (see learned.hpp): features, the network's forward and backward
// passes, the loss, Adam and Q x, as kernels; pairs and training control on the host.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "capi_util.hpp"
#include "learned.hpp"

namespace ddgrs {

LearnedLayout learned_layout() {
  LearnedLayout L{};
  size_t o = 0;
  L.w0 = o;
  o += (size_t)kLnFeat * kLnWidth;
  L.b0 = o;
  o += kLnWidth;
  for (int l = 0; l < kLnLayers; ++l) {
    L.w[l] = o;
    o += (size_t)2 * kLnWidth * kLnWidth;
    L.b[l] = o;
    o += kLnWidth;
  }
  L.wv = o;
  o += (size_t)kLnWidth * kLnOut;
  L.bv = o;
  o += kLnOut;
  L.we1 = o;
  o += (size_t)kLnWidth * kLnWidth;
  L.be1 = o;
  o += kLnWidth;
  L.we2 = o;
  o += (size_t)kLnWidth * kLnOut;
  L.be2 = o;
  o += kLnOut;
  L.total = o;
  return L;
}

namespace {

// Y[v][j] = Σ_i X[v][i] M(i, j) (+ b[j]), X = [X1 (I1 columns) | X2 (I − I1)], M(i, j) =
// M[i msi + j msj]; `relu`: Y = max(0, ·) with the pre-activation in `pre`.  M in LDS; a workgroup
// takes 64 rows, four at a time (a thread per (row, column)).
template <int I, int J>
__global__ __launch_bounds__(256) void ln_rows_kernel(int n, const float* __restrict__ X1, int ld1, int I1,
                                                      const float* __restrict__ X2, int ld2,
                                                      const float* __restrict__ M, int msi, int msj,
                                                      const float* __restrict__ bias, int relu, float* __restrict__ pre,
                                                      float* __restrict__ Y, int ldy) {
  __shared__ float sm[I * J];
  __shared__ float sx[4][I];
  const int tid = threadIdx.x, j = tid & 63, r = tid >> 6;
  for (int e = tid; e < I * J; e += 256) sm[e] = M[(size_t)(e / J) * msi + (size_t)(e % J) * msj];
  const int v1 = min(n, (int)blockIdx.x * 64 + 64);
  for (int v0 = blockIdx.x * 64; v0 < v1; v0 += 4) {
    __syncthreads();
    for (int e = tid; e < 4 * I; e += 256) {
      const int rr = e / I, i = e % I, v = v0 + rr;
      sx[rr][i] = v < n ? (i < I1 ? X1[(size_t)v * ld1 + i] : X2[(size_t)v * ld2 + (i - I1)]) : 0.f;
    }
    __syncthreads();
    const int v = v0 + r;
    if (v < n && j < J) {
      float acc = bias ? bias[j] : 0.f;
#pragma unroll 8
      for (int i = 0; i < I; ++i) acc = fmaf(sx[r][i], sm[i * J + j], acc);
      if (pre) pre[(size_t)v * ldy + j] = acc;
      Y[(size_t)v * ldy + j] = relu ? fmaxf(acc, 0.f) : acc;
    }
  }
}

// The same product for J = 64 on the matrix cores, BF16 inputs with FP32 accumulation (MFMA
// 16×16×16 bf16): X and M are rounded to BF16 (nearest, ties to even) as they are staged in LDS; a
// workgroup takes 64 rows in two passes of 32; per pass wave w owns the 16-row group w & 1 and the
// two 16-column tiles 2 (w >> 1) and 2 (w >> 1) + 1.  M (K × 64, K = I rounded up to 16, zero rows
// appended) and the pass's 32 rows of X (row stride K + 4) are staged in LDS.  Lane l supplies
// A(l % 16, k₀ + 4 (l / 16) + q) and B(k₀ + 4 (l / 16) + q, 16 t + l % 16), q < 4, and holds
// D(4 (l / 16) + v, l % 16).
using f32x4 = __attribute__((__vector_size__(4 * sizeof(float)))) float;
using bf16x4 = __attribute__((__vector_size__(4 * sizeof(short)))) short;

// FP32 → BF16, round to nearest, ties to even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int I>
__global__ __launch_bounds__(256) void ln_rows_mfma_kernel(int n, const float* __restrict__ X1, int ld1, int I1,
                                                           const float* __restrict__ X2, int ld2,
                                                           const float* __restrict__ M, int msi, int msj,
                                                           const float* __restrict__ bias, int relu,
                                                           float* __restrict__ pre, float* __restrict__ Y, int ldy) {
  constexpr int K = (I + 15) / 16 * 16, J = 64, SX = K + 4;
  __shared__ uint16_t sm[K * J];
  __shared__ uint16_t sx[32 * SX];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int e = tid; e < K * J; e += 256) {
    const int i = e / J, j = e % J;
    sm[e] = i < I ? to_bf16(M[(size_t)i * msi + (size_t)j * msj]) : (uint16_t)0;
  }
  const int rg = w & 1, c0 = 32 * (w >> 1), li = lane & 15, lk = lane >> 4;
  for (int p = 0; p < 2; ++p) {
    const int v0 = blockIdx.x * 64 + 32 * p;
    __syncthreads();
    for (int e = tid; e < 32 * K; e += 256) {
      const int rr = e / K, i = e % K, v = v0 + rr;
      float x = 0.f;
      if (v < n && i < I) x = i < I1 ? X1[(size_t)v * ld1 + i] : X2[(size_t)v * ld2 + (i - I1)];
      sx[rr * SX + i] = to_bf16(x);
    }
    __syncthreads();
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 16) {
      bf16x4 a, b0, b1;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = k0 + 4 * lk + q;
        a[q] = (short)sx[(16 * rg + li) * SX + k];
        b0[q] = (short)sm[k * J + c0 + li];
        b1[q] = (short)sm[k * J + c0 + 16 + li];
      }
      acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int j = c0 + 16 * t + li;
      const float bj = bias ? bias[j] : 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int v = v0 + 16 * rg + 4 * lk + q;
        if (v < n) {
          const float y = (t ? acc1[q] : acc0[q]) + bj;
          if (pre) pre[(size_t)v * ldy + j] = y;
          Y[(size_t)v * ldy + j] = relu ? fmaxf(y, 0.f) : y;
        }
      }
    }
  }
}

// Partial weight gradients of one workgroup's rows: P[i][j] = Σ_v A[v][i] B[v][j] for i < I and,
// as row I, the bias gradient Σ_v B[v][j]; A = [A1 (I1 columns) | A2].  Rows in chunks of 32 staged
// in LDS; each thread keeps its outputs (t, t + 256, …) in registers.
template <int I, int J>
__global__ __launch_bounds__(256) void ln_wgrad_kernel(int n, int rows_per_block, const float* __restrict__ A1, int lda1,
                                                       int I1, const float* __restrict__ A2, int lda2,
                                                       const float* __restrict__ B, int ldb, float* __restrict__ part) {
  constexpr int NO = (I + 1) * J, NQ = (NO + 255) / 256;
  __shared__ float sa[32][I + 1];
  __shared__ float sb[32][J];
  const int tid = threadIdx.x;
  float acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.f;
  const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
  for (int c0 = r0; c0 < r1; c0 += 32) {
    __syncthreads();
    for (int e = tid; e < 32 * (I + 1); e += 256) {
      const int rr = e / (I + 1), i = e % (I + 1), v = c0 + rr;
      float x = 0.f;
      if (v < r1) x = i == I ? 1.f : (i < I1 ? A1[(size_t)v * lda1 + i] : A2[(size_t)v * lda2 + (i - I1)]);
      sa[rr][i] = x;
    }
    for (int e = tid; e < 32 * J; e += 256) {
      const int rr = e / J, j = e % J, v = c0 + rr;
      sb[rr][j] = v < r1 ? B[(size_t)v * ldb + j] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int o = tid + 256 * q;
      if (o < NO) {
        const int i = o / J, j = o % J;
        float s = acc[q];
        for (int rr = 0; rr < 32; ++rr) s = fmaf(sa[rr][i], sb[rr][j], s);
        acc[q] = s;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int o = tid + 256 * q;
    if (o < NO) part[(size_t)blockIdx.x * NO + o] = acc[q];
  }
}

// The same partial gradients for J = 64 on the matrix cores, BF16 inputs with FP32 accumulation
// (MFMA 16×16×16 bf16; A and B rounded to BF16 as they are staged): the output (I + 1) × 64, rows
// padded to 16·NT, is NT × 4 tiles; wave w owns column tile w and all NT row tiles.  Per step of 16
// rows of the chunk, lane l supplies Aᵀ(16 t + l % 16, k₀ + 4 (l / 16) + q) =
// A(k₀ + 4 (l / 16) + q, 16 t + l % 16) and B(k₀ + 4 (l / 16) + q, 16 w + l % 16), q < 4.
template <int I>
__global__ __launch_bounds__(256) void ln_wgrad_mfma_kernel(int n, int rows_per_block, const float* __restrict__ A1,
                                                            int lda1, int I1, const float* __restrict__ A2, int lda2,
                                                            const float* __restrict__ B, int ldb,
                                                            float* __restrict__ part) {
  constexpr int J = 64, NT = (I + 16) / 16, SA = 16 * NT, NO = (I + 1) * J;
  __shared__ uint16_t sa[32][SA + 2];
  __shared__ uint16_t sb[32][J + 2];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
  for (int c0 = r0; c0 < r1; c0 += 32) {
    __syncthreads();
    for (int e = tid; e < 32 * SA; e += 256) {
      const int rr = e / SA, i = e % SA, v = c0 + rr;
      float x = 0.f;
      if (v < r1 && i <= I) x = i == I ? 1.f : (i < I1 ? A1[(size_t)v * lda1 + i] : A2[(size_t)v * lda2 + (i - I1)]);
      sa[rr][i] = to_bf16(x);
    }
    for (int e = tid; e < 32 * J; e += 256) {
      const int rr = e / J, j = e % J, v = c0 + rr;
      sb[rr][j] = to_bf16(v < r1 ? B[(size_t)v * ldb + j] : 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 16) {
      bf16x4 b;
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = (short)sb[k0 + 4 * lk + q][16 * w + li];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        bf16x4 a;
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = (short)sa[k0 + 4 * lk + q][16 * t + li];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = 16 * t + 4 * lk + q;
      if (i <= I) part[(size_t)blockIdx.x * NO + (size_t)i * J + 16 * w + li] = acc[t][q];
    }
}

// grad[w + i J + j] += Σ_b part[b][i][j] (i < I), grad[bias + j] += Σ_b part[b][I][j].  A workgroup
// takes 64 outputs; its four waves sum the blocks b ≡ wave (mod 4) in order, then the four sums are
// added in wave order (a fixed summation tree).
__global__ __launch_bounds__(256) void ln_wgrad_sum_kernel(int nb, int I, int J, const float* __restrict__ part,
                                                           float* __restrict__ gw, float* __restrict__ gb) {
  __shared__ float red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, NO = (I + 1) * J;
  const int o = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (o < NO)
    for (int b = w; b < nb; b += 4) s += part[(size_t)b * NO + o];
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && o < NO) {
    const float t = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (o < I * J)
      gw[o] += t;
    else
      gb[o - I * J] += t;
  }
}

// m[v][c] = mean of h[u][c] over the neighbours u of v.
__global__ __launch_bounds__(256) void ln_mean_kernel(int n, const uint32_t* __restrict__ ptr,
                                                      const uint32_t* __restrict__ nb, const float* __restrict__ h,
                                                      float* __restrict__ m) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * kLnWidth) return;
  const int v = t / kLnWidth, c = t % kLnWidth;
  const uint32_t q0 = ptr[v], q1 = ptr[v + 1];
  float s = 0.f;
  for (uint32_t q = q0; q < q1; ++q) s += h[(size_t)nb[q] * kLnWidth + c];
  m[t] = q1 > q0 ? s / (float)(q1 - q0) : 0.f;
}

// dh[u][c] += Σ_{v ∈ N(u)} dm[v][c] / deg v (the neighbourhoods are symmetric).
__global__ __launch_bounds__(256) void ln_mean_t_kernel(int n, const uint32_t* __restrict__ ptr,
                                                        const uint32_t* __restrict__ nb, const float* __restrict__ dm,
                                                        float* __restrict__ dh) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * kLnWidth) return;
  const int u = t / kLnWidth, c = t % kLnWidth;
  float s = 0.f;
  for (uint32_t q = ptr[u]; q < ptr[u + 1]; ++q) {
    const uint32_t v = nb[q];
    s += dm[(size_t)v * kLnWidth + c] / (float)(ptr[v + 1] - ptr[v]);
  }
  dh[t] += s;
}

__global__ __launch_bounds__(256) void ln_relu_bwd_kernel(int n, const float* __restrict__ pre, float* __restrict__ d) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < n && !(pre[t] > 0.f)) d[t] = 0.f;
}

__global__ __launch_bounds__(256) void ln_fill_kernel(size_t n, float* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] = 0.f;
}

// q_v = s (B_v r_v + Σ_u ½ (g_v + g_u) r_u); d_v = 2 (q_v − e_v) / ‖e‖² into qd; per workgroup
// Σ |q − e|² into lsum (FP64).
// q_v = B_v r_v + Σ_{u ∈ N(v)} E_uv r_u; qd = 2 (s q − e) / |e|², loss partial sums per block.
__global__ __launch_bounds__(256) void ln_loss_kernel(int n, const uint32_t* __restrict__ ptr,
                                                      const uint32_t* __restrict__ nb, const uint32_t* __restrict__ emap,
                                                      const float* __restrict__ o, const float* __restrict__ eo,
                                                      const float* __restrict__ r, const float* __restrict__ e,
                                                      float s, float inv_e2, float* __restrict__ qd,
                                                      double* __restrict__ lsum) {
  __shared__ double red[256];
  const int v = blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (v < n) {
    const float* ov = o + (size_t)v * kLnOut;
    float q[3];
    for (int a = 0; a < 3; ++a) {
      float acc = 0.f;
      for (int b = 0; b < 3; ++b) acc = fmaf(ov[3 * a + b], r[3 * v + b], acc);
      q[a] = acc;
    }
    for (uint32_t k = ptr[v]; k < ptr[v + 1]; ++k) {
      const uint32_t u = nb[k];
      const float* E = eo + (size_t)emap[k] * kLnOut;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) q[a] = fmaf(E[3 * a + b], r[3 * u + b], q[a]);
    }
    for (int a = 0; a < 3; ++a) {
      const float df = s * q[a] - e[3 * v + a];
      qd[3 * v + a] = 2.f * df * inv_e2;
      l += (double)df * (double)df;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lsum[blockIdx.x] = red[0];
}

// dB_v = s d_v ⊗ r_v.
__global__ __launch_bounds__(256) void ln_dout_kernel(int n, const float* __restrict__ r, const float* __restrict__ qd,
                                                      float s, float* __restrict__ dout) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  float* d = dout + (size_t)v * kLnOut;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) d[3 * a + b] = s * qd[3 * (size_t)v + a] * r[3 * (size_t)v + b];
}

// dE_uv = s (d_u ⊗ r_v + d_v ⊗ r_u) per edge {u, v}.
__global__ __launch_bounds__(256) void ln_edge_dout_kernel(int ne, const uint32_t* __restrict__ edge,
                                                           const float* __restrict__ r, const float* __restrict__ qd,
                                                           float s, float* __restrict__ dout) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= ne) return;
  const size_t u = edge[2 * (size_t)k], v = edge[2 * (size_t)k + 1];
  float* d = dout + (size_t)k * kLnOut;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) d[3 * a + b] = s * (qd[3 * u + a] * r[3 * v + b] + qd[3 * v + a] * r[3 * u + b]);
}

// The edge head's input h_u + h_v (64 channels) per edge.
__global__ __launch_bounds__(256) void ln_edge_in_kernel(int ne, const uint32_t* __restrict__ edge,
                                                         const float* __restrict__ h, float* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)ne * kLnWidth) return;
  const size_t k = t / kLnWidth, c = t % kLnWidth;
  x[t] = h[(size_t)edge[2 * k] * kLnWidth + c] + h[(size_t)edge[2 * k + 1] * kLnWidth + c];
}

// dh_v += Σ_{u ∈ N(v)} dx_{edge {u, v}} (neighbour order: a fixed sum per vertex).
__global__ __launch_bounds__(256) void ln_edge_gather_kernel(int n, const uint32_t* __restrict__ ptr,
                                                             const uint32_t* __restrict__ emap,
                                                             const float* __restrict__ dx, float* __restrict__ dh) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)n * kLnWidth) return;
  const size_t v = t / kLnWidth, c = t % kLnWidth;
  float acc = dh[t];
  for (uint32_t k = ptr[v]; k < ptr[v + 1]; ++k) acc += dx[(size_t)emap[k] * kLnWidth + c];
  dh[t] = acc;
}

// Adam (β₁ 0.9, β₂ 0.999, ε 1e-8) on g / batch, then g = 0.
__global__ __launch_bounds__(256) void ln_adam_kernel(size_t n, float* __restrict__ p, float* __restrict__ g,
                                                      float* __restrict__ m, float* __restrict__ v, float lr,
                                                      float inv_batch, float c1, float c2) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t] * inv_batch;
  const float mt = 0.9f * m[t] + 0.1f * gt, vt = 0.999f * v[t] + 0.001f * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + 1e-8f);
  g[t] = 0.f;
}

// z_v += s (B_v r_v + Σ_u E_uv r_u) (FP64 sums of the FP32 blocks).
__global__ __launch_bounds__(256) void ln_apply_kernel(int n, const uint32_t* __restrict__ ptr,
                                                       const uint32_t* __restrict__ nb, const uint32_t* __restrict__ emap,
                                                       const float* __restrict__ o, const float* __restrict__ eo,
                                                       double s, const double* __restrict__ r, double* __restrict__ z) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  const float* ov = o + (size_t)v * kLnOut;
  double q[3];
  for (int a = 0; a < 3; ++a) {
    double acc = 0.0;
    for (int b = 0; b < 3; ++b) acc += (double)ov[3 * a + b] * r[3 * (size_t)v + b];
    q[a] = acc;
  }
  for (uint32_t k = ptr[v]; k < ptr[v + 1]; ++k) {
    const uint32_t u = nb[k];
    const float* E = eo + (size_t)emap[k] * kLnOut;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) q[a] += (double)E[3 * a + b] * r[3 * (size_t)u + b];
  }
  for (int a = 0; a < 3; ++a) z[3 * (size_t)v + a] += s * q[a];
}

template <int I, int J>
void rows(ddgrs_ctx* ctx, int n, const float* X1, int ld1, int I1, const float* X2, int ld2, const float* M, int msi,
          int msj, const float* bias, int relu, float* pre, float* Y, int ldy) {
  if constexpr (J == 64)
    launch(ctx, "ln_rows_mfma", ln_rows_mfma_kernel<I>, dim3(blocks_for(n, 64)), dim3(256), 0, n, X1, ld1, I1, X2, ld2, M,
           msi, msj, bias, relu, pre, Y, ldy);
  else
    launch(ctx, "ln_rows_valu", ln_rows_kernel<I, J>, dim3(blocks_for(n, 64)), dim3(256), 0, n, X1, ld1, I1, X2, ld2, M, msi,
           msj, bias, relu, pre, Y, ldy);
}

constexpr int kGradBlocks = 512;

template <int I, int J>
void wgrad(ddgrs_ctx* ctx, int n, const float* A1, int lda1, int I1, const float* A2, int lda2, const float* B, int ldb,
           float* part, float* gw, float* gb) {
  const int per = (n + kGradBlocks - 1) / kGradBlocks;
  if constexpr (J == 64)
    launch(ctx, "ln_wgrad_mfma", ln_wgrad_mfma_kernel<I>, dim3(kGradBlocks), dim3(256), 0, n, per, A1, lda1, I1, A2, lda2, B,
           ldb, part);
  else
    launch(ctx, "ln_wgrad_valu", ln_wgrad_kernel<I, J>, dim3(kGradBlocks), dim3(256), 0, n, per, A1, lda1, I1, A2, lda2, B,
           ldb, part);
  launch(ctx, "ln_wgrad_sum", ln_wgrad_sum_kernel, dim3(blocks_for((I + 1) * J, 64)), dim3(256), 0, kGradBlocks, I, J,
         (const float*)part, gw, gb);
}

}  // namespace

void LearnedModel::init(ddgrs_ctx* ctx, const std::vector<uint32_t>& f, int n, uint64_t seed) {
  faces = f;
  nv = n;
  // neighbours: the mesh's edges, both directions, ascending
  std::vector<std::vector<uint32_t>> adj(nv);
  for (size_t t = 0; t + 2 < faces.size(); t += 3)
    for (int c = 0; c < 3; ++c) {
      const uint32_t a = faces[t + c], b = faces[t + (c + 1) % 3];
      adj[a].push_back(b);
      adj[b].push_back(a);
    }
  nbr_ptr_h.assign(1, 0);
  nbr_h.clear();
  for (int v = 0; v < nv; ++v) {
    std::sort(adj[v].begin(), adj[v].end());
    adj[v].erase(std::unique(adj[v].begin(), adj[v].end()), adj[v].end());
    nbr_h.insert(nbr_h.end(), adj[v].begin(), adj[v].end());
    nbr_ptr_h.push_back((uint32_t)nbr_h.size());
  }
  upload(ctx, nbr_ptr, nbr_ptr_h);
  upload(ctx, nbr, nbr_h);
  // undirected edges {u < v}: vertex u's larger neighbours in order, u ascending
  edge_h.clear();
  std::vector<uint32_t> first(nv + 1, 0);
  for (int u = 0; u < nv; ++u) {
    first[u] = (uint32_t)(edge_h.size() / 2);
    for (uint32_t k = nbr_ptr_h[u]; k < nbr_ptr_h[u + 1]; ++k)
      if (nbr_h[k] > (uint32_t)u) {
        edge_h.push_back((uint32_t)u);
        edge_h.push_back(nbr_h[k]);
      }
  }
  ne = (int)(edge_h.size() / 2);
  emap_h.assign(nbr_h.size(), 0);
  for (int v = 0; v < nv; ++v)
    for (uint32_t k = nbr_ptr_h[v]; k < nbr_ptr_h[v + 1]; ++k) {
      const uint32_t u = nbr_h[k], a = std::min<uint32_t>(u, v), b = std::max<uint32_t>(u, v);
      // the edge {a, b}: a's larger neighbours are in ascending order from first[a]
      uint32_t id = first[a];
      for (uint32_t q = nbr_ptr_h[a]; q < nbr_ptr_h[a + 1] && nbr_h[q] != b; ++q)
        if (nbr_h[q] > a) ++id;
      emap_h[k] = id;
    }
  upload(ctx, edge, edge_h);
  upload(ctx, emap, emap_h);
  // parameters: hidden layers uniform ±√(6 / (fan_in + fan_out)) from a fixed 64-bit LCG, biases 0;
  // the output layer 0 (Q = 0 before training)
  const LearnedLayout L = learned_layout();
  std::vector<float> p(L.total, 0.f);
  uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto uni = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  auto fill = [&](size_t off, size_t count, int fan_in, int fan_out) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    for (size_t i = 0; i < count; ++i) p[off + i] = (float)(a * uni());
  };
  fill(L.w0, (size_t)kLnFeat * kLnWidth, kLnFeat, kLnWidth);
  for (int l = 0; l < kLnLayers; ++l) fill(L.w[l], (size_t)2 * kLnWidth * kLnWidth, 2 * kLnWidth, kLnWidth);
  fill(L.we1, (size_t)kLnWidth * kLnWidth, kLnWidth, kLnWidth);
  upload(ctx, params, p);
  std::vector<float> z(L.total, 0.f);
  upload(ctx, grad, z);
  upload(ctx, adam_m, z);
  upload(ctx, adam_v, z);
  gpart.reserve((size_t)kGradBlocks * (2 * kLnWidth + 1) * kLnWidth);
  feat.reserve((size_t)nv * kLnFeat);
  act.reserve((size_t)(kLnLayers + 1) * nv * kLnWidth);
  pre.reserve((size_t)(kLnLayers + 1) * nv * kLnWidth);
  mean.reserve((size_t)kLnLayers * nv * kLnWidth);
  out.reserve((size_t)nv * kLnOut);
  dout.reserve((size_t)nv * kLnOut);
  const size_t EW = (size_t)std::max(ne, 1) * kLnWidth;
  e_in.reserve(EW);
  e_pre.reserve(EW);
  e_act.reserve(EW);
  e_d.reserve(EW);
  e_out.reserve((size_t)std::max(ne, 1) * kLnOut);
  e_dout.reserve((size_t)std::max(ne, 1) * kLnOut);
  dh.reserve((size_t)nv * kLnWidth);
  dtmp.reserve((size_t)2 * nv * kLnWidth);
  qd.reserve(3 * (size_t)nv);
  lsum.reserve(blocks_for(nv, 256));
  adam_t = 0;
  blocks_ready = false;
}

void LearnedModel::features(ddgrs_ctx* ctx, const double* pos, float* dst) {
  const size_t n = (size_t)nv;
  double c[3] = {0, 0, 0};
  for (size_t v = 0; v < n; ++v)
    for (int k = 0; k < 3; ++k) c[k] += pos[3 * v + k];
  for (int k = 0; k < 3; ++k) c[k] /= (double)n;
  double rr = 0.0;
  for (size_t v = 0; v < n; ++v)
    for (int k = 0; k < 3; ++k) rr += (pos[3 * v + k] - c[k]) * (pos[3 * v + k] - c[k]);
  const double R = std::sqrt(rr / (double)n) + 1e-300;
  std::vector<double> vn(3 * n, 0.0), area(n, 0.0);
  for (size_t t = 0; t + 2 < faces.size(); t += 3) {
    const double* p0 = pos + 3 * (size_t)faces[t];
    const double* p1 = pos + 3 * (size_t)faces[t + 1];
    const double* p2 = pos + 3 * (size_t)faces[t + 2];
    const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double nf[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double a2 = std::sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    for (int c2 = 0; c2 < 3; ++c2) {
      const size_t v = faces[t + c2];
      for (int k = 0; k < 3; ++k) vn[3 * v + k] += nf[k];
      area[v] += a2 / 6.0;
    }
  }
  double ma = 0.0, el = 0.0;
  size_t ne = 0;
  for (size_t v = 0; v < n; ++v) {
    ma += area[v];
    for (uint32_t q = nbr_ptr_h[v]; q < nbr_ptr_h[v + 1]; ++q) {
      const size_t u = nbr_h[q];
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) d2 += (pos[3 * u + k] - pos[3 * v + k]) * (pos[3 * u + k] - pos[3 * v + k]);
      el += d2;
      ++ne;
    }
  }
  ma /= (double)n;
  el = ne ? el / (double)ne : 1.0;
  float* f = (float*)ctx->arena_take(n * kLnFeat * sizeof(float));
  for (size_t v = 0; v < n; ++v) {
    double nn = std::sqrt(vn[3 * v] * vn[3 * v] + vn[3 * v + 1] * vn[3 * v + 1] + vn[3 * v + 2] * vn[3 * v + 2]);
    nn = nn > 0.0 ? nn : 1.0;
    double lap[3] = {0, 0, 0};
    const uint32_t q0 = nbr_ptr_h[v], q1 = nbr_ptr_h[v + 1];
    for (uint32_t q = q0; q < q1; ++q)
      for (int k = 0; k < 3; ++k) lap[k] += pos[3 * (size_t)nbr_h[q] + k];
    double h = 0.0;
    for (int k = 0; k < 3; ++k) {
      lap[k] = q1 > q0 ? lap[k] / (double)(q1 - q0) - pos[3 * v + k] : 0.0;
      h += lap[k] * vn[3 * v + k] / nn;
    }
    float* o = f + v * kLnFeat;
    for (int k = 0; k < 3; ++k) {
      o[k] = (float)((pos[3 * v + k] - c[k]) / R);
      o[3 + k] = (float)(vn[3 * v + k] / nn);
    }
    o[6] = (float)(area[v] / ma);
    o[7] = (float)(h / el * R);
  }
  HIPCHECK(hipMemcpyAsync(dst, f, n * kLnFeat * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
}

void LearnedModel::forward(ddgrs_ctx* ctx, const float* f) {
  const LearnedLayout L = learned_layout();
  const size_t VW = (size_t)nv * kLnWidth;
  const float* P = params.ptr;
  rows<kLnFeat, kLnWidth>(ctx, nv, f, kLnFeat, kLnFeat, nullptr, 0, P + L.w0, kLnWidth, 1, P + L.b0, 1, pre.ptr,
                          act.ptr, kLnWidth);
  for (int l = 0; l < kLnLayers; ++l) {
    const float* h = act.ptr + l * VW;
    float* m = mean.ptr + l * VW;
    launch(ctx, "ln_mean", ln_mean_kernel, dim3(blocks_for((int64_t)VW, 256)), dim3(256), 0, nv,
           (const uint32_t*)nbr_ptr.ptr, (const uint32_t*)nbr.ptr, h, m);
    rows<2 * kLnWidth, kLnWidth>(ctx, nv, h, kLnWidth, kLnWidth, m, kLnWidth, P + L.w[l], kLnWidth, 1, P + L.b[l], 1,
                                 pre.ptr + (l + 1) * VW, act.ptr + (l + 1) * VW, kLnWidth);
  }
  const float* h4 = act.ptr + kLnLayers * VW;
  rows<kLnWidth, kLnOut>(ctx, nv, h4, kLnWidth, kLnWidth, nullptr, 0, P + L.wv, kLnOut, 1, P + L.bv, 0, nullptr,
                         out.ptr, kLnOut);
  if (ne > 0) {
    const size_t EW = (size_t)ne * kLnWidth;
    launch(ctx, "ln_edge_in", ln_edge_in_kernel, dim3(blocks_for((int64_t)EW, 256)), dim3(256), 0, ne,
           (const uint32_t*)edge.ptr, h4, e_in.ptr);
    rows<kLnWidth, kLnWidth>(ctx, ne, e_in.ptr, kLnWidth, kLnWidth, nullptr, 0, P + L.we1, kLnWidth, 1, P + L.be1, 1,
                             e_pre.ptr, e_act.ptr, kLnWidth);
    rows<kLnWidth, kLnOut>(ctx, ne, e_act.ptr, kLnWidth, kLnWidth, nullptr, 0, P + L.we2, kLnOut, 1, P + L.be2, 0,
                           nullptr, e_out.ptr, kLnOut);
  }
  blocks_ready = true;
}

void LearnedModel::add_pair(ddgrs_ctx* ctx, const double* pos, const double* r, const double* pr, const double* y) {
  if (npairs >= cap_pairs) throw Error(DDGRS_INVALID_OPTION, "learned_add_pair", "the model holds no more pairs");
  const size_t n3 = 3 * (size_t)nv;
  features(ctx, pos, pair_feat.ptr + (size_t)npairs * nv * kLnFeat);
  float* h = (float*)ctx->arena_take(2 * n3 * sizeof(float));
  // the scalar c = ⟨y, P⁻¹ r⟩ / ‖P⁻¹ r‖² leaves GMRES's Krylov spaces unchanged: Q learns what
  // remains of y after it, from c P⁻¹ r
  double yp = 0.0, pp = 0.0;
  for (size_t i = 0; i < n3; ++i) {
    yp += y[i] * pr[i];
    pp += pr[i] * pr[i];
  }
  const double c = pp > 0.0 ? yp / pp : 1.0;
  double e2 = 0.0, r2 = 0.0;
  for (size_t i = 0; i < n3; ++i) {
    const double x = c * pr[i], e = y[i] - x;
    h[i] = (float)x;
    h[n3 + i] = (float)e;
    e2 += e * e;
    r2 += x * x;
  }
  HIPCHECK(hipMemcpyAsync(pair_r.ptr + npairs * n3, h, n3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIPCHECK(hipMemcpyAsync(pair_e.ptr + npairs * n3, h + n3, n3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  pair_e2.push_back(e2 > 0.0 ? e2 : 1.0);
  pair_ratio.push_back(r2 > 0.0 ? std::sqrt(e2 / r2) : 0.0);
  ++npairs;
}

double LearnedModel::loss_and_grad(ddgrs_ctx* ctx, int p, bool add_grad, bool read) {
  const LearnedLayout L = learned_layout();
  const size_t VW = (size_t)nv * kLnWidth, n3 = 3 * (size_t)nv;
  const float* f = pair_feat.ptr + (size_t)p * nv * kLnFeat;
  const float* r = pair_r.ptr + (size_t)p * n3;
  const float* e = pair_e.ptr + (size_t)p * n3;
  forward(ctx, f);
  const unsigned nbv = blocks_for(nv, 256);
  launch(ctx, "ln_loss", ln_loss_kernel, dim3(nbv), dim3(256), 0, nv, (const uint32_t*)nbr_ptr.ptr,
         (const uint32_t*)nbr.ptr, (const uint32_t*)emap.ptr, (const float*)out.ptr, (const float*)e_out.ptr, r, e,
         scale, (float)(1.0 / pair_e2[p]), qd.ptr, lsum.ptr);
  if (add_grad) {
    float* G = grad.ptr;
    const float* P = params.ptr;
    launch(ctx, "ln_dout", ln_dout_kernel, dim3(nbv), dim3(256), 0, nv, r, (const float*)qd.ptr, scale, dout.ptr);
    const float* h4 = act.ptr + kLnLayers * VW;
    wgrad<kLnWidth, kLnOut>(ctx, nv, h4, kLnWidth, kLnWidth, nullptr, 0, dout.ptr, kLnOut, gpart.ptr, G + L.wv, G + L.bv);
    // dh = dB W_vᵀ, then the edge head's share: dE → dt = dE W_e2ᵀ → ⊙ relu′ → dx = · W_e1ᵀ,
    // added to both ends of every edge
    rows<kLnOut, kLnWidth>(ctx, nv, dout.ptr, kLnOut, kLnOut, nullptr, 0, P + L.wv, 1, kLnOut, nullptr, 0, nullptr,
                           dh.ptr, kLnWidth);
    if (ne > 0) {
      const size_t EW = (size_t)ne * kLnWidth;
      launch(ctx, "ln_edge_dout", ln_edge_dout_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, ne,
             (const uint32_t*)edge.ptr, r, (const float*)qd.ptr, scale, e_dout.ptr);
      wgrad<kLnWidth, kLnOut>(ctx, ne, e_act.ptr, kLnWidth, kLnWidth, nullptr, 0, e_dout.ptr, kLnOut, gpart.ptr,
                              G + L.we2, G + L.be2);
      rows<kLnOut, kLnWidth>(ctx, ne, e_dout.ptr, kLnOut, kLnOut, nullptr, 0, P + L.we2, 1, kLnOut, nullptr, 0, nullptr,
                             e_d.ptr, kLnWidth);
      launch(ctx, "ln_relu_bwd", ln_relu_bwd_kernel, dim3(blocks_for((int64_t)EW, 256)), dim3(256), 0, (int)EW,
             (const float*)e_pre.ptr, e_d.ptr);
      wgrad<kLnWidth, kLnWidth>(ctx, ne, e_in.ptr, kLnWidth, kLnWidth, nullptr, 0, e_d.ptr, kLnWidth, gpart.ptr,
                                G + L.we1, G + L.be1);
      // dx = dpre W_e1ᵀ into e_in (its forward value is no longer needed)
      rows<kLnWidth, kLnWidth>(ctx, ne, e_d.ptr, kLnWidth, kLnWidth, nullptr, 0, P + L.we1, 1, kLnWidth, nullptr, 0,
                               nullptr, e_in.ptr, kLnWidth);
      launch(ctx, "ln_edge_gather", ln_edge_gather_kernel, dim3(blocks_for((int64_t)VW, 256)), dim3(256), 0, nv,
             (const uint32_t*)nbr_ptr.ptr, (const uint32_t*)emap.ptr, (const float*)e_in.ptr, dh.ptr);
    }
    for (int l = kLnLayers; l >= 1; --l) {
      float* dpre = dh.ptr;  // dh ⊙ (pre_l > 0), in place
      launch(ctx, "ln_relu_bwd", ln_relu_bwd_kernel, dim3(blocks_for((int64_t)VW, 256)), dim3(256), 0, (int)VW,
             (const float*)(pre.ptr + l * VW), dpre);
      const float* hp = act.ptr + (l - 1) * VW;
      const float* mp = mean.ptr + (l - 1) * VW;
      wgrad<2 * kLnWidth, kLnWidth>(ctx, nv, hp, kLnWidth, kLnWidth, mp, kLnWidth, dpre, kLnWidth, gpart.ptr,
                                    G + L.w[l - 1], G + L.b[l - 1]);
      // dh_{l−1} = dpre W_sᵀ + meanᵀ(dpre W_nᵀ)
      float* dnext = dtmp.ptr;
      float* dm = dtmp.ptr + VW;
      rows<kLnWidth, kLnWidth>(ctx, nv, dpre, kLnWidth, kLnWidth, nullptr, 0, P + L.w[l - 1], 1, kLnWidth, nullptr, 0,
                               nullptr, dnext, kLnWidth);
      rows<kLnWidth, kLnWidth>(ctx, nv, dpre, kLnWidth, kLnWidth, nullptr, 0, P + L.w[l - 1] + kLnWidth * kLnWidth, 1,
                               kLnWidth, nullptr, 0, nullptr, dm, kLnWidth);
      launch(ctx, "ln_mean_t", ln_mean_t_kernel, dim3(blocks_for((int64_t)VW, 256)), dim3(256), 0, nv,
             (const uint32_t*)nbr_ptr.ptr, (const uint32_t*)nbr.ptr, (const float*)dm, dnext);
      HIPCHECK(hipMemcpyAsync(dh.ptr, dnext, VW * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    launch(ctx, "ln_relu_bwd", ln_relu_bwd_kernel, dim3(blocks_for((int64_t)VW, 256)), dim3(256), 0, (int)VW,
           (const float*)pre.ptr, dh.ptr);
    wgrad<kLnFeat, kLnWidth>(ctx, nv, f, kLnFeat, kLnFeat, nullptr, 0, dh.ptr, kLnWidth, gpart.ptr, G + L.w0, G + L.b0);
  }
  if (!read) return 0.0;
  std::vector<double> ls(nbv);
  download(ctx, ls.data(), (const double*)lsum.ptr, nbv);
  ctx->sync();
  double s = 0.0;
  for (double x : ls) s += x;
  return s / pair_e2[p];
}

void LearnedModel::train(ddgrs_ctx* ctx, int iters, int batch, double lr, double* before, double* after) {
  if (npairs == 0) throw Error(DDGRS_INVALID_OPTION, "learned_train", "no training pairs");
  if (scale == 0.f) {
    std::vector<double> q = pair_ratio;
    std::nth_element(q.begin(), q.begin() + q.size() / 2, q.end());
    scale = (float)(q[q.size() / 2] > 0.0 ? q[q.size() / 2] : 1.0);
  }
  const LearnedLayout L = learned_layout();
  auto mean_loss = [&]() {
    double s = 0.0;
    for (int p = 0; p < npairs; ++p) s += loss_and_grad(ctx, p, false);
    return s / npairs;
  };
  // The parameters with the lowest mean loss over all pairs, checked once per epoch and at the end,
  // are the ones kept: training never ends above its starting loss (1 at Q = 0).
  double best = mean_loss();
  if (before) *before = best;
  best_params.reserve(L.total);
  HIPCHECK(hipMemcpyAsync(best_params.ptr, params.ptr, L.total * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  const int epoch = std::max(1, npairs / std::max(1, batch));
  launch(ctx, "ln_fill", ln_fill_kernel, dim3(blocks_for((int64_t)L.total, 256)), dim3(256), 0, L.total, grad.ptr);
  int next = 0;
  for (int it = 0; it < iters; ++it) {
    for (int b = 0; b < batch; ++b) {
      (void)loss_and_grad(ctx, next, true, false);
      next = (next + 1) % npairs;
    }
    ++adam_t;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)adam_t)), c2 = (float)(1.0 - std::pow(0.999, (double)adam_t));
    launch(ctx, "ln_adam", ln_adam_kernel, dim3(blocks_for((int64_t)L.total, 256)), dim3(256), 0, L.total, params.ptr,
           grad.ptr, adam_m.ptr, adam_v.ptr, (float)lr, 1.f / (float)batch, c1, c2);
    if ((it + 1) % epoch == 0 || it + 1 == iters) {
      const double l = mean_loss();
      if (l < best) {
        best = l;
        HIPCHECK(hipMemcpyAsync(best_params.ptr, params.ptr, L.total * sizeof(float), hipMemcpyDeviceToDevice,
                                ctx->stream));
      }
    }
  }
  HIPCHECK(hipMemcpyAsync(params.ptr, best_params.ptr, L.total * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  if (after) *after = best;
  blocks_ready = false;
}

void LearnedModel::apply(ddgrs_ctx* ctx, const double* r, double* z) {
  launch(ctx, "ln_apply", ln_apply_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const uint32_t*)nbr_ptr.ptr,
         (const uint32_t*)nbr.ptr, (const uint32_t*)emap.ptr, (const float*)out.ptr, (const float*)e_out.ptr,
         (double)scale, r, z);
}

}  // namespace ddgrs

// ---- C ABI -------------------------------------------------------------------------------------

struct ddgrs_learned {
  ddgrs_surface* s = nullptr;
  ddgrs::LearnedModel m;
};

using namespace ddgrs;

LearnedModel* ddgrs::learned_model(ddgrs_learned* l) { return &l->m; }

namespace {
void check_positions(ddgrs_learned* l, const double* pos) {
  need(pos, "positions");
  const size_t n = 3 * (size_t)l->s->nv;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(pos[i])) {
      Error e(DDGRS_NON_FINITE_POSITION, "learned", "a position is not finite");
      e.a = (int64_t)(i / 3);
      throw e;
    }
}
}  // namespace

extern "C" {

int ddgrs_learned_create(ddgrs_surface* s, uint64_t seed, int64_t max_pairs, ddgrs_learned** out, ddgrs_error* err) {
  return guarded(err, "learned_create", [&] {
    need(s, "surface");
    need(out, "out");
    *out = nullptr;
    if (max_pairs < 0) invalid("max_pairs", std::to_string(max_pairs), "non-negative");
    HIPCHECK(hipSetDevice(s->ctx->device));
    auto* l = new ddgrs_learned;
    l->s = s;
    surface_retain(s);
    try {
      LearnedModel& m = l->m;
      m.init(s->ctx, s->faces_h, (int)s->nv, seed);
      const size_t nv = (size_t)s->nv, cap = (size_t)max_pairs;
      m.cap_pairs = (int)cap;
      m.pair_feat.reserve(std::max<size_t>(1, cap) * nv * kLnFeat);
      m.pair_r.reserve(std::max<size_t>(1, cap) * 3 * nv);
      m.pair_e.reserve(std::max<size_t>(1, cap) * 3 * nv);
      m.feat.reserve(nv * kLnFeat);
      s->ctx->sync();
    } catch (...) {
      destroy_handle(l);
      throw;
    }
    *out = l;
  });
}

void ddgrs_learned_destroy(ddgrs_learned* l) { destroy_handle(l); }

int ddgrs_learned_info(ddgrs_learned* l, int64_t* out4, double* scale, ddgrs_error* err) {
  return guarded(err, "learned_info", [&] {
    need(l, "learned");
    if (out4) {
      out4[0] = (int64_t)learned_layout().total;
      out4[1] = l->m.npairs;
      out4[2] = l->m.adam_t;
      out4[3] = kLnFeat;
    }
    if (scale) *scale = l->m.scale;
  });
}

int ddgrs_learned_params(ddgrs_learned* l, float* get, const float* set, ddgrs_error* err) {
  return guarded(err, "learned_params", [&] {
    need(l, "learned");
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t n = learned_layout().total;
    if (get) {
      download(ctx, get, (const float*)l->m.params.ptr, n);
      ctx->sync();
    }
    if (set) {
      for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(set[i])) invalid("params", "a non-finite value", "finite parameters");
      upload(ctx, l->m.params, set, n);
      l->m.blocks_ready = false;
      ctx->sync();
    }
  });
}

int ddgrs_learned_set_scale(ddgrs_learned* l, double scale, ddgrs_error* err) {
  return guarded(err, "learned_set_scale", [&] {
    need(l, "learned");
    if (!(std::isfinite(scale) && scale > 0.0)) invalid("scale", std::to_string(scale), "a finite number > 0");
    l->m.scale = (float)scale;
  });
}

int ddgrs_learned_set_krylov_pairs(ddgrs_learned* l, int32_t krylov_pairs, ddgrs_error* err) {
  return guarded(err, "learned_set_krylov_pairs", [&] {
    need(l, "learned");
    if (krylov_pairs < 0 || krylov_pairs > 8) invalid("krylov_pairs", std::to_string(krylov_pairs), "0 to 8");
    l->m.krylov_pairs = krylov_pairs;
  });
}

int ddgrs_learned_add_pair(ddgrs_learned* l, const double* positions, const double* r, const double* pr,
                           const double* y, ddgrs_error* err) {
  return guarded(err, "learned_add_pair", [&] {
    need(l, "learned");
    need(r, "r");
    need(pr, "pr");
    need(y, "y");
    check_positions(l, positions);
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    l->m.add_pair(ctx, positions, r, pr, y);
    ctx->sync();
  });
}

int ddgrs_learned_features(ddgrs_learned* l, const double* positions, float* feat, ddgrs_error* err) {
  return guarded(err, "learned_features", [&] {
    need(l, "learned");
    need(feat, "feat");
    check_positions(l, positions);
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    l->m.features(ctx, positions, l->m.feat.ptr);
    download(ctx, feat, (const float*)l->m.feat.ptr, (size_t)l->s->nv * kLnFeat);
    ctx->sync();
  });
}

int ddgrs_learned_blocks(ddgrs_learned* l, const double* positions, float* blocks, ddgrs_error* err) {
  return guarded(err, "learned_blocks", [&] {
    need(l, "learned");
    check_positions(l, positions);
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    l->m.features(ctx, positions, l->m.feat.ptr);
    l->m.forward(ctx, l->m.feat.ptr);
    if (blocks) {
      download(ctx, blocks, (const float*)l->m.out.ptr, (size_t)l->s->nv * kLnOut);
      download(ctx, blocks + (size_t)l->s->nv * kLnOut, (const float*)l->m.e_out.ptr, (size_t)l->m.ne * kLnOut);
    }
    ctx->sync();
  });
}

int ddgrs_learned_loss_grad(ddgrs_learned* l, int64_t pair, double* loss, float* grad, ddgrs_error* err) {
  return guarded(err, "learned_loss_grad", [&] {
    need(l, "learned");
    if (pair < 0 || pair >= l->m.npairs) invalid("pair", std::to_string(pair), "an index of an added pair");
    if (!(l->m.scale > 0.f)) invalid("scale", "unset", "a scale (learned_set_scale or a training run)");
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t n = learned_layout().total;
    if (grad) launch(ctx, "ln_fill", ln_fill_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, n, l->m.grad.ptr);
    const double v = l->m.loss_and_grad(ctx, (int)pair, grad != nullptr);
    if (loss) *loss = v;
    if (grad) {
      download(ctx, grad, (const float*)l->m.grad.ptr, n);
      launch(ctx, "ln_fill", ln_fill_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, n, l->m.grad.ptr);
    }
    ctx->sync();
    l->m.blocks_ready = false;
  });
}

int ddgrs_learned_train(ddgrs_learned* l, int64_t iterations, int64_t batch, double lr, double* loss2,
                        ddgrs_error* err) {
  return guarded(err, "learned_train", [&] {
    need(l, "learned");
    if (iterations < 0) invalid("iterations", std::to_string(iterations), "non-negative");
    if (batch < 1) invalid("batch", std::to_string(batch), "at least 1");
    if (!(std::isfinite(lr) && lr > 0.0)) invalid("lr", std::to_string(lr), "a finite number > 0");
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    double b = 0.0, a = 0.0;
    l->m.train(ctx, (int)iterations, (int)batch, lr, &b, &a);
    ctx->sync();
    if (loss2) {
      loss2[0] = b;
      loss2[1] = a;
    }
  });
}

int ddgrs_learned_apply_batch(ddgrs_learned* l, const double* positions, int64_t nb, const double* rs, double* zs,
                              ddgrs_error* err) {
  return guarded(err, "learned_apply_batch", [&] {
    need(l, "learned");
    need(rs, "rs");
    need(zs, "zs");
    if (nb < 0) invalid("nb", std::to_string(nb), "non-negative");
    if (!(l->m.scale > 0.f)) invalid("scale", "unset", "a scale (learned_set_scale or a training run)");
    check_positions(l, positions);
    ddgrs_ctx* ctx = l->s->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    LearnedModel& m = l->m;
    m.features(ctx, positions, m.feat.ptr);
    m.forward(ctx, m.feat.ptr);
    const size_t n = 3 * (size_t)l->s->nv + 3;
    DevBuf<double> r(n), z(n);
    for (int64_t k = 0; k < nb; ++k) {
      HIPCHECK(hipMemcpyAsync(r.ptr, rs + k * n, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      HIPCHECK(hipMemsetAsync(z.ptr, 0, n * sizeof(double), ctx->stream));
      m.apply(ctx, r.ptr, z.ptr);
      download(ctx, zs + k * n, (const double*)z.ptr, n);
      ctx->sync();
    }
    m.blocks_ready = false;
  });
}

}  // extern "C"
