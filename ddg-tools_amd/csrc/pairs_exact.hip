// The exact all-pairs tangent-point energy and its per-face partials: FP32 pair kernel with
// compensated accumulation.
//
// Follows RS/src/energy/exact.rs:87 `face_partials`: the task of face S (its row) evaluates, for
// every T ≠ S, both directions of the pair with a shared r^{-β} (RS/src/kernel.rs:256
// `pair_terms`) and accumulates
//   e_S = Σ a_T K(S→T),  ∂/∂a_S = Σ a_T (K(S→T) + K(T→S)),
//   ∂/∂X_S = a_S Σ a_T (∂K(S→T)/∂v − ∂K(T→S)/∂v'),  ∂/∂N_S = a_S Σ a_T ∂K(S→T)/∂N_S.
//
// Layout and schedule (gfx950): one lane per row S, 256 rows per workgroup; the columns T are split
// into P chunks (grid.y) so that a launch has ≥ 2048 workgroups.  The T data of an iteration is
// wave-uniform, so it is read with scalar loads into SGPRs (no LDS, no barriers) and every VALU
// operation takes it as an SGPR operand.  Sums: 64 terms in plain FP32, then a Neumaier-compensated
// FP32 sum over the tiles of a chunk; the P chunk sums of a row are combined as a double-single
// (TwoSum) and only then widened to FP64.  The partition (tile 64, P) depends on nf only, so results
// are bitwise reproducible run to run.
#include "kernels.hpp"

namespace ddgrs {

constexpr int kPairTile = 64;
constexpr int kPairRows = 256;

// MODE 0: (α, β) = (6, 12) in the overflow-safe form K = (ν²/r²)³ (1/r²)³.
// MODE 1: integer α ≤ 64, β ≤ 128 by repeated products (RS/src/kernel.rs:149-166 with ipow).
// MODE 2: other exponents through exp2/log2.
template <int MODE>
struct PairPow {
  float alpha, beta;
  int aint, bint;
  // |ν|^α and sign(ν)|ν|^(α−1)
  __device__ __forceinline__ void nu(float v, float& pa, float& sp) const {
    float a = fabsf(v);
    if constexpr (MODE == 1) {
      float p = 1.f;
      for (int k = 0; k < aint - 1; ++k) p *= a;
      pa = p * a;
      sp = copysignf(p, v);
    } else {
      float l = __log2f(a);
      float p = exp2f((alpha - 1.f) * l);
      pa = p * a;
      sp = copysignf(p, v);
    }
  }
  // r^{-β} from r² and 1/r²
  __device__ __forceinline__ float rb(float r2, float inv) const {
    if constexpr (MODE == 1) {
      float p = 1.f;
      for (int k = 0; k < bint / 2; ++k) p *= inv;
      if (bint & 1) p *= __frsqrt_rn(r2);
      return p;
    } else {
      return exp2f(-0.5f * beta * __log2f(r2));
    }
  }
};

struct Acc8 {
  float e, da, dx0, dx1, dx2, dn0, dn1, dn2;
};

template <int MODE>
__device__ __forceinline__ void pair_both(const PairPow<MODE>& pw, float vx, float vy, float vz, float4 ns, float4 nt,
                                          float w, float r2, Acc8& a) {
  float inv = __builtin_amdgcn_rcpf(r2);
  float nu_s = ns.x * vx + ns.y * vy + ns.z * vz;
  float nu_t = -(nt.x * vx + nt.y * vy + nt.z * vz);
  float k_st, k_ts, c_s, c_t;
  if constexpr (MODE == 0) {
    float i3 = inv * inv * inv;
    float ts = nu_s * nu_s * inv, tt = nu_t * nu_t * inv;
    float ts2 = ts * ts, tt2 = tt * tt;
    k_st = ts2 * ts * i3;
    k_ts = tt2 * tt * i3;
    c_s = pw.alpha * (nu_s * inv) * ts2 * i3;
    c_t = pw.alpha * (nu_t * inv) * tt2 * i3;
  } else {
    float rbv = pw.rb(r2, inv);
    float pa_s, sp_s, pa_t, sp_t;
    pw.nu(nu_s, pa_s, sp_s);
    pw.nu(nu_t, pa_t, sp_t);
    k_st = pa_s * rbv;
    k_ts = pa_t * rbv;
    c_s = pw.alpha * sp_s * rbv;
    c_t = pw.alpha * sp_t * rbv;
  }
  float b = pw.beta * (k_st + k_ts) * inv;
  float wcs = w * c_s, wct = w * c_t, wb = w * b;
  a.e += w * k_st;
  a.da += w * (k_st + k_ts);
  a.dx0 += wcs * ns.x - wct * nt.x - wb * vx;
  a.dx1 += wcs * ns.y - wct * nt.y - wb * vy;
  a.dx2 += wcs * ns.z - wct * nt.z - wb * vz;
  a.dn0 += wcs * vx;
  a.dn1 += wcs * vy;
  a.dn2 += wcs * vz;
}

// Neumaier-compensated s += x (c collects the rounding errors).
__device__ __forceinline__ void comp_add(float& s, float& c, float x) {
  float t = s + x;
  float e = (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
  c += e;
  s = t;
}

template <int MODE>
__global__ __launch_bounds__(kPairRows) void exact_pairs_kernel(int nf, int chunk, const float4* __restrict__ xhi,
                                                                const float4* __restrict__ xlo,
                                                                const float4* __restrict__ nrm, PairPow<MODE> pw,
                                                                float2* __restrict__ partial) {
  const int s = blockIdx.x * kPairRows + threadIdx.x;
  const int sc = s < nf ? s : nf - 1;
  const float4 hs = xhi[sc], ls = xlo[sc], ns = nrm[sc];
  const int c0 = blockIdx.y * chunk;
  const int c1 = min(c0 + chunk, nf);
  float S[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float C[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int t0 = c0; t0 < c1; t0 += kPairTile) {
    const int t1 = min(t0 + kPairTile, c1);
    Acc8 a = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int t = t0; t < t1; ++t) {
      const float4 ht = xhi[t], lt = xlo[t], nt = nrm[t];
      const float vx = (hs.x - ht.x) + (ls.x - lt.x);
      const float vy = (hs.y - ht.y) + (ls.y - lt.y);
      const float vz = (hs.z - ht.z) + (ls.z - lt.z);
      const bool self = (t == sc);
      const float r2 = self ? 1.f : (vx * vx + vy * vy + vz * vz);
      const float w = self ? 0.f : ht.w;
      pair_both<MODE>(pw, vx, vy, vz, ns, nt, w, r2, a);
    }
    comp_add(S[0], C[0], a.e);
    comp_add(S[1], C[1], a.da);
    comp_add(S[2], C[2], a.dx0);
    comp_add(S[3], C[3], a.dx1);
    comp_add(S[4], C[4], a.dx2);
    comp_add(S[5], C[5], a.dn0);
    comp_add(S[6], C[6], a.dn1);
    comp_add(S[7], C[7], a.dn2);
  }
  if (s < nf) {
#pragma unroll
    for (int k = 0; k < 8; ++k) partial[((size_t)blockIdx.y * 8 + k) * nf + s] = make_float2(S[k], C[k]);
  }
}

// Per row: the P chunk sums as a double-single (TwoSum), widened to FP64, then the scaling by a_S:
// partials = (∂/∂a, a ∂/∂X, a ∂/∂N), row energy = a e.
__global__ __launch_bounds__(256) void exact_combine_kernel(int nf, int P, const float2* __restrict__ partial,
                                                            const double* __restrict__ area,
                                                            double* __restrict__ out, double* __restrict__ row_energy) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nf) return;
  double v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float hi = 0.f, lo = 0.f;
    for (int p = 0; p < P; ++p) {
      float2 q = partial[((size_t)p * 8 + k) * nf + s];
      float t = hi + q.x;
      float bp = t - hi;
      float err = (hi - (t - bp)) + (q.x - bp);
      hi = t;
      lo += err + q.y;
    }
    v[k] = (double)hi + (double)lo;
  }
  double a = area[s];
  row_energy[s] = a * v[0];
  out[7 * (size_t)s] = v[1];
  for (int k = 0; k < 3; ++k) {
    out[7 * (size_t)s + 1 + k] = a * v[2 + k];
    out[7 * (size_t)s + 4 + k] = a * v[5 + k];
  }
}

namespace {
// The column partition: P chunks of `chunk` columns (a multiple of the tile), P · ⌈nf/256⌉ ≥ 2048
// workgroups where nf allows; it depends on nf only, so results do not depend on the device.
void pair_partition(int nf, int* P_out, int* chunk_out) {
  const int rb = (nf + kPairRows - 1) / kPairRows;
  int P = (2048 + rb - 1) / rb;
  P = std::max(1, std::min(P, (nf + kPairRows - 1) / kPairRows));
  int chunk = (nf + P - 1) / P;
  chunk = (chunk + kPairTile - 1) / kPairTile * kPairTile;
  *P_out = (nf + chunk - 1) / chunk;
  *chunk_out = chunk;
}
}  // namespace

size_t exact_pairs_chunk_elems(int64_t nf) {
  if (nf <= 0 || nf > INT32_MAX / 8) return 0;
  int P, chunk;
  pair_partition((int)nf, &P, &chunk);
  return (size_t)P * 8 * (size_t)nf;
}

void exact_pairs(ddgrs_ctx* ctx, int64_t nf64, const FacePacks& packs, const double* area, const Exps& e,
                 double* partials, double* row_energy, float2* chunk_partials) {
  if (nf64 > INT32_MAX / 8) invalid("faces", std::to_string(nf64), "at most 2^28 faces for the all-pairs kernel");
  const int nf = (int)nf64;
  const int rb = (nf + kPairRows - 1) / kPairRows;
  int P, chunk;
  pair_partition(nf, &P, &chunk);
  float2* part = chunk_partials ? chunk_partials : scratch_of(ctx).f2((size_t)P * 8 * nf);
  dim3 grid(rb, P);
  if (e.alpha_int == 6 && e.beta_int == 12) {
    PairPow<0> pw{e.alpha, e.beta, 6, 12};
    launch(ctx, "exact_pairs", exact_pairs_kernel<0>, grid, dim3(kPairRows), 0, nf, chunk,
           (const float4*)packs.xhi.ptr, (const float4*)packs.xlo.ptr, (const float4*)packs.nrm.ptr, pw, part);
  } else if (e.alpha_int >= 0 && e.beta_int >= 0) {
    PairPow<1> pw{e.alpha, e.beta, e.alpha_int, e.beta_int};
    launch(ctx, "exact_pairs", exact_pairs_kernel<1>, grid, dim3(kPairRows), 0, nf, chunk,
           (const float4*)packs.xhi.ptr, (const float4*)packs.xlo.ptr, (const float4*)packs.nrm.ptr, pw, part);
  } else {
    PairPow<2> pw{e.alpha, e.beta, e.alpha_int, e.beta_int};
    launch(ctx, "exact_pairs", exact_pairs_kernel<2>, grid, dim3(kPairRows), 0, nf, chunk,
           (const float4*)packs.xhi.ptr, (const float4*)packs.xlo.ptr, (const float4*)packs.nrm.ptr, pw, part);
  }
  launch(ctx, "exact_combine", exact_combine_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, P,
         (const float2*)part, area, partials, row_energy);
}

}  // namespace ddgrs
