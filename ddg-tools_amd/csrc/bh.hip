// Barnes–Hut energy Ẽ and its exact gradient for frozen lists on the device.
//
// Follows RS/src/energy/barnes_hut.rs: leaf_with_gradient :627 (own, node and partner terms),
// face_partials :478 (node sums over the transposed lumped lists :506-520, downward pass
// :521-538, per-leaf final :540-567), leaf_energy :583 (value only, line-search trials), refit
// :204 (face data and node aggregates A_C, X_C; boxes, radii and lists frozen).
//
// Schedule (gfx950): one wave per leaf ℓ, lane = 8 i + j.  Lumped entries are spread over j
// (lane (i, j) evaluates face S_i against node C_{e}, e ≡ j mod 8; the node terms b, u are reduced
// over i with shuffles); an exact block (ℓ, C) puts the ≤ 8 × 8 pairs (S_i, T_j) on the 64 lanes;
// the partner terms are gathered on the target side: the wave of leaf C walks the transposed list
// (the leaves that open C) and evaluates K(S'_j → T_i) itself, so no partner array is stored.
// Own and partner sums are reduced over j by an xor butterfly (bitwise identical on all lanes,
// independent of scheduling).  Pair arithmetic is FP32 with the hi/lo barycenter difference; the
// node sums, the downward pass and the per-face epilogue are FP64.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "bh.hpp"

namespace ddgrs {

#pragma clang fp contract(fast)

template <int MODE>
struct KPow {
  float alpha, beta;
  int aint, bint;
};

// K, ∂K/∂v (vector), ∂K/∂N factor c_n (∂K/∂N = c_n v) for v and the source normal n
// (RS/src/kernel.rs:238 kernel_with_gradient_of_v).
template <int MODE>
__device__ __forceinline__ void kgrad(const KPow<MODE>& p, float vx, float vy, float vz, float nx, float ny, float nz,
                                      float r2, float& k, float& cn, float& bv) {
  float inv = __builtin_amdgcn_rcpf(r2);
  float nu = nx * vx + ny * vy + nz * vz;
  if constexpr (MODE == 0) {  // (6, 12)
    float i3 = inv * inv * inv;
    float t = nu * nu * inv;
    float t2 = t * t;
    k = t2 * t * i3;
    cn = p.alpha * (nu * inv) * t2 * i3;
  } else if constexpr (MODE == 1) {
    float a = fabsf(nu), q = 1.f;
    for (int i = 0; i < p.aint - 1; ++i) q *= a;
    float rb = 1.f;
    for (int i = 0; i < p.bint / 2; ++i) rb *= inv;
    if (p.bint & 1) rb *= __frsqrt_rn(r2);
    k = q * a * rb;
    cn = p.alpha * copysignf(q, nu) * rb;
  } else {
    float a = fabsf(nu);
    float q = exp2f((p.alpha - 1.f) * __log2f(a));
    float rb = exp2f(-0.5f * p.beta * __log2f(r2));
    k = q * a * rb;
    cn = p.alpha * copysignf(q, nu) * rb;
  }
  bv = p.beta * k * inv;  // ∂K/∂v = cn n − bv v
}

template <int MODE>
__device__ __forceinline__ float kval(const KPow<MODE>& p, float vx, float vy, float vz, float nx, float ny, float nz,
                                      float r2) {
  float inv = __builtin_amdgcn_rcpf(r2);
  float nu = nx * vx + ny * vy + nz * vz;
  if constexpr (MODE == 0) {
    float i3 = inv * inv * inv;
    float t = nu * nu * inv;
    return t * t * t * i3;
  } else if constexpr (MODE == 1) {
    float a = fabsf(nu), q = 1.f;
    for (int i = 0; i < p.aint; ++i) q *= a;
    float rb = 1.f;
    for (int i = 0; i < p.bint / 2; ++i) rb *= inv;
    if (p.bint & 1) rb *= __frsqrt_rn(r2);
    return q * rb;
  } else {
    float a = fabsf(nu);
    return exp2f(p.alpha * __log2f(a)) * exp2f(-0.5f * p.beta * __log2f(r2));
  }
}

__device__ __forceinline__ float xsum8(float x) {  // sum over the 8 lanes j of a row i
  x += __shfl_xor(x, 1);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 4);
  return x;
}
__device__ __forceinline__ float isum8(float x) {  // sum over the 8 rows i (lanes j, j+8, …)
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  return x;
}

struct BhArgs {
  int n_leaves, n_items;
  const uint32_t *item_leaf, *item_g0;  // per work item: leaf ordinal, first step
  const uint32_t *leaves, *nstart, *nend;
  const uint32_t *lumped_off, *lumped;
  const uint32_t *ex_off, *ex_pos;  // per leaf: BVH positions of the faces of its exact nodes (list order)
  const uint32_t *pt_off, *pt_pos;  // per leaf: BVH positions of the faces of the leaves that open it
  const float4 *xhi, *xlo, *nrm;    // faces, BVH order
  const float4 *chi, *clo;          // nodes: (X_C hi, A_C), (X_C − hi, 0)
};

// Entries [b, e) of a per-leaf list, 8 per step (lane slot j): the list's u32 values are loaded 64 at
// a time (one per lane) and handed to the slots with a lane shuffle.  U steps at a time: load(v)
// fetches a step's data for this lane's value v (every lane loads, past-the-end slots from a valid
// address), then step(data, v, valid, entry) consumes the U steps in order, so that each wave has
// U steps' loads in flight.  Every lane runs the same steps.
template <int U, class Load, class Step>
__device__ __forceinline__ void walk8(const uint32_t* __restrict__ list, uint32_t b, uint32_t e, int lane, int j,
                                      Load&& load, Step&& step) {
  using Data = decltype(load(0u));
  for (uint32_t b0 = b; b0 < e; b0 += 64) {
    const uint32_t mine = b0 + lane < e ? list[b0 + lane] : 0u;
    const uint32_t n = min(64u, e - b0);
    for (uint32_t s = 0; s < n; s += 8 * U) {
      Data d[U];
      uint32_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t src = min(s + 8 * u + j, 63u);
        v[u] = (uint32_t)__shfl((int)mine, (int)src);
        d[u] = load(v[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (s + 8 * u < n) step(d[u], v[u], s + 8 * u + j < n, b0 + s + 8 * u + j);
    }
  }
}

// The entries of list [b, e) that fall in the item's steps [g0, g0 + kBhItemSteps), the list's
// steps numbered from o (8 entries a step).
__device__ __forceinline__ void item_piece(uint32_t b, uint32_t e, uint32_t o, uint32_t g0, uint32_t& pb,
                                           uint32_t& pe) {
  const uint32_t st = (e - b + 7) / 8, g1 = g0 + kBhItemSteps;
  const uint32_t lo = g0 > o ? min(g0 - o, st) : 0u, hi = g1 > o ? min(g1 - o, st) : 0u;
  pb = b + 8 * lo;
  pe = max(pb, min(b + 8 * hi, e));
}

struct NodeData {
  float4 h, l;
};
struct FaceData {
  float4 h, l;
};
struct FaceDataN {
  float4 h, l, n;
};

template <int MODE>
__global__ __launch_bounds__(256) void bh_gradient_kernel(BhArgs A, KPow<MODE> pw, float4* __restrict__ node_bu,
                                                          float* __restrict__ part, double* __restrict__ item_energy) {
  const int lane = threadIdx.x & 63;
  const int it = (int)xcd_wave_id();
  if (it >= A.n_items) return;  // whole waves exit together
  const int k = (int)A.item_leaf[it];
  const uint32_t g0 = A.item_g0[it];
  const int i = lane >> 3, j = lane & 7;
  const uint32_t L = A.leaves[k];
  const uint32_t s0 = A.nstart[L];
  const int ns = (int)(A.nend[L] - s0);
  const bool vi = i < ns;
  const uint32_t ps = s0 + (vi ? i : 0);
  const float4 sh = A.xhi[ps], sl = A.xlo[ps], sn = A.nrm[ps];
  const float as = vi ? sh.w : 0.f;
  float own_a = 0.f, ox = 0.f, oy = 0.f, oz = 0.f, onx = 0.f, ony = 0.f, onz = 0.f, en = 0.f;
  // 1. lumped nodes: lane (i, j) evaluates face S_i against node C of slot j
  auto load_node = [&](uint32_t C) { return NodeData{A.chi[C], A.clo[C]}; };
  auto load_face = [&](uint32_t p) { return FaceData{A.xhi[p], A.xlo[p]}; };
  auto load_face_n = [&](uint32_t p) { return FaceDataN{A.xhi[p], A.xlo[p], A.nrm[p]}; };
  const uint32_t lb = A.lumped_off[k], le = A.lumped_off[k + 1], eb = A.ex_off[k], ee = A.ex_off[k + 1];
  const uint32_t sl_steps = (le - lb + 7) / 8, se_steps = (ee - eb + 7) / 8;
  uint32_t b1, e1, b2, e2, b3, e3;
  item_piece(lb, le, 0, g0, b1, e1);
  item_piece(eb, ee, sl_steps, g0, b2, e2);
  item_piece(A.pt_off[k], A.pt_off[k + 1], sl_steps + se_steps, g0, b3, e3);
  walk8<4>(A.lumped, b1, e1, lane, j, load_node, [&](const NodeData& nd, uint32_t, bool act, uint32_t e) {
    const float4 ch = nd.h, cl = nd.l;
    const float vx = (sh.x - ch.x) + (sl.x - cl.x);
    const float vy = (sh.y - ch.y) + (sl.y - cl.y);
    const float vz = (sh.z - ch.z) + (sl.z - cl.z);
    float kk, cn, bv;
    kgrad<MODE>(pw, vx, vy, vz, sn.x, sn.y, sn.z, vx * vx + vy * vy + vz * vz, kk, cn, bv);
    const float w = (act && vi) ? 1.f : 0.f;
    const float ac = ch.w;
    const float dvx = cn * sn.x - bv * vx, dvy = cn * sn.y - bv * vy, dvz = cn * sn.z - bv * vz;
    own_a += w * kk * ac;
    const float sa = w * as * ac;
    ox += sa * dvx;
    oy += sa * dvy;
    oz += sa * dvz;
    onx += sa * cn * vx;
    ony += sa * cn * vy;
    onz += sa * cn * vz;
    en += sa * kk;
    const float b = isum8(w * as * kk);
    const float ux = isum8(-w * as * dvx), uy = isum8(-w * as * dvy), uz = isum8(-w * as * dvz);
    if (i == 0 && act) node_bu[e] = make_float4(b, ux, uy, uz);
  });
  // 2. exact blocks, own side: pairs (S_i, T) over the faces T of the leaf's exact nodes, slot j
  float row = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
  walk8<4>(A.ex_pos, b2, e2, lane, j, load_face, [&](const FaceData& fd, uint32_t pt, bool valid, uint32_t) {
    const bool act = vi && valid && pt != ps;
    const float4 th = fd.h, tl = fd.l;
    const float vx = (sh.x - th.x) + (sl.x - tl.x);
    const float vy = (sh.y - th.y) + (sl.y - tl.y);
    const float vz = (sh.z - th.z) + (sl.z - tl.z);
    const float r2 = act ? (vx * vx + vy * vy + vz * vz) : 1.f;
    float kk, cn, bv;
    kgrad<MODE>(pw, vx, vy, vz, sn.x, sn.y, sn.z, r2, kk, cn, bv);
    const float at = act ? th.w : 0.f;
    row += at * kk;
    dx += at * (cn * sn.x - bv * vx);
    dy += at * (cn * sn.y - bv * vy);
    dz += at * (cn * sn.z - bv * vz);
    nx += at * cn * vx;
    ny += at * cn * vy;
    nz += at * cn * vz;
  });
  // 3. exact blocks, partner side: pairs (S', T_i) over the faces S' of the leaves that open ℓ
  float pa = 0.f, px = 0.f, py = 0.f, pz = 0.f;
  walk8<4>(A.pt_pos, b3, e3, lane, j, load_face_n, [&](const FaceDataN& fd, uint32_t pu, bool valid, uint32_t) {
    const bool act = vi && valid && pu != ps;
    const float4 uh = fd.h, ul = fd.l, un = fd.n;
    const float vx = (uh.x - sh.x) + (ul.x - sl.x);
    const float vy = (uh.y - sh.y) + (ul.y - sl.y);
    const float vz = (uh.z - sh.z) + (ul.z - sl.z);
    const float r2 = act ? (vx * vx + vy * vy + vz * vz) : 1.f;
    float kk, cn, bv;
    kgrad<MODE>(pw, vx, vy, vz, un.x, un.y, un.z, r2, kk, cn, bv);
    const float au = act ? uh.w : 0.f;
    pa += au * kk;
    const float w = -au * as;
    px += w * (cn * un.x - bv * vx);
    py += w * (cn * un.y - bv * vy);
    pz += w * (cn * un.z - bv * vz);
  });
  // reduce over j and finish the own terms of S_i
  own_a = xsum8(own_a);
  ox = xsum8(ox); oy = xsum8(oy); oz = xsum8(oz);
  onx = xsum8(onx); ony = xsum8(ony); onz = xsum8(onz);
  row = xsum8(row);
  dx = xsum8(dx); dy = xsum8(dy); dz = xsum8(dz);
  nx = xsum8(nx); ny = xsum8(ny); nz = xsum8(nz);
  pa = xsum8(pa);
  px = xsum8(px); py = xsum8(py); pz = xsum8(pz);
  if (j == 0 && vi) {
    float* o = part + 7 * ((size_t)it * 8 + i);
    o[0] = (own_a + row) + pa;
    o[1] = (ox + as * dx) + px;
    o[2] = (oy + as * dy) + py;
    o[3] = (oz + as * dz) + pz;
    o[4] = onx + as * nx;
    o[5] = ony + as * ny;
    o[6] = onz + as * nz;
  }
  // leaf energy: lumped terms (spread over all lanes) + a_S · row (lanes j == 0)
  float el = en + ((j == 0 && vi) ? as * row : 0.f);
  el = xsum8(el);
  el = isum8(el);
  if (lane == 0) item_energy[it] = (double)el;
}

template <int MODE>
__global__ __launch_bounds__(256) void bh_energy_kernel(BhArgs A, KPow<MODE> pw, double* __restrict__ item_energy) {
  const int lane = threadIdx.x & 63;
  const int it = (int)xcd_wave_id();
  if (it >= A.n_items) return;
  const int k = (int)A.item_leaf[it];
  const uint32_t g0 = A.item_g0[it];
  const int i = lane >> 3, j = lane & 7;
  const uint32_t L = A.leaves[k];
  const uint32_t s0 = A.nstart[L];
  const int ns = (int)(A.nend[L] - s0);
  const bool vi = i < ns;
  const uint32_t ps = s0 + (vi ? i : 0);
  const float4 sh = A.xhi[ps], sl = A.xlo[ps], sn = A.nrm[ps];
  const float as = vi ? sh.w : 0.f;
  float en = 0.f;
  auto load_node = [&](uint32_t C) { return NodeData{A.chi[C], A.clo[C]}; };
  auto load_face = [&](uint32_t p) { return FaceData{A.xhi[p], A.xlo[p]}; };
  const uint32_t lb = A.lumped_off[k], le = A.lumped_off[k + 1];
  uint32_t b1, e1, b2, e2;
  item_piece(lb, le, 0, g0, b1, e1);
  item_piece(A.ex_off[k], A.ex_off[k + 1], (le - lb + 7) / 8, g0, b2, e2);
  walk8<4>(A.lumped, b1, e1, lane, j, load_node, [&](const NodeData& nd, uint32_t, bool act, uint32_t) {
    const float4 ch = nd.h, cl = nd.l;
    const float vx = (sh.x - ch.x) + (sl.x - cl.x);
    const float vy = (sh.y - ch.y) + (sl.y - cl.y);
    const float vz = (sh.z - ch.z) + (sl.z - cl.z);
    const float kk = kval<MODE>(pw, vx, vy, vz, sn.x, sn.y, sn.z, vx * vx + vy * vy + vz * vz);
    en += (act && vi) ? as * ch.w * kk : 0.f;
  });
  walk8<4>(A.ex_pos, b2, e2, lane, j, load_face, [&](const FaceData& fd, uint32_t pt, bool valid, uint32_t) {
    const bool act = vi && valid && pt != ps;
    const float4 th = fd.h, tl = fd.l;
    const float vx = (sh.x - th.x) + (sl.x - tl.x);
    const float vy = (sh.y - th.y) + (sl.y - tl.y);
    const float vz = (sh.z - th.z) + (sl.z - tl.z);
    const float r2 = act ? (vx * vx + vy * vy + vz * vz) : 1.f;
    const float kk = kval<MODE>(pw, vx, vy, vz, sn.x, sn.y, sn.z, r2);
    en += act ? as * th.w * kk : 0.f;
  });
  en = xsum8(en);
  en = isum8(en);
  if (lane == 0) item_energy[it] = (double)en;
}

#pragma clang fp contract(off)

// Faces into BVH order: the FP32 packs of position p are those of face order[p].
__global__ __launch_bounds__(256) void permute_packs_kernel(int nf, const uint32_t* __restrict__ order,
                                                            const float4* __restrict__ a, const float4* __restrict__ b,
                                                            const float4* __restrict__ c, float4* __restrict__ pa,
                                                            float4* __restrict__ pb, float4* __restrict__ pc) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nf) return;
  uint32_t f = order[p];
  pa[p] = a[f];
  pb[p] = b[f];
  pc[p] = c[f];
}

// One level of the bottom-up node aggregates (bvh.rs:221-254): a leaf sums its faces in BVH order
// (area, area-weighted barycenter), an internal node adds its children.
__global__ __launch_bounds__(256) void node_aggregate_kernel(int count, const uint32_t* __restrict__ nodes,
                                                             const uint32_t* __restrict__ nstart,
                                                             const uint32_t* __restrict__ nend,
                                                             const uint32_t* __restrict__ right,
                                                             const uint32_t* __restrict__ order,
                                                             const double* __restrict__ area,
                                                             const double* __restrict__ bary, double* __restrict__ mass,
                                                             double* __restrict__ moment) {
  int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= count) return;
  uint32_t a = nodes[q];
  if (right[a] == 0) {
    double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (uint32_t p = nstart[a]; p < nend[a]; ++p) {
      uint32_t f = order[p];
      double w = area[f];
      m += w;
      mx += w * bary[3 * f];
      my += w * bary[3 * f + 1];
      mz += w * bary[3 * f + 2];
    }
    mass[a] = m;
    moment[3 * a] = mx;
    moment[3 * a + 1] = my;
    moment[3 * a + 2] = mz;
  } else {
    uint32_t l = a + 1, r = right[a];
    mass[a] = mass[l] + mass[r];
    for (int k = 0; k < 3; ++k) moment[3 * a + k] = moment[3 * l + k] + moment[3 * r + k];
  }
}

// Centroids X_C = moment / A_C and the FP32 node packs.
__global__ __launch_bounds__(256) void node_pack_kernel(int nn, const double* __restrict__ mass,
                                                        const double* __restrict__ moment, double* __restrict__ cen,
                                                        float4* __restrict__ chi, float4* __restrict__ clo) {
  int a = blockIdx.x * 256 + threadIdx.x;
  if (a >= nn) return;
  double m = mass[a];
  double x[3];
  for (int k = 0; k < 3; ++k) {
    x[k] = moment[3 * a + k] / m;
    cen[3 * a + k] = x[k];
  }
  float hx = (float)x[0], hy = (float)x[1], hz = (float)x[2];
  chi[a] = make_float4(hx, hy, hz, (float)m);
  clo[a] = make_float4((float)(x[0] - (double)hx), (float)(x[1] - (double)hy), (float)(x[2] - (double)hz), 0.f);
}

// Node sums b_C, u_C over the transposed lumped lists, FP64: one wave per node, lanes strided over
// the entries, then a fixed-order butterfly (the result does not depend on timing).
__global__ __launch_bounds__(256) void bh_node_sums_kernel(int nn, const uint32_t* __restrict__ off,
                                                           const uint32_t* __restrict__ ent,
                                                           const float4* __restrict__ node_bu,
                                                           double4* __restrict__ sums) {
  const int a = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (a >= nn) return;
  double b = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
  for (uint32_t q = off[a] + lane; q < off[a + 1]; q += 64) {
    float4 t = node_bu[ent[q]];
    b += (double)t.x;
    ux += (double)t.y;
    uy += (double)t.z;
    uz += (double)t.w;
  }
  for (int o = 32; o > 0; o >>= 1) {
    b += __shfl_xor(b, o);
    ux += __shfl_xor(ux, o);
    uy += __shfl_xor(uy, o);
    uz += __shfl_xor(uz, o);
  }
  if (lane == 0) sums[a] = make_double4(b, ux, uy, uz);
}

// Per face (BVH position p): the downward pass along the chain root → leaf (barnes_hut.rs:521-538)
// and the final partials (:540-567), written in face order (7 FP64 per face).
__global__ __launch_bounds__(256) void bh_final_kernel(int nf, const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ leaf_of_pos,
                                                       const uint32_t* __restrict__ leaves,
                                                       const uint32_t* __restrict__ nstart,
                                                       const uint32_t* __restrict__ item_ptr,
                                                       const uint32_t* __restrict__ chain_ptr,
                                                       const uint32_t* __restrict__ chain,
                                                       const double4* __restrict__ sums, const double* __restrict__ cen,
                                                       const float* __restrict__ part, const double* __restrict__ area,
                                                       const double* __restrict__ bary, double* __restrict__ out) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nf) return;
  uint32_t k = leaf_of_pos[p];
  uint32_t c0 = chain_ptr[k], c1 = chain_ptr[k + 1];
  uint32_t a = chain[c0];
  double4 s = sums[a];
  double beta = s.x, U[3] = {s.y, s.z, s.w};
  for (uint32_t q = c0 + 1; q < c1; ++q) {
    uint32_t c = chain[q];
    double sh[3] = {cen[3 * c] - cen[3 * a], cen[3 * c + 1] - cen[3 * a + 1], cen[3 * c + 2] - cen[3 * a + 2]};
    double4 sc = sums[c];
    beta = beta + (U[0] * sh[0] + U[1] * sh[1] + U[2] * sh[2]) + sc.x;
    U[0] = U[0] + sc.y;
    U[1] = U[1] + sc.z;
    U[2] = U[2] + sc.w;
    a = c;
  }
  uint32_t f = order[p];
  // the pair partials of face p: the sum over its leaf's work items, in item order
  const uint32_t row = p - nstart[leaves[k]];
  double pp[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t q = item_ptr[k]; q < item_ptr[k + 1]; ++q) {
    const float* src = part + 7 * ((size_t)q * 8 + row);
    for (int c = 0; c < 7; ++c) pp[c] += (double)src[c];
  }
  double at = area[f];
  double d[3] = {bary[3 * f] - cen[3 * a], bary[3 * f + 1] - cen[3 * a + 1], bary[3 * f + 2] - cen[3 * a + 2]};
  double* o = out + 7 * (size_t)f;
  o[0] = pp[0] + (beta + (U[0] * d[0] + U[1] * d[1] + U[2] * d[2]));
  for (int k2 = 0; k2 < 3; ++k2) o[1 + k2] = pp[1 + k2] + at * U[k2];
  for (int k2 = 0; k2 < 3; ++k2) o[4 + k2] = pp[4 + k2];
}

// ---- host ------------------------------------------------------------------------------------

void BhEval::upload_structure(ddgrs_ctx* ctx) {
  const Topology& t = topo;
  nf = (int)t.order.size();
  nn = (int)t.n_nodes();
  nl = (int)t.leaves.size();
  upload(ctx, d_order, t.order);
  upload(ctx, d_start, t.start);
  upload(ctx, d_end, t.end);
  upload(ctx, d_right, t.right);
  upload(ctx, d_leaves, t.leaves);
  std::vector<uint32_t> leaf_of_pos(nf), chain_ptr{0}, chain;
  for (int k = 0; k < nl; ++k) {
    uint32_t L = t.leaves[k];
    for (uint32_t p = t.start[L]; p < t.end[L]; ++p) leaf_of_pos[p] = (uint32_t)k;
    size_t c0 = chain.size();
    for (uint32_t x = L;; x = t.parent[x]) {
      chain.push_back(x);
      if (x == 0) break;
    }
    std::reverse(chain.begin() + c0, chain.end());
    chain_ptr.push_back((uint32_t)chain.size());
  }
  upload(ctx, d_leaf_of_pos, leaf_of_pos);
  upload(ctx, d_chain_ptr, chain_ptr);
  upload(ctx, d_chain, chain);
  // nodes grouped by depth, deepest first, for the bottom-up aggregate
  levels.assign(t.height, {});
  for (int a = 0; a < nn; ++a) levels[t.depth[a]].push_back((uint32_t)a);
  std::vector<uint32_t> flat;
  level_off.assign(1, 0);
  for (int d = (int)t.height - 1; d >= 0; --d) {
    flat.insert(flat.end(), levels[d].begin(), levels[d].end());
    level_off.push_back((uint32_t)flat.size());
  }
  upload(ctx, d_levels, flat);
  pxhi.reserve(nf);
  pxlo.reserve(nf);
  pnrm.reserve(nf);
  mass.reserve(nn);
  moment.reserve(3 * nn);
  cen.reserve(3 * nn);
  chi.reserve(nn);
  clo.reserve(nn);
  sums.reserve(nn);
}

void BhEval::refresh(ddgrs_surface* s) {
  ddgrs_ctx* ctx = s->ctx;
  launch(ctx, "bh_permute", permute_packs_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf,
         (const uint32_t*)d_order.ptr, (const float4*)s->packs.xhi.ptr, (const float4*)s->packs.xlo.ptr,
         (const float4*)s->packs.nrm.ptr, pxhi.ptr, pxlo.ptr, pnrm.ptr);
  for (size_t lv = 0; lv + 1 < level_off.size(); ++lv) {
    int cnt = (int)(level_off[lv + 1] - level_off[lv]);
    launch(ctx, "node_aggregate", node_aggregate_kernel, dim3(blocks_for(cnt, 256)), dim3(256), 0, cnt,
           (const uint32_t*)(d_levels.ptr + level_off[lv]), (const uint32_t*)d_start.ptr, (const uint32_t*)d_end.ptr,
           (const uint32_t*)d_right.ptr, (const uint32_t*)d_order.ptr, (const double*)s->area.ptr,
           (const double*)s->bary.ptr, mass.ptr, moment.ptr);
  }
  launch(ctx, "node_pack", node_pack_kernel, dim3(blocks_for(nn, 256)), dim3(256), 0, nn, (const double*)mass.ptr,
         (const double*)moment.ptr, cen.ptr, chi.ptr, clo.ptr);
}

BhArgs BhEval::args() const {
  BhArgs A;
  A.n_leaves = nl;
  A.n_items = n_items;
  A.item_leaf = item_leaf.ptr;
  A.item_g0 = item_g0.ptr;
  A.leaves = d_leaves.ptr;
  A.nstart = d_start.ptr;
  A.nend = d_end.ptr;
  A.lumped_off = d_lumped_off.ptr;
  A.lumped = d_lumped.ptr;
  A.ex_off = d_ex_off.ptr;
  A.ex_pos = d_ex_pos.ptr;
  A.pt_off = d_pt_off.ptr;
  A.pt_pos = d_pt_pos.ptr;
  A.xhi = pxhi.ptr;
  A.xlo = pxlo.ptr;
  A.nrm = pnrm.ptr;
  A.chi = chi.ptr;
  A.clo = clo.ptr;
  return A;
}

void BhEval::energy(ddgrs_ctx* ctx, const Exps& e, double* total) {
  BhArgs A = args();
  dim3 grid(blocks_for(n_items, 4));
  if (e.alpha_int == 6 && e.beta_int == 12)
    launch(ctx, "bh_energy", bh_energy_kernel<0>, grid, dim3(256), 0, A, KPow<0>{e.alpha, e.beta, 6, 12}, item_energy.ptr);
  else if (e.alpha_int >= 0 && e.beta_int >= 0)
    launch(ctx, "bh_energy", bh_energy_kernel<1>, grid, dim3(256), 0, A, KPow<1>{e.alpha, e.beta, e.alpha_int, e.beta_int},
           item_energy.ptr);
  else
    launch(ctx, "bh_energy", bh_energy_kernel<2>, grid, dim3(256), 0, A, KPow<2>{e.alpha, e.beta, -1, -1}, item_energy.ptr);
  sum_f64(ctx, n_items, item_energy.ptr, total);
}

void BhEval::gradient(ddgrs_surface* s, const Exps& e, double* total, double* partials) {
  ddgrs_ctx* ctx = s->ctx;
  BhArgs A = args();
  dim3 grid(blocks_for(n_items, 4));
  if (e.alpha_int == 6 && e.beta_int == 12)
    launch(ctx, "bh_gradient", bh_gradient_kernel<0>, grid, dim3(256), 0, A, KPow<0>{e.alpha, e.beta, 6, 12}, node_bu.ptr,
           part.ptr, item_energy.ptr);
  else if (e.alpha_int >= 0 && e.beta_int >= 0)
    launch(ctx, "bh_gradient", bh_gradient_kernel<1>, grid, dim3(256), 0, A,
           KPow<1>{e.alpha, e.beta, e.alpha_int, e.beta_int}, node_bu.ptr, part.ptr, item_energy.ptr);
  else
    launch(ctx, "bh_gradient", bh_gradient_kernel<2>, grid, dim3(256), 0, A, KPow<2>{e.alpha, e.beta, -1, -1},
           node_bu.ptr, part.ptr, item_energy.ptr);
  launch(ctx, "bh_node_sums", bh_node_sums_kernel, dim3(blocks_for(nn, 4)), dim3(256), 0, nn,
         (const uint32_t*)d_lumped_t_off.ptr, (const uint32_t*)d_lumped_t.ptr, (const float4*)node_bu.ptr, sums.ptr);
  launch(ctx, "bh_final", bh_final_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, nf, (const uint32_t*)d_order.ptr,
         (const uint32_t*)d_leaf_of_pos.ptr, (const uint32_t*)d_leaves.ptr, (const uint32_t*)d_start.ptr,
         (const uint32_t*)item_ptr.ptr, (const uint32_t*)d_chain_ptr.ptr, (const uint32_t*)d_chain.ptr,
         (const double4*)sums.ptr, (const double*)cen.ptr, (const float*)part.ptr, (const double*)s->area.ptr,
         (const double*)s->bary.ptr, partials);
  sum_f64(ctx, n_items, item_energy.ptr, total);
}

HostGeometry download_geometry(ddgrs_surface* s) {
  ddgrs_ctx* ctx = s->ctx;
  const size_t nf_ = (size_t)s->nf, nv = (size_t)s->nv;
  HostGeometry g;
  double* pin = (double*)ctx->staging((3 * nv + nf_) * sizeof(double));
  download(ctx, pin, (const double*)s->positions.ptr, 3 * nv);
  download(ctx, pin + 3 * nv, (const double*)s->area.ptr, nf_);
  ctx->sync();
  g.pos.assign(pin, pin + 3 * nv);
  g.area.assign(pin + 3 * nv, pin + 3 * nv + nf_);
  // RS/src/geometry.rs:194 vertex_areas (faces ascending)
  g.vmass.assign(nv, 0.0);
  for (size_t f = 0; f < nf_; ++f) {
    double a = g.area[f] / 3.0;
    for (int c = 0; c < 3; ++c) g.vmass[s->faces_h[3 * f + c]] += a;
  }
  return g;
}

// Face boxes: per-axis min and max of the corners (FP64, exact).
__global__ __launch_bounds__(256) void face_boxes_kernel(int64_t nf, const uint32_t* __restrict__ faces,
                                                         const double* __restrict__ pos, double* __restrict__ lo,
                                                         double* __restrict__ hi) {
  const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= nf) return;
  for (int k = 0; k < 3; ++k) {
    double a = INFINITY, b = -INFINITY;
    for (int c = 0; c < 3; ++c) {
      const double x = pos[3 * (size_t)faces[3 * f + c] + k];
      a = fmin(a, x);
      b = fmax(b, x);
    }
    lo[3 * f + k] = a;
    hi[3 * f + k] = b;
  }
}

// ---- cluster aggregates on the device ------------------------------------------------------------
//
// The sums of build_clusters (trees_host.hip) with the same operations in the same order: each
// leaf over its elements in tree order, each internal node as left + right (one launch per depth,
// deepest first), then centroid, radius and projector per node.  Box corners use the C library's
// fmin / fmax (the first argument when it is smaller, or when the second is NaN).  No contraction:
// every product is rounded before its sum.
namespace {

__device__ __forceinline__ double lib_fmin(double x, double y) { return (x < y || y != y) ? x : y; }
__device__ __forceinline__ double lib_fmax(double x, double y) { return (x > y || y != y) ? x : y; }

struct ClusterOut {  // offsets into one buffer of 27 doubles per node
  double *mass, *centroid, *lo, *hi, *radius, *projector, *mom, *ps;
};
ClusterOut cluster_out(double* b, size_t nn) {
  return {b, b + nn, b + 4 * nn, b + 7 * nn, b + 10 * nn, b + 11 * nn, b + 17 * nn, b + 20 * nn};
}

__global__ __launch_bounds__(256) void cl_leaf_kernel(uint32_t nl, const uint32_t* __restrict__ leaves,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ order, const double* __restrict__ pts,
                                                      const double* __restrict__ lo, const double* __restrict__ hi,
                                                      const double* __restrict__ mass, const double* __restrict__ nrm,
                                                      ClusterOut C) {
#pragma clang fp contract(off)
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nl) return;
  const uint32_t a = leaves[k];
  double bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
  double m = 0.0, mo[3] = {0.0, 0.0, 0.0}, o[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t q = start[a]; q < end[a]; ++q) {
    const uint32_t e = order[q];
    const double w = mass[e];
    for (int j = 0; j < 3; ++j) {
      bl[j] = lib_fmin(bl[j], lo[3 * (size_t)e + j]);
      bh[j] = lib_fmax(bh[j], hi[3 * (size_t)e + j]);
      mo[j] += w * pts[3 * (size_t)e + j];
    }
    m += w;
    if (nrm) {
      const double n0 = nrm[3 * (size_t)e], n1 = nrm[3 * (size_t)e + 1], n2 = nrm[3 * (size_t)e + 2];
      o[0] += w;
      o[1] += w * n0 * n0;
      o[2] += w * n0 * n1;
      o[3] += w * n0 * n2;
      o[4] += w * n1 * n1;
      o[5] += w * n1 * n2;
      o[6] += w * n2 * n2;
    }
  }
  C.mass[a] = m;
  for (int j = 0; j < 3; ++j) {
    C.mom[3 * (size_t)a + j] = mo[j];
    C.lo[3 * (size_t)a + j] = bl[j];
    C.hi[3 * (size_t)a + j] = bh[j];
  }
  if (nrm)
    for (int j = 0; j < 7; ++j) C.ps[7 * (size_t)a + j] = o[j];
}

__global__ __launch_bounds__(256) void cl_internal_kernel(uint32_t nn, uint32_t d, const uint32_t* __restrict__ depth,
                                                          const uint32_t* __restrict__ right, int with_ps, ClusterOut C) {
#pragma clang fp contract(off)
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a >= nn || depth[a] != d || right[a] == 0) return;
  const size_t l = a + 1, r = right[a];
  C.mass[a] = C.mass[l] + C.mass[r];
  for (int j = 0; j < 3; ++j) {
    C.mom[3 * a + j] = C.mom[3 * l + j] + C.mom[3 * r + j];
    C.lo[3 * a + j] = lib_fmin(lib_fmin(INFINITY, C.lo[3 * l + j]), C.lo[3 * r + j]);
    C.hi[3 * a + j] = lib_fmax(lib_fmax(-INFINITY, C.hi[3 * l + j]), C.hi[3 * r + j]);
  }
  if (with_ps)
    for (int j = 0; j < 7; ++j) C.ps[7 * a + j] = C.ps[7 * l + j] + C.ps[7 * r + j];
}

__global__ __launch_bounds__(256) void cl_final_kernel(uint32_t nn, int with_ps, ClusterOut C) {
#pragma clang fp contract(off)
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a >= nn) return;
  if (with_ps)
    for (int j = 0; j < 6; ++j) C.projector[6 * (size_t)a + j] = C.ps[7 * (size_t)a + j + 1] / C.ps[7 * (size_t)a];
  double x[3], r2 = 0.0;
  for (int j = 0; j < 3; ++j) {
    x[j] = C.mom[3 * (size_t)a + j] / C.mass[a];
    C.centroid[3 * (size_t)a + j] = x[j];
  }
  for (int j = 0; j < 3; ++j) {
    const double dd = lib_fmax(C.hi[3 * (size_t)a + j] - x[j], x[j] - C.lo[3 * (size_t)a + j]);
    r2 += dd * dd;
  }
  C.radius[a] = sqrt(r2);
}

}  // namespace

// build_clusters (trees_host.hip) of a tree built by `b`, computed on the device from element
// data on the device and read back.
Clusters build_clusters_device(ddgrs_ctx* ctx, GpuBvhBuilder& b, const Topology& t, const double* pts,
                               const double* lo, const double* hi, const double* mass, const double* nrm) {
  const size_t n = t.order.size(), nn = t.n_nodes(), nl = t.leaves.size();
  Clusters c;
  if (nn == 0) return c;
  const size_t words = n + 5 * nn + nl;
  uint32_t* h = (uint32_t*)ctx->arena_take(words * sizeof(uint32_t));
  uint32_t* p = h;
  for (const std::vector<uint32_t>* v : {&t.order, &t.start, &t.end, &t.right, &t.depth, &t.leaves, &t.parent}) {
    std::copy(v->begin(), v->end(), p);
    p += v->size();
  }
  b.cl_topo.reserve(words);
  HIPCHECK(hipMemcpyAsync(b.cl_topo.ptr, h, words * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
  const uint32_t* order = b.cl_topo.ptr;
  const uint32_t *start = order + n, *end = start + nn, *right = end + nn, *depth = right + nn, *leaves = depth + nn;
  b.cl_buf.reserve(27 * nn);
  const ClusterOut C = cluster_out(b.cl_buf.ptr, nn);
  const int with_ps = nrm != nullptr;
  launch(ctx, "cl_leaf", cl_leaf_kernel, dim3(blocks_for((int64_t)nl, 256)), dim3(256), 0, (uint32_t)nl, leaves, start,
         end, order, pts, lo, hi, mass, nrm, C);
  for (uint32_t d = t.height; d-- > 0;)
    launch(ctx, "cl_internal", cl_internal_kernel, dim3(blocks_for((int64_t)nn, 256)), dim3(256), 0, (uint32_t)nn, d,
           depth, right, with_ps, C);
  launch(ctx, "cl_final", cl_final_kernel, dim3(blocks_for((int64_t)nn, 256)), dim3(256), 0, (uint32_t)nn, with_ps, C);
  const size_t nout = (with_ps ? 17 : 11) * nn;
  double* out = (double*)ctx->staging(nout * sizeof(double));
  download(ctx, out, (const double*)b.cl_buf.ptr, nout);
  ctx->sync();
  c.mass.assign(out, out + nn);
  c.centroid.assign(out + nn, out + 4 * nn);
  c.lo.assign(out + 4 * nn, out + 7 * nn);
  c.hi.assign(out + 7 * nn, out + 10 * nn);
  c.radius.assign(out + 10 * nn, out + 11 * nn);
  if (with_ps) c.projector.assign(out + 11 * nn, out + 17 * nn);
  return c;
}

DeviceTree device_tree(const GpuBvhBuilder& b, const Topology& t) {
  const size_t n = t.order.size(), nn = t.n_nodes(), nl = t.leaves.size();
  const uint32_t* o = b.cl_topo.ptr;
  const ClusterOut C = cluster_out(b.cl_buf.ptr, nn);
  return {o, o + n, o + n + nn, o + n + 2 * nn, o + n + 4 * nn + nl, o + n + 4 * nn,
          C.mass, C.centroid, C.lo, C.hi, C.radius, C.projector};
}

void build_face_tree(ddgrs_surface* s, size_t leaf_size, GpuBvhBuilder& b, Topology* topo, Clusters* clusters) {
  ddgrs_ctx* ctx = s->ctx;
  const size_t nf_ = (size_t)s->nf;
  b.box_lo.reserve(3 * nf_);
  b.box_hi.reserve(3 * nf_);
  launch(ctx, "face_boxes", face_boxes_kernel, dim3(blocks_for(s->nf, 256)), dim3(256), 0, s->nf,
         (const uint32_t*)s->faces.ptr, (const double*)s->positions.ptr, b.box_lo.ptr, b.box_hi.ptr);
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  *topo = b.build(ctx, nf_, s->bary.ptr, b.box_lo.ptr, b.box_hi.ptr, leaf_size);
  auto t1 = std::chrono::steady_clock::now();
  *clusters = build_clusters_device(ctx, b, *topo, s->bary.ptr, b.box_lo.ptr, b.box_hi.ptr, s->area.ptr,
                                    s->normal.ptr);
  if (trace)
    std::fprintf(stderr, "step:   face BVH %.2f ms, clusters %.2f ms\n",
                 std::chrono::duration<double>(t1 - t0).count() * 1e3,
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count() * 1e3);
}

void build_face_tree(ddgrs_surface* s, size_t leaf_size, Topology* topo, Clusters* clusters) {
  GpuBvhBuilder b;
  build_face_tree(s, leaf_size, b, topo, clusters);
}

Topology build_vertex_tree(ddgrs_surface* s, size_t leaf_size, GpuBvhBuilder& b) {
  return b.build(s->ctx, (size_t)s->nv, s->positions.ptr, s->positions.ptr, s->positions.ptr, leaf_size);
}

void vertex_areas_host(ddgrs_surface* s, std::vector<double>* pos, std::vector<double>* mass) {
  HostGeometry g = download_geometry(s);
  *pos = std::move(g.pos);
  *mass = std::move(g.vmass);
}

void BhEval::build(ddgrs_surface* s, double theta_, size_t leaf_size_) {
  Topology t;
  Clusters c;
  build_face_tree(s, leaf_size_, &t, &c);
  build_from(s, t, c, theta_, leaf_size_);
}

void BhEval::build_from(ddgrs_surface* s, const Topology& t, const Clusters& c, double theta_, size_t leaf_size_) {
  topo = t;
  clusters = c;
  build_from_members(s, theta_, leaf_size_);
}

void BhEval::build_from_members(ddgrs_surface* s, double theta_, size_t leaf_size_) {
  theta = theta_;
  leaf_size = leaf_size_;
  ddgrs_ctx* ctx = s->ctx;
  // DDGRS_TRACE_STEP=1: wall times of the build's parts on stderr (each part synchronized)
  static const bool trace = std::getenv("DDGRS_TRACE_STEP") != nullptr;
  auto t_part = std::chrono::steady_clock::now();
  auto mark = [&](const char* what) {
    if (!trace) return;
    ctx->sync();
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "bh: %-18s %8.2f ms\n", what, std::chrono::duration<double>(now - t_part).count() * 1e3);
    t_part = now;
  };
  upload_structure(ctx);
  mark("structure");
  const char* host = std::getenv("DDGRS_BH_LISTS_ON_HOST");
  if (host && host[0] == '1') {
    lists = build_bh_lists(topo, clusters, theta);
    upload(ctx, d_lumped_off, lists.lumped_off);
    upload(ctx, d_lumped, lists.lumped);
    upload(ctx, d_exact_off, lists.exact_off);
    upload(ctx, d_exact, lists.exact);
    upload(ctx, d_exact_t_off, lists.exact_t_off);
    upload(ctx, d_exact_t, lists.exact_t);
    upload(ctx, d_lumped_t_off, lists.lumped_t_off);
    upload(ctx, d_lumped_t, lists.lumped_t);
    n_lumped = lists.lumped.size();
    n_exact = lists.exact.size();
    pairs_exact = lists.pairs_exact;
  } else {
    lists = BhLists();
    build_lists_device(ctx);
  }
  mark("lists");
  flatten_lists(ctx);
  mark("flatten");
  build_items(ctx);
  mark("items");
  node_bu.reserve(std::max<size_t>(1, n_lumped));
  refresh(s);
  mark("node data");
  if (trace) trace_lists(ctx);
}

// Per-leaf list lengths on stderr (lumped nodes, exact faces, partner faces): mean, p99, largest,
// and the 8-entry steps of the longest leaf, for reading the pair kernels' balance.
void BhEval::trace_lists(ddgrs_ctx* ctx) const {
  const uint32_t* offs[3] = {d_lumped_off.ptr, d_ex_off.ptr, d_pt_off.ptr};
  const char* names[3] = {"lumped", "exact", "partner"};
  std::vector<uint64_t> steps(nl, 0);
  for (int q = 0; q < 3; ++q) {
    std::vector<uint32_t> o(nl + 1);
    download(ctx, o.data(), offs[q], o.size());
    ctx->sync();
    std::vector<uint32_t> len(nl);
    double sum = 0.0;
    for (int k = 0; k < nl; ++k) {
      len[k] = o[k + 1] - o[k];
      sum += len[k];
      steps[k] += (len[k] + 7) / 8;
    }
    std::sort(len.begin(), len.end());
    std::fprintf(stderr, "bh lists: %-8s mean %.1f p50 %u p99 %u max %u\n", names[q], sum / nl, len[nl / 2],
                 len[(size_t)(0.99 * nl)], len.back());
  }
  uint64_t tot = 0, mx = 0;
  for (uint64_t x : steps) {
    tot += x;
    mx = std::max(mx, x);
  }
  std::fprintf(stderr, "bh lists: steps per leaf mean %.1f max %llu (total %llu)\n", (double)tot / nl,
               (unsigned long long)mx, (unsigned long long)tot);
}

BhLists BhEval::download_lists(ddgrs_ctx* ctx) const {
  BhLists L;
  L.lumped_off.resize((size_t)nl + 1);
  L.exact_off.resize((size_t)nl + 1);
  L.lumped.resize(n_lumped);
  L.exact.resize(n_exact);
  L.lumped_t_off.resize((size_t)nn + 1);
  L.lumped_t.resize(n_lumped);
  L.exact_t_off.resize((size_t)nl + 1);
  L.exact_t.resize(n_exact);
  download(ctx, L.lumped_off.data(), (const uint32_t*)d_lumped_off.ptr, L.lumped_off.size());
  download(ctx, L.exact_off.data(), (const uint32_t*)d_exact_off.ptr, L.exact_off.size());
  download(ctx, L.lumped.data(), (const uint32_t*)d_lumped.ptr, L.lumped.size());
  download(ctx, L.exact.data(), (const uint32_t*)d_exact.ptr, L.exact.size());
  download(ctx, L.lumped_t_off.data(), (const uint32_t*)d_lumped_t_off.ptr, L.lumped_t_off.size());
  download(ctx, L.lumped_t.data(), (const uint32_t*)d_lumped_t.ptr, L.lumped_t.size());
  download(ctx, L.exact_t_off.data(), (const uint32_t*)d_exact_t_off.ptr, L.exact_t_off.size());
  download(ctx, L.exact_t.data(), (const uint32_t*)d_exact_t.ptr, L.exact_t.size());
  ctx->sync();
  L.pairs_exact = pairs_exact;
  return L;
}

}  // namespace ddgrs
