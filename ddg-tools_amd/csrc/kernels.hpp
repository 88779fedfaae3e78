// Shared declarations of the library's kernels' host wrappers and device helpers.
#pragma once
#include "runtime.hpp"

namespace ddgrs {


inline Scratch& scratch_of(ddgrs_ctx* ctx) { return ctx->scratch; }

// MI355X places consecutive workgroups on its 8 XCDs in turn, each XCD with its own L2.  For
// work ordered by locality (BVH leaves, panels) this gives each XCD a contiguous eighth of the
// blocks instead: a bijection of [0, n) taking the blocks that run on XCD x to
// [x·⌊n/8⌋ + min(x, n mod 8), …).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t x = b & 7u, i = b >> 3, per = n >> 3, rem = n & 7u;
  return x * per + (x < rem ? x : rem) + i;
}
// The wave's index in a 256-thread launch with XCD-contiguous blocks (wave-uniform).
__device__ __forceinline__ uint32_t xcd_wave_id() {
  return __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
}

// Hand-off of values between waves of one launch (last arriver of a counter reads what the others
// wrote) without a device-scope fence, which on this GPU writes back and invalidates the whole L2
// of the XCD.  The values are written with device-scope atomic stores and read with device-scope
// atomic loads (both coherent across the XCDs' L2s); `published()` waits until the wave's stores
// have completed, which orders them before the arrival counter's atomic add.  This is the
// compiler's own release/acquire sequence for device scope minus the L2 write-back and
// invalidation, which only concern data written with plain stores.
template <class T>
__device__ __forceinline__ void publish(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T collect(T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void published() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Exponents as the kernels take them (RS/src/kernel.rs:26 Exponents): integer exponents use
// products, others powf.
struct Exps {
  float alpha, beta;
  int alpha_int, beta_int;  // -1 when not an integer (kernel.rs:37 as_small_int)
  double alpha64, beta64;
};
Exps make_exps(double alpha, double beta);  // throws DDGRS_INVALID_EXPONENTS

// Geometry of the surface's current positions into s->area/normal/bary and the packs;
// `validate` checks the positions as RS/src/surface.rs:513 (synchronizes).
void compute_face_geometry(ddgrs_surface* s, FacePacks& packs, bool validate);
// out[3v+k] = Σ_{S∋v} face_pull_back(S)[corner]; partials: 7 per face.
void pull_back(ddgrs_surface* s, const double* partials, double* out);
// out[0] = Σ x (fixed order).
void sum_f64(ddgrs_ctx* ctx, int64_t n, const double* x, double* out);

// Exact all-pairs energy: per-face partials (7·nf, FP64) and per-face energy rows (nf, FP64).
// chunk_partials: exact_pairs_chunk_elems(nf) float2 of device memory for the per-column-chunk sums,
// or null for the context's scratch.
void exact_pairs(ddgrs_ctx* ctx, int64_t nf, const FacePacks& packs, const double* area, const Exps& e,
                 double* partials, double* row_energy, float2* chunk_partials = nullptr);
// The float2 count of exact_pairs' per-column-chunk sums for nf faces (the partition depends on nf only).
size_t exact_pairs_chunk_elems(int64_t nf);
// Looks up the current device's address of the no-allocation entry point's scratch array (once per
// device; called when a context is created on it).
void noalloc_setup(int device);

}  // namespace ddgrs
