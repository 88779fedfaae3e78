// The exact all-pairs energy and differential without any allocation: the same kernels as
// ddgrs_exact_energy_and_differential, with their scratch memory (the per-column-chunk FP32 sums,
// the FP64 per-face partials and rows, the differential and the energy before they are copied out)
// in one file-scope device array instead of the context's on-demand buffers.
//
// Each device holds its own copy of `scratch` at its own address; the host name of the array is a
// stand-in, not a device address, so the address is looked up once per device when a context is
// created on it (noalloc_setup, from ddgrs_ctx_create) and only that address is used afterwards.
// Every call on a device shares the array and it keeps the last call's contents; no kernel here
// reads a scratch value before writing it, so nothing needs clearing between calls.  Calls that
// overlap on one device race on the array.
#include <cmath>
#include <cstring>
#include <mutex>

#include "capi_util.hpp"
#include "kernels.hpp"

namespace ddgrs {

constexpr int64_t kNoAllocMaxFaces = 131072;
constexpr int64_t kNoAllocMaxVertices = 131072;
// Chunk sums: exact_pairs_chunk_elems(nf) float2 reach 39.92 MiB for nf ≤ 2^17 (at most
// P · 8 · nf with P · ⌈nf / 256⌉ ≈ 2048 workgroups); FP64 part: 8 nf + 3 nv + 1 doubles ≤ 11.0 MiB.
constexpr size_t kChunkBytes = (size_t)40 << 20;
constexpr size_t kF64Bytes = (size_t)12 << 20;
constexpr size_t kScratchBytes = kChunkBytes + kF64Bytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_setup;
void* g_scratch[kMaxDevices] = {};  // per device: the device address of `scratch`
}  // namespace

// Looks up this device's address of `scratch` (the device must be current); once per device.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_setup);
  if (g_scratch[device]) return;
  void* ptr = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || !ptr)
    throw Error(DDGRS_GPU, "hipGetSymbolAddress(scratch)", hipGetErrorString(rc));
  g_scratch[device] = ptr;
}

namespace {
void* scratch_for(int device) {
  std::lock_guard<std::mutex> lock(g_setup);
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}
}  // namespace

}  // namespace ddgrs

using namespace ddgrs;

extern "C" {

int ddgrs_exact_energy_and_differential_noalloc(ddgrs_surface* s, double alpha, double beta, double* energy,
                                                double* differential, double* face_partials, ddgrs_error* err) {
  return guarded(err, "exact_energy_noalloc", [&] {
    need(s, "surface");
    need(energy, "energy");
    Exps e = make_exps(alpha, beta);
    ddgrs_ctx* ctx = s->ctx;
    const int64_t nf = s->nf, nv = s->nv;
    if (nf > kNoAllocMaxFaces)
      invalid("faces", std::to_string(nf), "at most 131072 faces for the no-allocation entry point");
    if (nv > kNoAllocMaxVertices)
      invalid("vertices", std::to_string(nv), "at most 131072 vertices for the no-allocation entry point");
    const size_t chunk_bytes = exact_pairs_chunk_elems(nf) * sizeof(float2);
    const size_t f64_bytes = (size_t)(8 * nf + 3 * nv + 1) * sizeof(double);
    if (chunk_bytes > kChunkBytes || f64_bytes > kF64Bytes)
      invalid("faces", std::to_string(nf), "a mesh whose scratch fits the no-allocation entry point's 52 MiB");
    unsigned char* base = (unsigned char*)scratch_for(ctx->device);
    if (!base) throw Error(DDGRS_GPU, "exact_energy_noalloc", "the device's scratch address was not set up");
    float2* chunks = (float2*)base;
    double* partials = (double*)(base + kChunkBytes);
    double* rows = partials + 7 * nf;
    double* diff = rows + nf;
    double* total = diff + 3 * nv;
    exact_pairs(ctx, nf, s->packs, s->area.ptr, e, partials, rows, chunks);
    sum_f64(ctx, nf, rows, total);
    if (differential) pull_back(s, partials, diff);
    download(ctx, energy, (const double*)total, 1);
    if (differential) download(ctx, differential, (const double*)diff, (size_t)(3 * nv));
    if (face_partials) download(ctx, face_partials, (const double*)partials, (size_t)(7 * nf));
    ctx->sync();
    bool finite = std::isfinite(*energy);
    if (finite && differential)
      for (int64_t i = 0; i < 3 * nv && finite; ++i) finite = std::isfinite(differential[i]);
    if (!finite)
      throw Error(DDGRS_NON_FINITE_ENERGY, "exact_energy_noalloc",
                  "the energy or its differential is not finite (coincident barycenters?)");
  });
}

}  // extern "C"
