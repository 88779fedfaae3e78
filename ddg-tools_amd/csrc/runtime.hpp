// Runtime of the ddgrs library: context (device + one in-order stream), device buffers, kernel
// launches with optional per-kernel timing, and the error type that the C ABI turns into a status.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ddgrs.h"

namespace ddgrs {

// An error that crosses the C ABI as `ddgrs_error` (status, operation, detail).
struct Error : std::runtime_error {
  int status;
  std::string operation;
  int64_t a = -1, b = -1;
  double value = 0.0;
  Error(int st, std::string op, const std::string& detail) : std::runtime_error(detail), status(st), operation(std::move(op)) {}
};

inline void hip_check(hipError_t e, const char* op) {
  if (e != hipSuccess) throw Error(DDGRS_GPU, op, hipGetErrorString(e));
}
#define HIPCHECK(call) ::ddgrs::hip_check((call), #call)

inline void invalid(const std::string& name, const std::string& value, const std::string& requirement) {
  throw Error(DDGRS_INVALID_OPTION, name, value + ": " + requirement);
}

// A device array owned by its handle.
template <class T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIPCHECK(hipMalloc(&ptr, count * sizeof(T)));
  }
  // Grows (never shrinks) the allocation, with a quarter of headroom: sizes that creep up from one
  // step to the next then stop reallocating after a few steps (hipFree waits for the whole device,
  // which would serialize the step's streams).  Contents are not kept.
  void reserve(size_t count) {
    if (count > n || !ptr) alloc(count > 0 ? count + count / 4 : 1);
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Per-context scratch buffers (grown on demand, reused between calls).
struct Scratch {
  DevBuf<int> ints_;
  DevBuf<double> f64_;
  DevBuf<float2> f2_;
  int* ints(size_t n) {
    ints_.reserve(n);
    return ints_.ptr;
  }
  float2* f2(size_t n) {
    f2_.reserve(n);
    return f2_.ptr;
  }
  double* f64(size_t n) {
    f64_.reserve(n);
    return f64_.ptr;
  }
};
// The FP32 face data read by the pair kernels, one float4 each:
//   xhi = (X_S rounded to f32, a_S), xlo = (X_S − xhi, 0), nrm = (N_S, 0).
// X_S = xhi + xlo to about 2^-48 relative: a difference X_S − X_T is formed as
// (xhi_S − xhi_T) + (xlo_S − xlo_T), exact for close faces (Sterbenz), so the tiny N·(X_S − X_T) of
// neighbouring faces keeps its digits in FP32.
struct FacePacks {
  DevBuf<float4> xhi, xlo, nrm;
  void ensure(int64_t nf);
};
}  // namespace ddgrs

struct KernelTime {
  int64_t launches = 0;
  double ms = 0.0;
};

struct ddgrs_ctx {
  std::atomic<int> refs{1};  // the owner's reference plus one per surface
  int device = 0;
  int determinism = 1;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
  bool profiling = false;
  struct Pending {
    std::string name;
    hipEvent_t start, stop;
  };
  std::vector<Pending> pending;
  std::map<std::string, KernelTime> times;
  ddgrs::Scratch scratch;
  // Pinned staging memory for small host↔device transfers (scalars, reductions).
  void* pinned = nullptr;
  size_t pinned_bytes = 0;

  void* staging(size_t bytes) {
    if (bytes > pinned_bytes) {
      if (pinned) (void)hipHostFree(pinned);
      pinned = nullptr;
      HIPCHECK(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
      pinned_bytes = bytes;
    }
    return pinned;
  }
  // Pinned memory that uploads are copied through: a copy from pageable memory is staged by the
  // runtime through buffers of its own, shared by every host thread, and waits for them.  A region
  // of the arena is reused only after a synchronization of the stream (sync() starts the arena
  // over: the copies reading it are done).  When the arena is full before that, the uploads
  // continue in a new arena twice as large and the full one is kept until the context is released
  // (copies may still be reading it): waiting for the stream instead would also wait for every
  // kernel queued before the upload.  After the first steps one arena holds a step's uploads.
  // Past kArenaGrowLimit bytes the arena stops growing and a full one waits for the stream.
  static constexpr size_t kArenaGrowLimit = (size_t)256 << 20;
  char* arena = nullptr;
  size_t arena_bytes = 0, arena_used = 0;
  std::vector<char*> arena_retired;
  void* arena_take(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (arena_used + bytes > arena_bytes && bytes <= arena_bytes && 2 * arena_bytes > kArenaGrowLimit) {
      HIPCHECK(hipStreamSynchronize(stream));
      arena_used = 0;
    }
    if (arena_used + bytes > arena_bytes) {
      const size_t grown = std::max<size_t>(bytes, std::max<size_t>(2 * arena_bytes, (size_t)16 << 20));
      char* fresh = nullptr;
      HIPCHECK(hipHostMalloc((void**)&fresh, grown, hipHostMallocDefault));
      if (arena) arena_retired.push_back(arena);
      arena = fresh;
      arena_bytes = grown;
      arena_used = 0;
    }
    void* p = arena + arena_used;
    arena_used += bytes;
    return p;
  }
  // Collects finished event pairs into `times` (called after a synchronization).
  void harvest() {
    for (auto& p : pending) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, p.start, p.stop));
      auto& t = times[p.name];
      t.launches += 1;
      t.ms += ms;
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
    pending.clear();
  }
  void sync() {
    HIPCHECK(hipStreamSynchronize(stream));
    arena_used = 0;
    if (!pending.empty()) harvest();
  }
};

namespace ddgrs {

// Launches `kernel` on the context's stream; with profiling on, brackets it with events.
template <class K, class... Args>
inline void launch(ddgrs_ctx* ctx, const char* name, K kernel, dim3 grid, dim3 block, size_t shmem, Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->profiling) {
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, ctx->stream));
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
  HIPCHECK(hipGetLastError());
  if (ctx->profiling) {
    HIPCHECK(hipEventRecord(e1, ctx->stream));
    ctx->pending.push_back({name, e0, e1});
  }
}

template <class T>
inline void upload(ddgrs_ctx* ctx, DevBuf<T>& d, const T* h, size_t n) {
  d.reserve(n);
  if (!n) return;
  void* a = ctx->arena_take(n * sizeof(T));
  std::memcpy(a, h, n * sizeof(T));
  HIPCHECK(hipMemcpyAsync(d.ptr, a, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}
template <class T>
inline void upload(ddgrs_ctx* ctx, DevBuf<T>& d, const std::vector<T>& h) {
  upload(ctx, d, h.data(), h.size());
}
template <class T>
inline void download(ddgrs_ctx* ctx, T* h, const T* d, size_t n) {
  if (n) HIPCHECK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace ddgrs

// The triangle mesh and its positions on the device.
struct ddgrs_surface {
  std::atomic<int> refs{1};  // the owner's reference plus one per handle built on it
  ddgrs_ctx* ctx = nullptr;
  int64_t nv = 0, nf = 0;
  std::vector<uint32_t> faces_h;          // 3·nf
  ddgrs::DevBuf<uint32_t> faces;          // 3·nf
  ddgrs::DevBuf<double> positions;        // 3·nv
  // Vertex → (face, corner) incidence, faces ascending: entry = 4·face + corner.
  ddgrs::DevBuf<uint32_t> inc_ptr, inc;   // nv+1, 3·nf
  // Per-face geometry of the current positions (FP64): area, normal, barycenter.
  ddgrs::DevBuf<double> area, normal, bary;  // nf, 3nf, 3nf
  ddgrs::FacePacks packs;                 // FP32 face data of the current positions
};

namespace ddgrs {
// Lifetimes: a context lives until its owner and every surface on it have released it, a surface
// until its owner and every handle built on it have; the destroy calls may come in any order.
void ctx_release(ddgrs_ctx* ctx);
// A second context on the same device with a stream of its own (default priority), for work that
// a step runs beside its main stream; released with ctx_release.
ddgrs_ctx* ctx_create_aux(const ddgrs_ctx* main);
// Adds ctx's per-kernel times into main's and clears them (after both are synchronized).
void ctx_merge_times(ddgrs_ctx* main, ddgrs_ctx* ctx);
void surface_retain(ddgrs_surface* s);
void surface_release(ddgrs_surface* s);

// Destroys a handle that holds `ddgrs_surface* s` and releases the surface.
template <class H>
void destroy_handle(H* h) {
  if (!h) return;
  ddgrs_surface* s = h->s;
  try {
    s->ctx->sync();
  } catch (...) {
  }
  delete h;
  surface_release(s);
}
}  // namespace ddgrs
