// The C ABI of include/ddgrs.h: context, surface and energy entry points.  Every entry point
// catches the library's errors and returns them as a status with (operation, detail).
#include <algorithm>
#include <cmath>

#include "capi_util.hpp"

using namespace ddgrs;


namespace ddgrs {

Exps make_exps(double alpha, double beta) {
  auto bad = [&](const char* req) {
    Error e(DDGRS_INVALID_EXPONENTS, "exponents", req);
    throw e;
  };
  // RS/src/kernel.rs:49 Exponents::new
  if (!std::isfinite(alpha) || !std::isfinite(beta)) bad("both exponents must be finite");
  if (alpha <= 1.0) bad("alpha must be greater than 1");
  if (beta <= alpha + 2.0) bad("beta must be greater than alpha + 2 (s = (beta - 2)/alpha > 1)");
  if (beta >= 2.0 * alpha + 2.0) bad("beta must be smaller than 2 alpha + 2 (s = (beta - 2)/alpha < 2)");
  auto small_int = [](double x, double max) {
    return (x == std::trunc(x) && x > 0.0 && x <= max) ? (int)(uint32_t)x : -1;
  };
  Exps e;
  e.alpha = (float)alpha;
  e.beta = (float)beta;
  e.alpha64 = alpha;
  e.beta64 = beta;
  e.alpha_int = small_int(alpha, 64.0);
  e.beta_int = small_int(beta, 128.0);
  return e;
}

}  // namespace ddgrs

namespace ddgrs {

void ctx_release(ddgrs_ctx* ctx) {
  if (!ctx || --ctx->refs > 0) return;
  try {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->arena) (void)hipHostFree(ctx->arena);
    for (char* a : ctx->arena_retired) (void)hipHostFree(a);
    ctx->scratch = Scratch();
    (void)hipStreamDestroy(ctx->stream);
  } catch (...) {
  }
  delete ctx;
}

ddgrs_ctx* ctx_create_aux(const ddgrs_ctx* main) {
  auto ctx = std::make_unique<ddgrs_ctx>();
  ctx->device = main->device;
  ctx->determinism = main->determinism;
  ctx->prop = main->prop;
  HIPCHECK(hipSetDevice(main->device));
  HIPCHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  return ctx.release();
}

void ctx_merge_times(ddgrs_ctx* main, ddgrs_ctx* ctx) {
  for (auto& kv : ctx->times) {
    auto& t = main->times[kv.first];
    t.launches += kv.second.launches;
    t.ms += kv.second.ms;
  }
  ctx->times.clear();
}

void surface_retain(ddgrs_surface* s) { ++s->refs; }

void surface_release(ddgrs_surface* s) {
  if (!s || --s->refs > 0) return;
  ddgrs_ctx* ctx = s->ctx;
  try {
    ctx->sync();
  } catch (...) {
  }
  delete s;
  ctx_release(ctx);
}

}  // namespace ddgrs

extern "C" {

int ddgrs_ctx_create(int32_t ordinal, int32_t determinism, ddgrs_ctx** out, ddgrs_error* err) {
  return guarded(err, "ctx_create", [&] {
    need(out, "out");
    *out = nullptr;
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n)
      invalid("ordinal", std::to_string(ordinal), "a device index below " + std::to_string(n));
    if (determinism != 1) invalid("determinism", std::to_string(determinism), "1 (run-to-run reproducible)");
    auto ctx = std::make_unique<ddgrs_ctx>();
    ctx->device = ordinal;
    ctx->determinism = determinism;
    HIPCHECK(hipSetDevice(ordinal));
    HIPCHECK(hipGetDeviceProperties(&ctx->prop, ordinal));
    HIPCHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    noalloc_setup(ordinal);
    *out = ctx.release();
  });
}

void ddgrs_ctx_destroy(ddgrs_ctx* ctx) { ctx_release(ctx); }

int ddgrs_device_info(ddgrs_ctx* ctx, char* name, int64_t name_len, int64_t* compute_units, int64_t* memory_bytes,
                      ddgrs_error* err) {
  return guarded(err, "device_info", [&] {
    need(ctx, "ctx");
    if (name && name_len > 0) {
      if (ctx->prop.name[0])
        std::snprintf(name, (size_t)name_len, "%s (%s)", ctx->prop.name, ctx->prop.gcnArchName);
      else  // some runtimes leave the marketing name empty
        std::snprintf(name, (size_t)name_len, "AMD GPU (%s)", ctx->prop.gcnArchName);
    }
    if (compute_units) *compute_units = ctx->prop.multiProcessorCount;
    if (memory_bytes) *memory_bytes = (int64_t)ctx->prop.totalGlobalMem;
  });
}

int ddgrs_ctx_synchronize(ddgrs_ctx* ctx, ddgrs_error* err) {
  return guarded(err, "synchronize", [&] {
    need(ctx, "ctx");
    ctx->sync();
  });
}

int ddgrs_ctx_set_profiling(ddgrs_ctx* ctx, int32_t enabled, ddgrs_error* err) {
  return guarded(err, "set_profiling", [&] {
    need(ctx, "ctx");
    ctx->sync();
    ctx->profiling = enabled != 0;
    ctx->times.clear();
  });
}

int ddgrs_ctx_kernel_times(ddgrs_ctx* ctx, int64_t capacity, char* names, int64_t* launches, double* ms, int64_t* count,
                           ddgrs_error* err) {
  return guarded(err, "kernel_times", [&] {
    need(ctx, "ctx");
    ctx->sync();
    int64_t i = 0;
    for (auto& kv : ctx->times) {
      if (i < capacity) {
        std::snprintf(names + 64 * i, 64, "%s", kv.first.c_str());
        launches[i] = kv.second.launches;
        ms[i] = kv.second.ms;
      }
      ++i;
    }
    if (count) *count = i;
  });
}

int ddgrs_surface_create(ddgrs_ctx* ctx, int64_t nv, int64_t nf, const uint32_t* faces, const double* positions,
                         ddgrs_surface** out, ddgrs_error* err) {
  return guarded(err, "surface_create", [&] {
    need(ctx, "ctx");
    need(out, "out");
    need(faces, "faces");
    need(positions, "positions");
    *out = nullptr;
    if (nv <= 0 || nf <= 0) invalid("mesh", std::to_string(nv) + " vertices, " + std::to_string(nf) + " faces", "a non-empty mesh");
    if (nf >= (int64_t)1 << 30 || nv >= (int64_t)1 << 31) invalid("mesh", std::to_string(nf) + " faces", "fewer than 2^30 faces");
    auto s = std::make_unique<ddgrs_surface>();
    s->ctx = ctx;
    s->nv = nv;
    s->nf = nf;
    s->faces_h.assign(faces, faces + 3 * nf);
    // Incidence vertex → (face, corner), faces ascending (counting sort in face order).
    std::vector<uint32_t> ptr(nv + 1, 0), inc(3 * nf);
    for (int64_t f = 0; f < nf; ++f)
      for (int c = 0; c < 3; ++c) {
        uint32_t v = faces[3 * f + c];
        if ((int64_t)v >= nv) {
          Error e(DDGRS_INVALID_OPTION, "faces", "face " + std::to_string(f) + " references vertex " + std::to_string(v) + " of " + std::to_string(nv));
          e.a = f;
          throw e;
        }
        ptr[v + 1]++;
      }
    for (int64_t v = 0; v < nv; ++v) {
      if (ptr[v + 1] == 0) {
        Error e(DDGRS_INVALID_OPTION, "faces", "vertex " + std::to_string(v) + " is not referenced by any face");
        e.a = v;
        throw e;
      }
      ptr[v + 1] += ptr[v];
    }
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t f = 0; f < nf; ++f)
      for (int c = 0; c < 3; ++c) inc[fill[faces[3 * f + c]]++] = (uint32_t)(4 * f + c);
    HIPCHECK(hipSetDevice(ctx->device));
    upload(ctx, s->faces, s->faces_h);
    upload(ctx, s->inc_ptr, ptr);
    upload(ctx, s->inc, inc);
    upload(ctx, s->positions, positions, (size_t)(3 * nv));
    compute_face_geometry(s.get(), s->packs, true);
    ++ctx->refs;
    *out = s.release();
  });
}

void ddgrs_surface_destroy(ddgrs_surface* s) { surface_release(s); }

int ddgrs_surface_set_positions(ddgrs_surface* s, const double* positions, ddgrs_error* err) {
  return guarded(err, "set_positions", [&] {
    need(s, "surface");
    need(positions, "positions");
    upload(s->ctx, s->positions, positions, (size_t)(3 * s->nv));
    compute_face_geometry(s, s->packs, true);
  });
}

int ddgrs_surface_get_positions(ddgrs_surface* s, double* positions, ddgrs_error* err) {
  return guarded(err, "get_positions", [&] {
    need(s, "surface");
    need(positions, "positions");
    download(s->ctx, positions, s->positions.ptr, (size_t)(3 * s->nv));
    s->ctx->sync();
  });
}

int ddgrs_exact_energy_and_differential(ddgrs_surface* s, double alpha, double beta, double* energy, double* differential,
                                        double* face_partials, ddgrs_error* err) {
  return guarded(err, "exact_energy", [&] {
    need(s, "surface");
    need(energy, "energy");
    Exps e = make_exps(alpha, beta);
    ddgrs_ctx* ctx = s->ctx;
    const int64_t nf = s->nf, nv = s->nv;
    double* buf = ctx->scratch.f64((size_t)(7 * nf + nf + 3 * nv + 1));
    double* partials = buf;
    double* rows = buf + 7 * nf;
    double* diff = rows + nf;
    double* total = diff + 3 * nv;
    exact_pairs(ctx, nf, s->packs, s->area.ptr, e, partials, rows);
    sum_f64(ctx, nf, rows, total);
    if (differential) pull_back(s, partials, diff);
    double* pin = (double*)ctx->staging(sizeof(double));
    download(ctx, pin, total, 1);
    if (differential) download(ctx, differential, (const double*)diff, (size_t)(3 * nv));
    if (face_partials) download(ctx, face_partials, (const double*)partials, (size_t)(7 * nf));
    ctx->sync();
    *energy = pin[0];
    bool finite = std::isfinite(*energy);
    if (finite && differential)
      for (int64_t i = 0; i < 3 * nv && finite; ++i) finite = std::isfinite(differential[i]);
    if (!finite) {
      Error er(DDGRS_NON_FINITE_ENERGY, "exact_energy", "the energy or its differential is not finite (coincident barycenters?)");
      throw er;
    }
  });
}

}  // extern "C"
