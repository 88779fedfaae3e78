// The dense stage of the step's recovery chain (RS/src/step/direction.rs:259-338 RobustHierarchical,
// `escalate` :277, stage 2): when GMRES has diverged twice, the saddle system with the exact metric
// (every face pair in the near field: the block cluster tree at χ = 0 with the whole surface as one
// dense block) is formed column by column on the device and solved by an LU factorization with
// partial pivoting (rocSOLVER getrf / getrs).  Meshes of at most DENSE_FALLBACK_MAX_VERTICES
// (direction.rs:238: 6000) vertices; the matrix is (3 nv + 3)² FP64 (2.6 GB at the limit).
//
// rocSOLVER is opened with dlopen when this stage first runs, so the library does not load it (or
// rocBLAS) for the steps that never need it.
#include <dlfcn.h>

#include <cmath>
#include <mutex>
#include <string>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "robust.hpp"

namespace ddgrs {

namespace {

struct RocSolver {
  decltype(&rocblas_create_handle) create = nullptr;
  decltype(&rocblas_destroy_handle) destroy = nullptr;
  decltype(&rocblas_set_stream) set_stream = nullptr;
  decltype(&rocsolver_dgetrf) getrf = nullptr;
  decltype(&rocsolver_dgetrs) getrs = nullptr;
};

const RocSolver& rocsolver() {
  static std::once_flag once;
  static RocSolver r;
  static std::string failure;
  std::call_once(once, [] {
    void* h = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("librocsolver.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      failure = dlerror();
      return;
    }
    r.create = (decltype(r.create))dlsym(h, "rocblas_create_handle");
    r.destroy = (decltype(r.destroy))dlsym(h, "rocblas_destroy_handle");
    r.set_stream = (decltype(r.set_stream))dlsym(h, "rocblas_set_stream");
    r.getrf = (decltype(r.getrf))dlsym(h, "rocsolver_dgetrf");
    r.getrs = (decltype(r.getrs))dlsym(h, "rocsolver_dgetrs");
    if (!r.create || !r.destroy || !r.set_stream || !r.getrf || !r.getrs) failure = "a rocSOLVER entry point is missing";
  });
  if (!failure.empty()) throw Error(DDGRS_GPU, "dense fallback", "rocSOLVER: " + failure);
  return r;
}

__global__ void set_entry_kernel(double* __restrict__ v, size_t i, double value) {
  if (threadIdx.x == 0) v[i] = value;
}

__global__ __launch_bounds__(256) void zero_kernel(size_t n, double* __restrict__ v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

__global__ __launch_bounds__(256) void sub_kernel(size_t n, const double* __restrict__ b, double* __restrict__ r) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

void check_rocblas(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    throw Error(DDGRS_GPU, "dense fallback", std::string(what) + " failed (rocblas_status " + std::to_string((int)st) + ")");
}

}  // namespace

double dense_saddle_solve(ddgrs_ctx* ctx, ddgrs_surface* s, const Topology& t, const Clusters& c, double s_exp,
                          double w_hi, double w_lo, const double* crow, const double* b, double* x) {
  const size_t nv = (size_t)s->nv, n = 3 * nv + 3;
  // the exact metric: no admissible block, the root block dense
  FaceMetricDev fm;
  fm.w_hi = w_hi;
  fm.w_lo = w_lo;
  const size_t all_pairs = (size_t)s->nf * (size_t)s->nf;
  fm.build(ctx, s, t, c, 0.0, all_pairs, s_exp);
  // FP64 differences and sums on the FP32 values: the operator is then linear to FP64 rounding, so
  // its columns and the residual check describe one matrix (the FP32 products are linear only to
  // their rounding, about 1e-7 relative)
  fm.extended = true;
  SaddleOp op;
  op.fm = &fm;
  op.crow = crow;
  op.nv = nv;
  DevBuf<double> A(n * n), e(n), rhs(n), r(n);
  DevBuf<rocblas_int> ipiv(n), info(1);
  launch(ctx, "dense_zero", zero_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, n, e.ptr);
  for (size_t j = 0; j < n; ++j) {
    launch(ctx, "dense_unit", set_entry_kernel, dim3(1), dim3(64), 0, e.ptr, j, 1.0);
    op.apply(ctx, e.ptr, A.ptr + j * n);  // column j of A′ (column-major)
    launch(ctx, "dense_unit", set_entry_kernel, dim3(1), dim3(64), 0, e.ptr, j, 0.0);
  }
  HIPCHECK(hipMemcpyAsync(rhs.ptr, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  const RocSolver& rs = rocsolver();
  rocblas_handle h = nullptr;
  check_rocblas(rs.create(&h), "rocblas_create_handle");
  rocblas_int inf = 0;
  try {
    check_rocblas(rs.set_stream(h, ctx->stream), "rocblas_set_stream");
    check_rocblas(rs.getrf(h, (rocblas_int)n, (rocblas_int)n, A.ptr, (rocblas_int)n, ipiv.ptr, info.ptr), "rocsolver_dgetrf");
    check_rocblas(rs.getrs(h, rocblas_operation_none, (rocblas_int)n, 1, A.ptr, (rocblas_int)n, ipiv.ptr, rhs.ptr,
                           (rocblas_int)n),
                  "rocsolver_dgetrs");
    download(ctx, &inf, (const rocblas_int*)info.ptr, 1);
    ctx->sync();
  } catch (...) {
    rs.destroy(h);
    throw;
  }
  rs.destroy(h);
  if (inf != 0) {
    Error er(DDGRS_SINGULAR_SADDLE, "dense fallback", "zero pivot in the LU of the dense saddle matrix");
    er.a = inf;
    throw er;
  }
  HIPCHECK(hipMemcpyAsync(x, rhs.ptr, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  // the explicit relative residual ‖b − A′ x‖ / ‖b‖ of the exact-metric system
  op.apply(ctx, x, r.ptr);
  launch(ctx, "dense_residual", sub_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, n, b, r.ptr);
  DeviceVecs vec;
  vec.init(n);
  const double rn = vec.norm_host(ctx, r.ptr), bn = vec.norm_host(ctx, b);
  return bn > 0.0 ? rn / bn : rn;
}

}  // namespace ddgrs
