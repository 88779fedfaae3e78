// Plain device copy of a known size: the reference kernel for scaling HBM byte counters
// (FETCH_SIZE, WRITE_SIZE) and a check of the achievable copy bandwidth.
//   hbm_copy [MiB=256] [repeats=5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                  \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));          \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

__global__ __launch_bounds__(256) void copy_kernel(size_t n4, const float4* __restrict__ src, float4* __restrict__ dst) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 256;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  const size_t bytes = mib << 20, n4 = bytes / sizeof(float4);
  float4 *src = nullptr, *dst = nullptr;
  CHECK(hipMalloc(&src, bytes));
  CHECK(hipMalloc(&dst, bytes));
  CHECK(hipMemset(src, 0, bytes));
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  const unsigned blocks = 256 * 32;
  copy_kernel<<<blocks, 256>>>(n4, src, dst);
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    CHECK(hipEventRecord(a));
    copy_kernel<<<blocks, 256>>>(n4, src, dst);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    if (ms < best) best = ms;
  }
  std::printf("hbm_copy: %zu MiB read + %zu MiB written per launch, best %.3f ms, %.1f GB/s\n", mib, mib, best,
              2.0 * (double)bytes / (best * 1e-3) / 1e9);
  CHECK(hipFree(src));
  CHECK(hipFree(dst));
  return 0;
}
