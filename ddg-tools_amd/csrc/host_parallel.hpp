// Host-side parallel loop for the per-step structure builds (at most 16 threads: the share of
// host cores one GPU process is given).
#pragma once
#include <algorithm>
#include <thread>
#include <vector>

namespace ddgrs {

// Calls f(i) for i in [0, n), interleaved over the threads.
template <class F>
void parallel(size_t n, F&& f) {
  unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 64 || t == 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned k = 0; k < t; ++k)
    th.emplace_back([&, k] {
      for (size_t i = k; i < n; i += t) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace ddgrs
