// Host-side parallel loop for the per-step structure builds: one pool of persistent workers (at
// most 15, plus the calling thread: the 16 host cores one GPU process is given), shared by every
// thread of the step.  A loop hands out chunks of its index range from an atomic counter, so several
// threads can run loops at once; each f(i) writes its own outputs, so the results do not depend on
// which thread runs which index.  Starting threads per loop instead costs tens of microseconds a
// thread, several times per step.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace ddgrs {

class HostPool {
 public:
  static HostPool& get() {
    static HostPool pool(std::max(1u, std::min(16u, std::thread::hardware_concurrency())) - 1);
    return pool;
  }

  // f(i) for i in [0, n); returns when every call has returned.  The bodies must not throw.
  void run(size_t n, const std::function<void(size_t)>& f) {
    Job j;
    j.n = n;
    j.chunk = std::max<size_t>(1, n / (8 * (threads_.size() + 1)));
    j.f = &f;
    {
      std::lock_guard<std::mutex> lk(m_);
      jobs_.push_back(&j);
      j.users++;  // the caller
    }
    work_cv_.notify_all();
    const size_t done = drain(&j);
    std::unique_lock<std::mutex> lk(m_);
    j.users--;
    j.finished += done;
    // no worker may pick the job up once this call has returned
    auto it = std::find(jobs_.begin(), jobs_.end(), &j);
    if (it != jobs_.end()) jobs_.erase(it);
    done_cv_.wait(lk, [&] { return j.finished == j.n && j.users == 0; });
  }

  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    work_cv_.notify_all();
    for (auto& t : threads_) t.join();
  }

 private:
  struct Job {
    size_t n = 0, chunk = 1;
    const std::function<void(size_t)>* f = nullptr;
    std::atomic<size_t> next{0};
    int users = 0;        // threads inside drain (guarded by m_)
    size_t finished = 0;  // indices done (guarded by m_)
  };

  explicit HostPool(unsigned workers) {
    for (unsigned k = 0; k < workers; ++k) threads_.emplace_back([this] { worker(); });
  }

  static size_t drain(Job* j) {
    size_t done = 0;
    for (;;) {
      const size_t i = j->next.fetch_add(j->chunk, std::memory_order_relaxed);
      if (i >= j->n) return done;
      const size_t e = std::min(j->n, i + j->chunk);
      for (size_t k = i; k < e; ++k) (*j->f)(k);
      done += e - i;
    }
  }

  void worker() {
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      work_cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
      if (stop_) return;
      Job* j = jobs_.front();
      if (j->next.load(std::memory_order_relaxed) >= j->n) {  // nothing left to hand out
        jobs_.pop_front();
        continue;
      }
      j->users++;
      lk.unlock();
      const size_t done = drain(j);
      lk.lock();
      j->users--;
      j->finished += done;
      if (j->finished == j->n && j->users == 0) done_cv_.notify_all();
    }
  }

  std::mutex m_;
  std::condition_variable work_cv_, done_cv_;
  std::deque<Job*> jobs_;
  std::vector<std::thread> threads_;
  bool stop_ = false;
};

// Calls f(i) for i in [0, n) on the pool (small loops on the calling thread).
template <class F>
void parallel(size_t n, F&& f) {
  if (n < 64) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  const std::function<void(size_t)> fn = [&f](size_t i) { f(i); };
  HostPool::get().run(n, fn);
}

}  // namespace ddgrs
