// grn_pool.hpp -- persistent host worker threads for the library's parallel
// loops (compiling a batch's models, copying programs to the staging buffer,
// breeding and screening candidates).  Starting and joining 15 threads costs
// a few hundred microseconds, and a screened generation runs several such
// loops; the pool's threads wait on a condition variable between loops.
#pragma once

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace grn {

// Host threads for the parallel loops: the CPUs this process may run on (its
// affinity mask), at most 16 -- the host share that goes with one GPU.
inline int host_threads() {
  cpu_set_t set;
  int n = (int)std::thread::hardware_concurrency();
  if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
  return std::max(1, std::min(16, n));
}

class HostPool {
 public:
  // f(0) ... f(n - 1), dynamically scheduled over the calling thread and up
  // to max_threads - 1 pool threads; returns when every f(i) has returned.
  // A loop started while the pool runs another one (from another thread, or
  // from inside f) runs on its calling thread alone, so loops never wait for
  // each other.  f must not throw.
  void run(int n, int max_threads, const std::function<void(int)> &f) {
    const int nt = std::min(n, std::min(max_threads, (int)workers_.size() + 1));
    std::unique_lock<std::mutex> busy(run_mu_, std::defer_lock);
    if (nt <= 1 || inside() || !busy.try_lock()) {
      for (int i = 0; i < n; ++i) f(i);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &f;
      n_ = n;
      next_.store(0, std::memory_order_relaxed);
      seats_ = nt - 1;
      ++epoch_;
    }
    wake_.notify_all();
    inside() = true;
    drain(f, n);
    inside() = false;
    std::unique_lock<std::mutex> lk(mu_);
    seats_ = 0;  // a helper that wakes from now on does not join this loop
    idle_.wait(lk, [this] { return running_ == 0; });
    job_ = nullptr;
  }

  static HostPool &instance() {
    static HostPool pool(host_threads() - 1);
    return pool;
  }

  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    wake_.notify_all();
    for (std::thread &t : workers_) t.join();
  }

  HostPool(const HostPool &) = delete;
  HostPool &operator=(const HostPool &) = delete;

 private:
  explicit HostPool(int n_workers) {
    for (int i = 0; i < n_workers; ++i) workers_.emplace_back([this] { loop(); });
  }

  // True on a thread that runs a loop's f: a loop started there runs inline
  // (the pool's own lock is never taken twice by one thread).
  static bool &inside() {
    static thread_local bool v = false;
    return v;
  }

  void drain(const std::function<void(int)> &f, int n) {
    for (int i; (i = next_.fetch_add(1, std::memory_order_relaxed)) < n;) f(i);
  }

  void loop() {
    inside() = true;
    unsigned long seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      wake_.wait(lk, [&] { return stop_ || epoch_ != seen; });
      if (stop_) return;
      seen = epoch_;
      if (seats_ <= 0 || !job_) continue;
      --seats_;
      ++running_;
      const std::function<void(int)> *f = job_;
      const int n = n_;
      lk.unlock();
      drain(*f, n);
      lk.lock();
      if (--running_ == 0) idle_.notify_all();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex run_mu_;  // held by the loop that owns the pool
  std::mutex mu_;      // guards the fields below
  std::condition_variable wake_, idle_;
  const std::function<void(int)> *job_ = nullptr;
  int n_ = 0;
  int seats_ = 0;    // helpers that may still join the current loop
  int running_ = 0;  // helpers inside the current loop
  unsigned long epoch_ = 0;
  bool stop_ = false;
  std::atomic<int> next_{0};
};

// f(0) ... f(n - 1) on up to host_threads() threads (the pool's).
template <class F>
void parallel_for(int n, F f) {
  if (n <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  const std::function<void(int)> fn(std::move(f));
  HostPool::instance().run(n, host_threads(), fn);
}

}  // namespace grn
