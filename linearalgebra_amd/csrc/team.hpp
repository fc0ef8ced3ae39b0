// A team of host threads for the parallel host phases (RCM, ILU(k)
// symbolic): the caller is member 0, run(f) calls f(t, nt) on every member
// and returns when all have finished; a spin barrier for the members of one
// run; an even split of a range.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

namespace laamd {
namespace team {

// A team of threads for the parallel phases; the caller is member 0.
// run(f) calls f(t, nt) on every member and returns when all have finished.
class Team {
 public:
  explicit Team(int nt) : nt_(std::max(1, nt)) {
    for (int t = 1; t < nt_; ++t) th_.emplace_back([this, t] { worker(t); });
  }
  ~Team() {
    quit_.store(true, std::memory_order_relaxed);
    gen_.fetch_add(1, std::memory_order_release);
    for (auto &t : th_) t.join();
  }
  int size() const { return nt_; }
  void run(const std::function<void(int, int)> &f) {
    if (nt_ == 1) {
      f(0, 1);
      return;
    }
    job_ = &f;
    done_.store(0, std::memory_order_relaxed);
    gen_.fetch_add(1, std::memory_order_release);
    f(0, nt_);
    for (int spins = 0; done_.load(std::memory_order_acquire) != nt_ - 1;)
      if (++spins > 256) std::this_thread::yield();
  }

 private:
  void worker(int t) {
    uint64_t seen = 0;
    for (;;) {
      uint64_t g;
      for (int spins = 0; (g = gen_.load(std::memory_order_acquire)) == seen;)
        if (++spins > 256) std::this_thread::yield();
      seen = g;
      if (quit_.load(std::memory_order_relaxed)) return;
      (*job_)(t, nt_);
      done_.fetch_add(1, std::memory_order_release);
    }
  }
  int nt_;
  std::vector<std::thread> th_;
  const std::function<void(int, int)> *job_ = nullptr;
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> done_{0};
  std::atomic<bool> quit_{false};
};

// Sense-reversing barrier for the members of one Team::run (spins, then
// yields, so an oversubscribed host still makes progress).
class Barrier {
 public:
  explicit Barrier(int nt) : nt_(nt) {}
  void wait() {
    if (nt_ == 1) return;
    const int gen = gen_.load(std::memory_order_acquire);
    if (count_.fetch_add(1, std::memory_order_acq_rel) == nt_ - 1) {
      count_.store(0, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_release);
      return;
    }
    for (int spins = 0; gen_.load(std::memory_order_acquire) == gen;)
      if (++spins > (1 << 16)) std::this_thread::yield();
  }

 private:
  int nt_;
  std::atomic<int> count_{0}, gen_{0};
};

inline void split(int64_t m, int t, int nt, int64_t &b, int64_t &e) {
  b = m * t / nt;
  e = m * (t + 1) / nt;
}

}  // namespace team

// threads for the host phases: LAAMD_HOST_THREADS, else the hardware's, at
// most 16
int host_threads();

}  // namespace laamd
