// A team of host threads for the parallel host phases (RCM, ILU(k)
// symbolic): the caller is member 0, run(f) calls f(t, nt) on every member
// and returns when all have finished; a spin barrier for the members of one
// run; an even split of a range.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace laamd {
namespace team {

// Thrown out of Barrier::wait in the other members once one member's f has
// thrown, so every member leaves the run instead of waiting for it.
struct Aborted {};

// A team of threads for the parallel phases; the caller is member 0.
// run(f) calls f(t, nt) on every member and returns when all have finished.
// If some member's f throws, the run is aborted (members waiting on a
// Barrier of this team leave it) and run() rethrows the first exception
// once every member is out; the team stays usable.
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
  const std::atomic<bool> *abort_flag() const { return &abort_; }
  void run(const std::function<void(int, int)> &f) {
    if (nt_ == 1) {
      f(0, 1);
      return;
    }
    job_ = &f;
    err_ = nullptr;
    abort_.store(false, std::memory_order_relaxed);
    done_.store(0, std::memory_order_relaxed);
    gen_.fetch_add(1, std::memory_order_release);
    call(0);
    for (int spins = 0; done_.load(std::memory_order_acquire) != nt_ - 1;)
      if (++spins > 256) std::this_thread::yield();
    if (err_) std::rethrow_exception(err_);
  }

 private:
  void call(int t) {
    try {
      (*job_)(t, nt_);
    } catch (const Aborted &) {
    } catch (...) {
      {
        std::lock_guard<std::mutex> g(mu_);
        if (!err_) err_ = std::current_exception();
      }
      abort_.store(true, std::memory_order_release);
    }
  }
  void worker(int t) {
    uint64_t seen = 0;
    for (;;) {
      uint64_t g;
      for (int spins = 0; (g = gen_.load(std::memory_order_acquire)) == seen;)
        if (++spins > 256) std::this_thread::yield();
      seen = g;
      if (quit_.load(std::memory_order_relaxed)) return;
      call(t);
      done_.fetch_add(1, std::memory_order_release);
    }
  }
  int nt_;
  std::vector<std::thread> th_;
  const std::function<void(int, int)> *job_ = nullptr;
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> done_{0};
  std::atomic<bool> quit_{false}, abort_{false};
  std::mutex mu_;
  std::exception_ptr err_;
};

// Sense-reversing barrier for the members of one Team::run (spins, then
// yields, so an oversubscribed host still makes progress). Throws Aborted
// when the team's run has been aborted.
class Barrier {
 public:
  explicit Barrier(const Team &T) : nt_(T.size()), abort_(T.abort_flag()) {}
  void wait() {
    if (nt_ == 1) return;
    const int gen = gen_.load(std::memory_order_acquire);
    if (count_.fetch_add(1, std::memory_order_acq_rel) == nt_ - 1) {
      count_.store(0, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_release);
      return;
    }
    for (int spins = 0; gen_.load(std::memory_order_acquire) == gen;) {
      if (abort_->load(std::memory_order_acquire)) throw Aborted{};
      if (++spins > (1 << 16)) std::this_thread::yield();
    }
  }

 private:
  int nt_;
  const std::atomic<bool> *abort_;
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
