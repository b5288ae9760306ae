// Host worker threads shared by the library's parallel host loops (path
// re-validation, negotiation planes and overrides, per-field override
// tables).  Creating and joining threads per call costs 60-120 us on the GPU
// hosts, several times per negotiation iteration; the workers here are
// created once and parked on a condition variable between jobs.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include <pthread.h>
#include <sched.h>

namespace mrf {
namespace host {

class HostPool {
 public:
  // body(begin, end) over [0, n) in chunks of at least min_chunk, on the
  // calling thread and up to max_threads - 1 workers; about four chunks per
  // thread, claimed in turn, so uneven items (paths of different lengths)
  // even out.  One job runs at a time (concurrent callers wait their turn); a
  // call from inside a job runs inline.  The first exception of any chunk is
  // rethrown here.
  void run(size_t n, size_t min_chunk, size_t max_threads, const std::function<void(size_t, size_t)>& body) {
    if (n == 0) return;
    const size_t want = std::min<size_t>(max_threads, (n + min_chunk - 1) / std::max<size_t>(min_chunk, 1));
    if (want <= 1 || in_job_) {
      body(0, n);
      return;
    }
    std::lock_guard<std::mutex> one_job(submit_mu_);
    auto job = std::make_shared<Job>();
    job->body = &body;
    job->n = n;
    job->per = std::max<size_t>(std::max<size_t>(min_chunk, 1), (n + 4 * want - 1) / (4 * want));
    job->chunks = (n + job->per - 1) / job->per;
    job->pending = job->chunks;
    {
      std::lock_guard<std::mutex> lk(mu_);
      while (workers_.size() < want - 1) workers_.emplace_back([this] { loop(); });
      current_ = job;
      ++generation_;
    }
    cv_work_.notify_all();
    work(*job);
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return job->pending == 0; });
    current_.reset();
    if (job->error) std::rethrow_exception(job->error);
  }

  // One pool per process, never destroyed: parked workers end with the
  // process, and nothing joins threads during static destruction.
  static HostPool& get() {
    static HostPool* pool = [] {
      HostPool* p = new HostPool();
      pthread_atfork(nullptr, nullptr, [] { get().after_fork_in_child(); });
      return p;
    }();
    return *pool;
  }

 private:
  // One call's work: chunks are claimed from `next`; `pending` (under mu_)
  // counts chunks not yet finished.  A worker that wakes late holds its own
  // reference and finds no chunk left.
  struct Job {
    const std::function<void(size_t, size_t)>* body = nullptr;
    size_t n = 0, per = 0, chunks = 0, pending = 0;
    std::atomic<size_t> next{0};
    std::exception_ptr error;
  };

  // A forked child has the parent's pool but only the forking thread: the
  // locks and condition variables may record the parent's waiters, so they
  // are made new, and the parent's thread handles are set aside (never
  // joined); the child starts workers of its own on its first job.
  void after_fork_in_child() {
    new (&submit_mu_) std::mutex();
    new (&mu_) std::mutex();
    new (&cv_work_) std::condition_variable();
    new (&cv_done_) std::condition_variable();
    if (!workers_.empty()) new std::vector<std::thread>(std::move(workers_));
    workers_.clear();
    current_.reset();
    stop_ = false;
  }

  void work(Job& job) {
    const bool outer = in_job_;
    in_job_ = true;
    for (;;) {
      const size_t c = job.next.fetch_add(1);
      if (c >= job.chunks) break;
      const size_t a = c * job.per, b = std::min(job.n, a + job.per);
      std::exception_ptr err;
      try {
        (*job.body)(a, b);
      } catch (...) {
        err = std::current_exception();
      }
      std::lock_guard<std::mutex> lk(mu_);
      if (err && !job.error) job.error = err;
      if (--job.pending == 0) cv_done_.notify_all();
    }
    in_job_ = outer;
  }

  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::shared_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] { return stop_ || generation_ != seen; });
        if (stop_) return;
        seen = generation_;
        job = current_;
      }
      if (job) work(*job);
    }
  }

  std::mutex submit_mu_, mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::thread> workers_;
  std::shared_ptr<Job> current_;
  uint64_t generation_ = 0;
  bool stop_ = false;
  static thread_local bool in_job_;
};

inline thread_local bool HostPool::in_job_ = false;

// CPUs this process may run on (its affinity mask), read once.
inline size_t usable_cpus() {
  static const size_t n = [] {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return (size_t)std::max(1, CPU_COUNT(&set));
    return (size_t)std::max(1u, std::thread::hardware_concurrency());
  }();
  return n;
}

// body(begin, end) over [0, n) on up to 16 host threads (the caller's and
// the shared workers), no more than the process's CPUs.
template <typename F>
void parallel_for(size_t n, size_t min_chunk, F&& body) {
  HostPool::get().run(n, min_chunk, std::min<size_t>(16, usable_cpus()), std::function<void(size_t, size_t)>(body));
}

}  // namespace host
}  // namespace mrf
