// Copy-thread pool for host memcpy (pageable <-> pinned staging).  Plain
// C++: no HIP, so the CPU tests build it on its own.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

namespace fddm {

// Copy workers started once per process and kept until it ends.  A call of
// run() queues one job of `parts` slices; the calling thread copies slices
// too, so a job finishes even when every worker is busy with other
// callers' jobs.  Creating and joining threads per call cost about 40 us per
// thread, paid on every 64 MiB staging chunk.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();  // never destroyed: its workers live until the process ends
    return *pool;
  }
  unsigned workers() const { return nworkers_; }
  // Copies n bytes from src to dst (not overlapping) in `parts` >= 1 slices;
  // returns once every slice is copied.  Safe to call from several threads.
  void run(char* dst, const char* src, size_t n, unsigned parts) {
    Job job;
    job.dst = dst;
    job.src = src;
    job.n = n;
    job.parts = parts;
    job.slice = (n + parts - 1) / parts;
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_.push_back(&job);
    }
    cv_.notify_all();
    work(job);
    {
      std::lock_guard<std::mutex> g(mu_);  // no worker joins the job after this
      auto it = std::find(jobs_.begin(), jobs_.end(), &job);
      if (it != jobs_.end()) jobs_.erase(it);
    }
    std::unique_lock<std::mutex> lk(job.mu);
    job.cv.wait(lk, [&] { return job.done == job.parts && job.users == 0; });
  }

 private:
  struct Job {
    char* dst;
    const char* src;
    size_t n, slice;
    unsigned parts;
    std::atomic<unsigned> next{0};  // next unclaimed slice
    unsigned done = 0, users = 0;   // slices copied, workers inside work(); guarded by mu
    std::mutex mu;
    std::condition_variable cv;
  };
  static void work(Job& j) {
    for (;;) {
      const unsigned i = j.next.fetch_add(1);
      if (i >= j.parts) return;
      const size_t b = i * j.slice, e = std::min(j.n, b + j.slice);
      if (b < e) std::memcpy(j.dst + b, j.src + b, e - b);
      std::lock_guard<std::mutex> g(j.mu);
      if (++j.done == j.parts) j.cv.notify_all();
    }
  }
  void worker() {
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !jobs_.empty(); });
        j = jobs_.front();
        if (j->next.load() >= j->parts) {  // every slice claimed: the job leaves the queue
          jobs_.pop_front();
          continue;
        }
        std::lock_guard<std::mutex> g(j->mu);
        ++j->users;
      }
      work(*j);
      std::lock_guard<std::mutex> g(j->mu);
      --j->users;
      j->cv.notify_all();
    }
  }
  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    nworkers_ = std::min(15u, hw > 1 ? hw - 1 : 1u);
    for (unsigned i = 0; i < nworkers_; ++i) std::thread([this] { worker(); }).detach();
  }
  unsigned nworkers_ = 0;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job*> jobs_;
};

}  // namespace fddm
