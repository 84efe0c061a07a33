// host_pool.h -- a process-wide pool of host worker threads for the library's parallel host
// copies (pinned staging <-> the caller's pageable memory).
#pragma once
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace storm {

// Host worker threads shared by every mapper of the process, started once: a call hands them the
// parts of one job and works on it itself until every part is done (a thread created per copy
// costs tens of microseconds, as much as copying a megabyte).  Jobs of concurrent callers (the
// serving loop's lanes) queue and are served in turn.  The workers live until the process ends
// (the library is not meant to be unloaded while the process runs).
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* p = new HostPool(7);   // never destroyed: workers may outlive static destructors
    return *p;
  }
  // runs fn(0) .. fn(n - 1), each once, on the caller and the workers; returns when all are done
  void run(int n, const std::function<void(int)>& fn) {
    if (n <= 1) {
      if (n == 1) fn(0);
      return;
    }
    Job j{&fn, n};
    {
      std::lock_guard<std::mutex> lk(mu_);
      q_.push_back(&j);
    }
    cv_.notify_all();
    work(j);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return j.done == j.n; });
  }

 private:
  struct Job {
    const std::function<void(int)>* fn;
    int n;
    int next = 0, done = 0;   // under mu_
  };
  explicit HostPool(int workers) {
    for (int t = 0; t < workers; ++t) th_.emplace_back([this] { loop(); });
    for (auto& t : th_) t.detach();
  }
  // claims the next part of job j (under mu_): -1 when every part is handed out; the job leaves
  // the queue with its last part, so no thread holds a job without a part of it in hand
  int claim(Job& j) {
    if (j.next >= j.n) return -1;
    const int i = j.next++;
    if (j.next == j.n) q_.erase(std::find(q_.begin(), q_.end(), &j));
    return i;
  }
  void finish(Job& j) {
    std::lock_guard<std::mutex> lk(mu_);
    if (++j.done == j.n) done_cv_.notify_all();   // j may be gone once its caller sees this
  }
  void work(Job& j) {   // the caller's share
    for (;;) {
      int i;
      {
        std::lock_guard<std::mutex> lk(mu_);
        i = claim(j);
      }
      if (i < 0) return;
      (*j.fn)(i);
      finish(j);
    }
  }
  void loop() {
    for (;;) {
      Job* j;
      int i;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty(); });
        j = q_.front();
        i = claim(*j);
      }
      (*j->fn)(i);
      finish(*j);
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Job*> q_;
  std::vector<std::thread> th_;
};

}  // namespace storm
