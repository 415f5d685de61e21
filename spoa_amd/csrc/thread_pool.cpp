// Thread pool (see thread_pool.hpp).

#include "thread_pool.hpp"

#include <algorithm>

namespace spoa_amd {

ThreadPool::ThreadPool(unsigned threads) {
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  for (unsigned t = 1; t < threads; ++t) workers_.emplace_back([this] { Work(); });
}

ThreadPool::~ThreadPool() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& t : workers_) t.join();
}

void ThreadPool::Work() {
  std::uint64_t seen = 0;
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    cv_.wait(lk, [&] { return stop_ || epoch_ != seen; });
    if (stop_) return;
    seen = epoch_;
    ++active_;
    // items are taken only while the loop that published them is current
    while (epoch_ == seen && next_ < n_) {
      const std::size_t i = next_++;
      const std::function<void(std::size_t)>* fn = fn_;
      lk.unlock();
      (*fn)(i);
      lk.lock();
      ++finished_;
    }
    --active_;
    done_cv_.notify_all();
  }
}

void ThreadPool::For(std::size_t n, const std::function<void(std::size_t)>& fn) {
  if (n == 0) return;
  std::unique_lock<std::mutex> lk(mu_);
  fn_ = &fn;
  n_ = n;
  next_ = 0;
  finished_ = 0;
  ++epoch_;
  cv_.notify_all();
  while (next_ < n_) {
    const std::size_t i = next_++;
    lk.unlock();
    fn(i);
    lk.lock();
    ++finished_;
  }
  done_cv_.wait(lk, [&] { return finished_ == n_ && active_ == 0; });
  fn_ = nullptr;
}

}  // namespace spoa_amd
