// A fixed pool of host threads for index-parallel loops (window merges,
// job packing and unpacking).
#pragma once

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace spoa_amd {

// A fixed pool of host threads running index-parallel loops.
class ThreadPool {
 public:
  explicit ThreadPool(unsigned threads);
  ~ThreadPool();
  unsigned size() const { return static_cast<unsigned>(workers_.size()) + 1; }
  // runs fn(i) for i in [0, n), the calling thread included
  void For(std::size_t n, const std::function<void(std::size_t)>& fn);

 private:
  void Work();
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(std::size_t)>* fn_ = nullptr;
  std::size_t n_ = 0, next_ = 0, finished_ = 0;
  std::uint64_t epoch_ = 0;
  unsigned active_ = 0;
  bool stop_ = false;
};

}  // namespace spoa_amd
