#include "Framework/TaskPool.h"

#include <exception>
#include <utility>

namespace edm {
  TaskPool::TaskPool(int threads) {
    for (int t = 0; t < (threads > 0 ? threads : 1); ++t) threads_.emplace_back([this] { loop(); });
  }

  TaskPool::~TaskPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }

  void TaskPool::push(std::function<void()> task) {
    // Notify under the lock: a thread outside the pool (a device callback) may push the task that
    // lets the owner finish and destroy the pool, which cannot happen while this lock is held.
    std::lock_guard<std::mutex> g(m_);
    queue_.push_back(std::move(task));
    cv_.notify_one();
  }

  void TaskPool::loop() {
    for (;;) {
      std::function<void()> task;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [this] { return stop_ || !queue_.empty(); });
        if (queue_.empty()) return;  // stopping and drained
        task = std::move(queue_.front());
        queue_.pop_front();
      }
      try {
        task();
      } catch (...) {
        std::terminate();
      }
    }
  }
}  // namespace edm
