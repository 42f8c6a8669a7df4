// Completion handle given to EDProducerExternalWork::acquire, with the reference's counting
// semantics (src/serial/Framework/WaitingTaskWithArenaHolder.h:18-66): every copy holds one
// reference to a pending continuation; a copy gives its reference back when doneWaiting() is called
// on it (optionally with an exception) or when it is destroyed; when the last reference is given back
// the continuation runs once, with the first exception recorded. A module copies the handle into its
// asynchronous work and calls doneWaiting() there; the framework keeps one copy across acquire() and
// gives it back afterwards, so produce() follows acquire() even when the module kept no copy. This
// build has no TBB arena: the continuation runs on the thread that gives back the last reference
// (the framework only queues a task there).
#ifndef CLUE_FW_WaitingTaskWithArenaHolder_h
#define CLUE_FW_WaitingTaskWithArenaHolder_h

#include <atomic>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <utility>

namespace edm {
  class WaitingTaskWithArenaHolder {
  public:
    WaitingTaskWithArenaHolder() = default;
    explicit WaitingTaskWithArenaHolder(std::function<void(std::exception_ptr)> f)
        : s_(std::make_shared<State>()), live_(true) {
      s_->f = std::move(f);
      s_->refs.store(1);
    }
    WaitingTaskWithArenaHolder(WaitingTaskWithArenaHolder const& o) : s_(o.s_), live_(o.live_) {
      if (live_) s_->refs.fetch_add(1);
    }
    WaitingTaskWithArenaHolder(WaitingTaskWithArenaHolder&& o) noexcept : s_(std::move(o.s_)), live_(o.live_) {
      o.live_ = false;
    }
    WaitingTaskWithArenaHolder& operator=(WaitingTaskWithArenaHolder o) {
      release(nullptr);
      s_ = std::move(o.s_);
      live_ = o.live_;
      o.live_ = false;
      return *this;
    }
    ~WaitingTaskWithArenaHolder() { release(nullptr); }

    void doneWaiting(std::exception_ptr e) { release(e); }

  private:
    struct State {
      std::function<void(std::exception_ptr)> f;
      std::atomic<int> refs{0};
      std::mutex m;
      std::exception_ptr first;
    };

    void release(std::exception_ptr e) {
      if (!live_) return;
      live_ = false;
      if (e) {
        std::lock_guard<std::mutex> g(s_->m);
        if (!s_->first) s_->first = e;
      }
      if (s_->refs.fetch_sub(1) == 1) {
        std::exception_ptr first;
        {
          std::lock_guard<std::mutex> g(s_->m);
          first = s_->first;
        }
        auto f = std::move(s_->f);
        f(first);
      }
    }

    std::shared_ptr<State> s_;
    bool live_ = false;
  };
}  // namespace edm

#endif
