// Completion handle given to EDProducerExternalWork::acquire. The module hands it to its
// asynchronous work and calls doneWaiting() when that work is finished (optionally with the
// exception it raised); the framework then runs produce(). This build has no TBB arena, so the
// handle wraps a plain callback (reference: src/serial/Framework/WaitingTaskWithArenaHolder.h).
#ifndef CLUE_FW_WaitingTaskWithArenaHolder_h
#define CLUE_FW_WaitingTaskWithArenaHolder_h

#include <exception>
#include <functional>
#include <memory>

namespace edm {
  class WaitingTaskWithArenaHolder {
  public:
    WaitingTaskWithArenaHolder() = default;
    explicit WaitingTaskWithArenaHolder(std::function<void(std::exception_ptr)> f)
        : f_(std::make_shared<std::function<void(std::exception_ptr)>>(std::move(f))) {}

    void doneWaiting(std::exception_ptr e) {
      if (f_ && *f_) {
        auto f = std::move(*f_);
        *f_ = nullptr;
        f(e);
      }
    }

  private:
    std::shared_ptr<std::function<void(std::exception_ptr)>> f_;
  };
}  // namespace edm

#endif
