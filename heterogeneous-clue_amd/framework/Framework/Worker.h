// Type-erased module instance owned by one EDM stream. WorkerT<T> builds T from the product
// registry (the reference requires `explicit T(edm::ProductRegistry&)`, Worker.h:51) and runs it on
// an event: for an external-work module, acquire() first, then produce() once the work signalled
// completion; an exception from either is rethrown to the caller.
#ifndef CLUE_FW_Worker_h
#define CLUE_FW_Worker_h

#include <condition_variable>
#include <exception>
#include <memory>
#include <mutex>

#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/ProductRegistry.h"

namespace edm {
  class Worker {
  public:
    virtual ~Worker() = default;
    virtual void doWork(Event& event, EventSetup const& eventSetup) = 0;
    virtual void doEndJob() = 0;
  };

  template <typename T>
  class WorkerT : public Worker {
  public:
    explicit WorkerT(ProductRegistry& reg) : producer_(reg) {}

    void doWork(Event& event, EventSetup const& eventSetup) override {
      if (producer_.hasAcquire()) {
        std::mutex m;
        std::condition_variable cv;
        bool done = false;
        std::exception_ptr err;
        producer_.doAcquire(event, eventSetup, WaitingTaskWithArenaHolder([&](std::exception_ptr e) {
                              std::lock_guard<std::mutex> g(m);
                              err = e;
                              done = true;
                              cv.notify_one();
                            }));
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return done; });
        if (err) std::rethrow_exception(err);
      }
      producer_.doProduce(event, eventSetup);
    }

    void doEndJob() override { producer_.doEndJob(); }

  private:
    T producer_;
  };
}  // namespace edm

#endif
