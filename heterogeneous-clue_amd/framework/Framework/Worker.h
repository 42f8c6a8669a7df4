// Type-erased module instance owned by one EDM stream. WorkerT<T> builds T from the product
// registry (the reference requires `explicit T(edm::ProductRegistry&)`, Worker.h:51) and runs it on
// an event, in one of two ways:
//  * doWork: on the calling thread; for an external-work module acquire() first, then produce()
//    once the work signalled completion; an exception from either is rethrown to the caller;
//  * doWorkAsync: as tasks of a TaskPool (the reference's Worker::doWorkAsync, Worker.h:53-96):
//    acquire() runs in the calling task and returns; when every copy of its holder has been given
//    back, produce() is queued as a new task; `done` is called exactly once, with the first
//    exception raised, if any (produce() is skipped after a failed acquire()).
#ifndef CLUE_FW_Worker_h
#define CLUE_FW_Worker_h

#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>

#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/ProductRegistry.h"
#include "Framework/TaskPool.h"

namespace edm {
  class Worker {
  public:
    virtual ~Worker() = default;
    using Done = std::function<void(std::exception_ptr)>;
    virtual void doWork(Event& event, EventSetup const& eventSetup) = 0;
    virtual void doWorkAsync(Event& event, EventSetup const& eventSetup, TaskPool& pool, Done done) = 0;
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
        acquire(event, eventSetup, [&](std::exception_ptr e) {
          std::lock_guard<std::mutex> g(m);
          err = e;
          done = true;
          cv.notify_one();
        });
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return done; });
        if (err) std::rethrow_exception(err);
      }
      producer_.doProduce(event, eventSetup);
    }

    void doWorkAsync(Event& event, EventSetup const& eventSetup, TaskPool& pool, Done done) override {
      if (!producer_.hasAcquire()) {
        done(produce(event, eventSetup));
        return;
      }
      acquire(event, eventSetup, [this, &event, &eventSetup, &pool, done](std::exception_ptr e) {
        if (e) {
          done(e);
          return;
        }
        pool.push([this, &event, &eventSetup, done] { done(produce(event, eventSetup)); });
      });
    }

    void doEndJob() override { producer_.doEndJob(); }

  private:
    // acquire() with a holder whose continuation is `then`; the framework's own copy is given back
    // after acquire() returns, with the exception acquire() threw, if any.
    void acquire(Event& event, EventSetup const& eventSetup, std::function<void(std::exception_ptr)> then) {
      WaitingTaskWithArenaHolder holder(std::move(then));
      std::exception_ptr err;
      try {
        producer_.doAcquire(event, eventSetup, holder);
      } catch (...) {
        err = std::current_exception();
      }
      holder.doneWaiting(err);
    }

    std::exception_ptr produce(Event& event, EventSetup const& eventSetup) {
      try {
        producer_.doProduce(event, eventSetup);
      } catch (...) {
        return std::current_exception();
      }
      return nullptr;
    }

    T producer_;
  };
}  // namespace edm

#endif
