// One EDM stream: its own product registry and module instances (as the reference's StreamSchedule
// builds its workers, src/serial/bin/StreamSchedule.cc:24-38), run on a TaskPool. For each event the
// stream takes from its source it runs the modules in order, one task per module step (an
// external-work module's produce() is a task of its own, queued when its acquire()'s work is done),
// then asks the source for the next event (StreamSchedule.cc:55-97). A stream never blocks a pool
// thread, so while one stream's event is on the device the threads serve the other streams.
#ifndef CLUE_FW_StreamSchedule_h
#define CLUE_FW_StreamSchedule_h

#include <atomic>
#include <exception>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "Framework/EDPutToken.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"
#include "Framework/ProductRegistry.h"
#include "Framework/TaskPool.h"
#include "Framework/Worker.h"

namespace edm {
  template <typename Input>
  class StreamSchedule {
  public:
    // next(): the 0-based number of the next event to process, or -1 when the source is done.
    // finished(e): called once when this stream stops, with the exception that stopped it, if any.
    using Next = std::function<int()>;
    using Finished = std::function<void(std::exception_ptr)>;

    StreamSchedule(int id, std::vector<std::string> const& modules) : id_(id) {
      token_ = reg_.produces<Input>();
      int index = 1;
      for (auto const& m : modules) {
        reg_.beginModuleConstruction(index++);
        workers_.push_back(PluginFactory::create(m, reg_));
      }
    }

    // Starts the stream on `pool`; events[k % events.size()] is the input of event number k (id k+1).
    void start(TaskPool& pool, std::vector<Input> const& events, EventSetup const& es, Next next,
               std::atomic<int>& processed, Finished finished) {
      pool_ = &pool;
      events_ = &events;
      es_ = &es;
      next_ = std::move(next);
      processed_ = &processed;
      finished_ = std::move(finished);
      pool.push([this] { nextEvent(); });
    }

    void endJob() {
      for (auto& w : workers_) w->doEndJob();
    }

  private:
    void nextEvent() {
      event_.reset();
      const int k = next_();
      if (k < 0) {
        finished_(nullptr);
        return;
      }
      event_ = std::make_unique<Event>(id_, k + 1, reg_);
      event_->emplace(token_, (*events_)[static_cast<size_t>(k) % events_->size()]);
      runModule(0);
    }

    void runModule(size_t m) {
      if (m == workers_.size()) {
        processed_->fetch_add(1);
        pool_->push([this] { nextEvent(); });
        return;
      }
      workers_[m]->doWorkAsync(*event_, *es_, *pool_, [this, m](std::exception_ptr e) {
        if (e) {
          event_.reset();
          finished_(e);
          return;
        }
        pool_->push([this, m] { runModule(m + 1); });
      });
    }

    int id_;
    ProductRegistry reg_;
    EDPutTokenT<Input> token_;
    std::vector<std::unique_ptr<Worker>> workers_;
    TaskPool* pool_ = nullptr;
    std::vector<Input> const* events_ = nullptr;
    EventSetup const* es_ = nullptr;
    Next next_;
    std::atomic<int>* processed_ = nullptr;
    Finished finished_;
    std::unique_ptr<Event> event_;
  };
}  // namespace edm

#endif
