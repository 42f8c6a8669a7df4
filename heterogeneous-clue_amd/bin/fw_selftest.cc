// Framework self-test (the role of the reference's fwtest program, src/fwtest): registry and
// conditions error paths, module creation by name, external-work modules whose asynchronous part
// finishes on another thread (with and without an exception, with and without a copy of the holder),
// the holder's counting, and streams on a task pool: overlap of external work with one thread, and a
// failing stream stopping the others. Prints "ok <check>" per check and exits non-zero on the first
// failure.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <iostream>
#include <mutex>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>

#include "Framework/EDProducer.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"
#include "Framework/ProductRegistry.h"
#include "Framework/StreamSchedule.h"
#include "Framework/TaskPool.h"

namespace {
  struct Input {
    int value;
  };
  struct Output {
    int value;
  };
  struct Config {
    int add;
  };

  class Doubler : public edm::EDProducer {
  public:
    explicit Doubler(edm::ProductRegistry& reg) : in_(reg.consumes<Input>()), out_(reg.produces<Output>()) {}
    void produce(edm::Event& ev, edm::EventSetup const& es) override {
      ev.emplace(out_, Output{2 * ev.get(in_).value + es.get<Config>().add});
    }

  private:
    edm::EDGetTokenT<Input> in_;
    edm::EDPutTokenT<Output> out_;
  };
  DEFINE_FWK_MODULE(Doubler);

  // acquire() hands the work to another thread, which signals completion through the holder.
  class AsyncSquare : public edm::EDProducerExternalWork {
  public:
    explicit AsyncSquare(edm::ProductRegistry& reg) : in_(reg.consumes<Input>()), out_(reg.produces<Output>()) {}
    void acquire(edm::Event const& ev, edm::EventSetup const&, edm::WaitingTaskWithArenaHolder holder) override {
      const int v = ev.get(in_).value;
      std::thread([this, v, holder]() mutable {
        std::exception_ptr e;
        if (v < 0)
          e = std::make_exception_ptr(std::runtime_error("negative input"));
        else
          result_ = v * v;
        holder.doneWaiting(e);
      }).detach();
    }
    void produce(edm::Event& ev, edm::EventSetup const&) override { ev.emplace(out_, Output{result_}); }

  private:
    edm::EDGetTokenT<Input> in_;
    edm::EDPutTokenT<Output> out_;
    int result_ = 0;
  };
  DEFINE_FWK_MODULE(AsyncSquare);

  // acquire() keeps no copy of the holder: produce() must still follow.
  class SyncAcquire : public edm::EDProducerExternalWork {
  public:
    explicit SyncAcquire(edm::ProductRegistry& reg) : in_(reg.consumes<Input>()), out_(reg.produces<Output>()) {}
    void acquire(edm::Event const& ev, edm::EventSetup const&, edm::WaitingTaskWithArenaHolder) override {
      v_ = ev.get(in_).value + 100;
    }
    void produce(edm::Event& ev, edm::EventSetup const&) override { ev.emplace(out_, Output{v_}); }

  private:
    edm::EDGetTokenT<Input> in_;
    edm::EDPutTokenT<Output> out_;
    int v_ = 0;
  };
  DEFINE_FWK_MODULE(SyncAcquire);

  // The asynchronous part sleeps 100 ms on its own thread (a stand-in for device work).
  class SlowAsync : public edm::EDProducerExternalWork {
  public:
    explicit SlowAsync(edm::ProductRegistry& reg) : in_(reg.consumes<Input>()), out_(reg.produces<Output>()) {}
    void acquire(edm::Event const& ev, edm::EventSetup const&, edm::WaitingTaskWithArenaHolder holder) override {
      const int v = ev.get(in_).value;
      std::thread([this, v, holder]() mutable {
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
        result_ = 3 * v;
        holder.doneWaiting(v == 13 ? std::make_exception_ptr(std::runtime_error("unlucky input")) : nullptr);
      }).detach();
    }
    void produce(edm::Event& ev, edm::EventSetup const&) override { ev.emplace(out_, Output{result_}); }

  private:
    edm::EDGetTokenT<Input> in_;
    edm::EDPutTokenT<Output> out_;
    int result_ = 0;
  };
  DEFINE_FWK_MODULE(SlowAsync);

  // Adds each event's Output to a global sum.
  std::atomic<long> g_sum{0};
  class Summer : public edm::EDProducer {
  public:
    explicit Summer(edm::ProductRegistry& reg) : in_(reg.consumes<Output>()) {}
    void produce(edm::Event& ev, edm::EventSetup const&) override { g_sum += ev.get(in_).value; }

  private:
    edm::EDGetTokenT<Output> in_;
  };
  DEFINE_FWK_MODULE(Summer);

  // Runs `modules` on inputs 0..n-1 with `streams` streams on `threads` threads; returns the seconds
  // taken and the number of events processed; rethrows the first exception.
  double run_streams(std::vector<std::string> const& modules, int n, int streams, int threads, int& processed,
                     edm::EventSetup const& es) {
    std::vector<Input> inputs;
    for (int i = 0; i < n; ++i) inputs.push_back(Input{i});
    std::vector<std::unique_ptr<edm::StreamSchedule<Input>>> ss;
    for (int s = 0; s < streams; ++s) ss.push_back(std::make_unique<edm::StreamSchedule<Input>>(s, modules));
    std::atomic<int> next{0}, done{0};
    std::atomic<bool> stop{false};
    std::exception_ptr err;
    std::mutex m;
    std::condition_variable cv;
    int running = streams;
    const auto t0 = std::chrono::steady_clock::now();
    {
      edm::TaskPool pool(threads);
      for (auto& s : ss)
        s->start(
            pool, inputs, es,
            [&] {
              const int k = next.fetch_add(1);
              return (stop.load() || k >= n) ? -1 : k;
            },
            done, [&](std::exception_ptr e) {
              std::lock_guard<std::mutex> g(m);
              if (e && !err) err = e;
              if (e) stop = true;
              if (--running == 0) cv.notify_one();
            });
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] { return running == 0; });
    }
    processed = done.load();
    if (err) std::rethrow_exception(err);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  void expect(bool ok, std::string const& what) {
    if (!ok) {
      std::cout << "FAILED " << what << std::endl;
      std::exit(EXIT_FAILURE);
    }
    std::cout << "ok " << what << std::endl;
  }

  template <typename F>
  bool throws(F&& f) {
    try {
      f();
    } catch (std::runtime_error const&) {
      return true;
    }
    return false;
  }

  int run(std::string const& module, int input, edm::EventSetup const& es) {
    edm::ProductRegistry reg;
    auto src = reg.produces<Input>();
    reg.beginModuleConstruction(1);
    auto w = edm::PluginFactory::create(module, reg);
    edm::ProductRegistry readReg = reg;
    auto outTok = readReg.consumes<Output>();
    edm::Event ev(0, 1, reg);
    ev.emplace(src, Input{input});
    w->doWork(ev, es);
    return ev.get(outTok).value;
  }
}  // namespace

int main() {
  edm::ProductRegistry reg;
  reg.produces<Input>();
  expect(throws([&] { reg.produces<Input>(); }), "duplicate product registration throws");
  expect(throws([&] { reg.consumes<Output>(); }), "consuming an unproduced type throws");

  edm::EventSetup es;
  expect(throws([&] { es.get<Config>(); }), "missing conditions product throws");
  es.put(std::make_unique<Config>(Config{1}));
  expect(throws([&] { es.put(std::make_unique<Config>(Config{2})); }), "duplicate conditions product throws");
  expect(es.get<Config>().add == 1, "conditions product is readable");

  expect(throws([&] { edm::PluginFactory::create("NoSuchModule", reg); }), "unknown module name throws");
  expect(throws([&] { edm::ESPluginFactory::create("NoSuchESModule", "x"); }), "unknown ES module name throws");

  expect(run("Doubler", 20, es) == 41, "synchronous module runs on the event");
  expect(run("AsyncSquare", 7, es) == 49, "external-work module: acquire, async completion, produce");
  expect(throws([&] { run("AsyncSquare", -1, es); }), "external-work exception reaches the caller");
  expect(run("SyncAcquire", 5, es) == 105, "external-work module keeping no holder copy still produces");

  {  // counting: the continuation runs once, after the last copy, with the first exception
    int calls = 0;
    std::exception_ptr got;
    {
      edm::WaitingTaskWithArenaHolder h([&](std::exception_ptr e) {
        ++calls;
        got = e;
      });
      auto c1 = h, c2 = h;
      c1.doneWaiting(std::make_exception_ptr(std::runtime_error("first")));
      c2.doneWaiting(std::make_exception_ptr(std::runtime_error("second")));
      expect(calls == 0, "holder continuation waits for every copy");
    }
    bool first = false;
    try {
      if (got) std::rethrow_exception(got);
    } catch (std::runtime_error const& e) {
      first = std::string(e.what()) == "first";
    }
    expect(calls == 1 && first, "holder continuation runs once with the first exception");
  }

  int processed = 0;
  g_sum = 0;
  const double t = run_streams({"SlowAsync", "Summer"}, 12, 4, 1, processed, es);
  expect(processed == 12 && g_sum == 3 * (11 * 12 / 2), "12 events on 4 streams and 1 thread: every product made");
  expect(t < 0.8, "external work overlaps across streams on one thread (" + std::to_string(t) + " s for 12 x 0.1 s)");
  expect(throws([&] { run_streams({"SlowAsync", "Summer"}, 40, 4, 2, processed, es); }) && processed < 40,
         "a failing event stops every stream and its exception reaches the caller");
  std::cout << "framework self-test passed" << std::endl;
  return EXIT_SUCCESS;
}
