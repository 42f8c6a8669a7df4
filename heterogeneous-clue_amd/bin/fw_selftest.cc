// Framework self-test (the role of the reference's fwtest program, src/fwtest): registry and
// conditions error paths, module creation by name, and external-work modules whose asynchronous
// part finishes on another thread, with and without an exception. Prints "ok <check>" per check
// and exits non-zero on the first failure.
#include <cstdlib>
#include <iostream>
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
  std::cout << "framework self-test passed" << std::endl;
  return EXIT_SUCCESS;
}
