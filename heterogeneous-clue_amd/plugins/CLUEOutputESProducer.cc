// Puts the output directory (a std::filesystem::path) into the EventSetup, the conditions product
// CLUEOutputProducer reads (reference: src/serial/plugin-CLUEOutputProducer/CLUEOutputESProducer.cc).
// The driver passes the directory itself.
#include <filesystem>
#include <memory>

#include "Framework/ESPluginFactory.h"
#include "Framework/ESProducer.h"
#include "Framework/EventSetup.h"

class CLUEOutputESProducer : public edm::ESProducer {
public:
  explicit CLUEOutputESProducer(std::filesystem::path const& dir) : dir_{dir} {}
  void produce(edm::EventSetup& eventSetup) override {
    std::filesystem::create_directories(dir_);
    eventSetup.put(std::make_unique<std::filesystem::path>(dir_));
  }

private:
  std::filesystem::path dir_;
};

DEFINE_FWK_EVENTSETUP_MODULE(CLUEOutputESProducer);
