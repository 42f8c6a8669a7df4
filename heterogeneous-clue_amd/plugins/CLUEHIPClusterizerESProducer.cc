// Reads the CLUE2D parameter CSV `dc,rhoc,outlierDeltaFactor,produceOutput` (one line; the last
// line read wins) into Parameters, as CLUESerialClusterizerESProducer does
// (src/serial/plugin-CLUEClusterizer/CLUESerialClusterizerESProducer.cc:17-34).
#include <filesystem>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>

#include "DataFormats/CLUE_config.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/ESProducer.h"
#include "Framework/EventSetup.h"

class CLUEHIPClusterizerESProducer : public edm::ESProducer {
public:
  explicit CLUEHIPClusterizerESProducer(std::filesystem::path const& config) : config_{config} {}

  void produce(edm::EventSetup& eventSetup) override {
    Parameters par;
    std::ifstream in(config_);
    if (!in) throw std::runtime_error("cannot open CLUE configuration " + config_.string());
    std::string v;
    while (std::getline(in, v, ',')) {
      par.dc = std::stof(v);
      std::getline(in, v, ',');
      par.rhoc = std::stof(v);
      std::getline(in, v, ',');
      par.outlierDeltaFactor = std::stof(v);
      std::getline(in, v);
      par.produceOutput = static_cast<bool>(std::stoi(v));
    }
    eventSetup.put(std::make_unique<Parameters>(par));
  }

private:
  std::filesystem::path config_;
};

DEFINE_FWK_EVENTSETUP_MODULE(CLUEHIPClusterizerESProducer);
