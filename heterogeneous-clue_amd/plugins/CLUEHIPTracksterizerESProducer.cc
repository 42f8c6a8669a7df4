// CLUE3D kernel parameters from a text file of `name,value` lines (names of Parameters3D; missing
// names keep the reference's defaults). The reference compiles these in as default arguments of
// its kernels (src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:26-32, 286-290, 509-515).
#include <filesystem>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include "DataFormats/CLUE_config.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/ESProducer.h"
#include "Framework/EventSetup.h"

class CLUEHIPTracksterizerESProducer : public edm::ESProducer {
public:
  explicit CLUEHIPTracksterizerESProducer(std::filesystem::path const& config) : config_{config} {}

  void produce(edm::EventSetup& eventSetup) override {
    Parameters3D p;
    std::ifstream in(config_);
    if (!in) throw std::runtime_error("cannot open CLUE3D configuration " + config_.string());
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] == '#') continue;
      const auto comma = line.find(',');
      if (comma == std::string::npos) throw std::runtime_error("bad CLUE3D configuration line: " + line);
      const std::string k = line.substr(0, comma);
      const std::string v = line.substr(comma + 1);
      if (k == "densitySiblingLayers") p.densitySiblingLayers = std::stoi(v);
      else if (k == "densityXYDistanceSqr") p.densityXYDistanceSqr = std::stof(v);
      else if (k == "kernelDensityFactor") p.kernelDensityFactor = std::stof(v);
      else if (k == "densityOnSameLayer") p.densityOnSameLayer = std::stoi(v) != 0;
      else if (k == "nearestHigherOnSameLayer") p.nearestHigherOnSameLayer = std::stoi(v) != 0;
      else if (k == "criticalXYDistance") p.criticalXYDistance = std::stof(v);
      else if (k == "criticalZDistanceLyr") p.criticalZDistanceLyr = std::stof(v);
      else if (k == "criticalDensity") p.criticalDensity = std::stof(v);
      else if (k == "criticalSelfDensity") p.criticalSelfDensity = std::stof(v);
      else if (k == "outlierMultiplier") p.outlierMultiplier = std::stof(v);
      else throw std::runtime_error("unknown CLUE3D parameter " + k);
    }
    eventSetup.put(std::make_unique<Parameters3D>(p));
  }

private:
  std::filesystem::path config_;
};

DEFINE_FWK_EVENTSETUP_MODULE(CLUEHIPTracksterizerESProducer);
