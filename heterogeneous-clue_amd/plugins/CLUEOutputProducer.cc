// CSV writer of CLUE2D results in the reference's format (src/serial/plugin-CLUEOutputProducer/
// CLUEOutputProducer.cc:25-49): Event_<id>_<dc>_<rhoc>_<odf>.csv in the output directory, columns
// index,x,y,layer,weight,rho,delta,nh,isSeed,clusterId, delta clamped at 999.
#include <filesystem>
#include <fstream>
#include <iostream>

#include "DataFormats/CLUECsvRow.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/PointsCloud.h"
#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"

class CLUEOutputProducer : public edm::EDProducer {
public:
  explicit CLUEOutputProducer(edm::ProductRegistry& reg) : clustersToken_(reg.consumes<PointsCloudSerial>()) {}

private:
  void produce(edm::Event& event, edm::EventSetup const& eventSetup) override {
    auto const& results = event.get(clustersToken_);
    Parameters const& par = eventSetup.get<Parameters>();
    if (!par.produceOutput) return;
    auto const& outDir = eventSetup.get<std::filesystem::path>();
    const auto outFile = outDir / create_outputfileName(event.eventID(), par.dc, par.rhoc, par.outlierDeltaFactor);
    std::ofstream out(outFile);
    out << kClue2DCsvHeader << "\n";
    for (size_t i = 0; i < results.x.size(); ++i) {
      write_clue2d_row(out, results, i);
      out << "\n";
    }
    std::cout << "Ouput was saved in " << outFile << std::endl;
  }

  edm::EDGetTokenT<PointsCloudSerial> clustersToken_;
};

DEFINE_FWK_MODULE(CLUEOutputProducer);
