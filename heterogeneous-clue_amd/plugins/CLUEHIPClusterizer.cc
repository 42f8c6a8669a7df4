// CLUE2D on the HIP backend behind the serial backend's module interface: consumes PointsCloud,
// produces PointsCloudSerial (src/serial/plugin-CLUEClusterizer/CLUESerialClusterizer.cc:10-38),
// with bit-identical outputs.
#include "ClueContext.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/PointsCloud.h"
#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"

class CLUEHIPClusterizer : public edm::EDProducer {
public:
  explicit CLUEHIPClusterizer(edm::ProductRegistry& reg)
      : pointsCloudToken_{reg.consumes<PointsCloud>()}, clusterToken_{reg.produces<PointsCloudSerial>()} {}

private:
  void produce(edm::Event& event, edm::EventSetup const& eventSetup) override {
    auto const& pc = event.get(pointsCloudToken_);
    Parameters const& par = eventSetup.get<Parameters>();
    const auto n = pc.x.size();
    if (pc.y.size() != n || pc.layer.size() != n || pc.weight.size() != n)
      throw std::runtime_error("CLUEHIPClusterizer: PointsCloud columns differ in length");
    PointsCloudSerial out;
    out.x = pc.x;
    out.y = pc.y;
    out.layer = pc.layer;
    out.weight = pc.weight;
    out.outResize();
    clue_ctx* ctx = ctx_.get(event.streamID());
    const clue2d_params p{par.dc, par.rhoc, par.outlierDeltaFactor};
    ctx_.check(clue2d_run_host(ctx, &p, static_cast<int32_t>(n), pc.x.data(), pc.y.data(), pc.layer.data(),
                               pc.weight.data(), out.rho.data(), out.delta.data(), out.nearestHigher.data(),
                               out.isSeed.data(), out.clusterIndex.data()),
               "CLUEHIPClusterizer");
    event.emplace(clusterToken_, std::move(out));
  }

  edm::EDGetTokenT<PointsCloud> pointsCloudToken_;
  edm::EDPutTokenT<PointsCloudSerial> clusterToken_;
  ClueContext ctx_;
};

DEFINE_FWK_MODULE(CLUEHIPClusterizer);
