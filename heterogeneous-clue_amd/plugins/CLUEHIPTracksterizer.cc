// CLUE3D on the HIP backend behind the serial tracksterizer's module interface: consumes
// ClusterCollection, produces ClusterCollectionSerial (src/serial/plugin-CLUETracksterizer/
// CLUESerialTracksterizer.cc:10-43, SoA path). Kernel parameters come from Parameters3D in the
// EventSetup when a module put one there, otherwise the reference's compiled-in defaults are used.
#include "ClueContext.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/ClusterCollection.h"
#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"

class CLUEHIPTracksterizer : public edm::EDProducer {
public:
  explicit CLUEHIPTracksterizer(edm::ProductRegistry& reg)
      : clusterCollectionToken_{reg.consumes<ClusterCollection>()},
        tracksterToken_{reg.produces<ClusterCollectionSerial>()} {}

private:
  void produce(edm::Event& event, edm::EventSetup const& eventSetup) override {
    auto const& pc = event.get(clusterCollectionToken_);
    Parameters3D par;
    try {
      par = eventSetup.get<Parameters3D>();
    } catch (std::runtime_error const&) {
      // no CLUE3D conditions module configured: the reference's defaults
    }
    const auto n = pc.layer.size();
    ClusterCollectionSerial out;
    out.x = pc.x;
    out.y = pc.y;
    out.z = pc.z;
    out.eta = pc.eta;
    out.phi = pc.phi;
    out.r_over_absz = pc.r_over_absz;
    out.radius = pc.radius;
    out.layer = pc.layer;
    out.energy = pc.energy;
    out.isSilicon = pc.isSilicon;
    out.outResize();
    std::vector<float> df(n);
    std::vector<int32_t> dl(n), nhl(n), nhi(n);
    clue3d_params p;
    p.densitySiblingLayers = par.densitySiblingLayers;
    p.densityXYDistanceSqr = par.densityXYDistanceSqr;
    p.kernelDensityFactor = par.kernelDensityFactor;
    p.densityOnSameLayer = par.densityOnSameLayer;
    p.nearestHigherOnSameLayer = par.nearestHigherOnSameLayer;
    p.criticalXYDistance = par.criticalXYDistance;
    p.criticalZDistanceLyr = par.criticalZDistanceLyr;
    p.criticalDensity = par.criticalDensity;
    p.criticalSelfDensity = par.criticalSelfDensity;
    p.outlierMultiplier = par.outlierMultiplier;
    clue_ctx* ctx = ctx_.get(event.streamID());
    ctx_.check(clue3d_run_host(ctx, &p, static_cast<int32_t>(n), pc.z.data(), pc.eta.data(), pc.phi.data(),
                               pc.r_over_absz.data(), pc.radius.data(), pc.energy.data(), pc.layer.data(),
                               pc.isSilicon.data(), out.rho.data(), df.data(), dl.data(), nhl.data(), nhi.data(),
                               out.isSeed.data(), out.tracksterIndex.data()),
               "CLUEHIPTracksterizer");
    for (size_t i = 0; i < n; ++i) {
      out.delta[i] = {df[i], dl[i]};
      out.nearestHigher[i] = {nhl[i], nhi[i]};
    }
    event.emplace(tracksterToken_, std::move(out));
  }

  edm::EDGetTokenT<ClusterCollection> clusterCollectionToken_;
  edm::EDPutTokenT<ClusterCollectionSerial> tracksterToken_;
  ClueContext ctx_;
};

DEFINE_FWK_MODULE(CLUEHIPTracksterizer);
