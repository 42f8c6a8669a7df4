// CLUE3D on the HIP backend behind the serial tracksterizer's module interface: consumes
// ClusterCollection, produces ClusterCollectionSerial (src/serial/plugin-CLUETracksterizer/
// CLUESerialTracksterizer.cc:10-43, SoA path). Kernel parameters come from Parameters3D in the
// EventSetup when a module put one there, otherwise the reference's compiled-in defaults are used.
// Like CLUEHIPClusterizer it is an external-work module: acquire() stages the event in pinned memory
// and enqueues copy in, clustering and copy out on the module's HIP stream; a stream callback hands
// the holder back; produce() checks the device's error flags and fills the product.
#include <cstring>

#include "ClueContext.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/ClusterCollection.h"
#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"

namespace {
  void give_back(void* holder) {
    auto* h = static_cast<edm::WaitingTaskWithArenaHolder*>(holder);
    h->doneWaiting(nullptr);
    delete h;
  }
}  // namespace

class CLUEHIPTracksterizer : public edm::EDProducerExternalWork {
public:
  explicit CLUEHIPTracksterizer(edm::ProductRegistry& reg)
      : clusterCollectionToken_{reg.consumes<ClusterCollection>()},
        tracksterToken_{reg.produces<ClusterCollectionSerial>()} {}

private:
  // staging columns: 8 inputs, then 7 outputs
  enum Column { kZ, kEta, kPhi, kRoz, kRadius, kEnergy, kLayer, kSilicon,
                kRho, kDeltaF, kDeltaL, kNhLayer, kNhIdx, kSeed, kTrackster, kColumns };

  void acquire(edm::Event const& event, edm::EventSetup const& eventSetup,
               edm::WaitingTaskWithArenaHolder holder) override {
    auto const& pc = event.get(clusterCollectionToken_);
    const size_t n = pc.layer.size();
    for (size_t len : {pc.z.size(), pc.eta.size(), pc.phi.size(), pc.r_over_absz.size(), pc.radius.size(),
                       pc.energy.size(), pc.isSilicon.size()})
      if (len != n) throw std::runtime_error("CLUEHIPTracksterizer: ClusterCollection columns differ in length");
    if (n == 0) return;
    Parameters3D par;
    try {
      par = eventSetup.get<Parameters3D>();
    } catch (std::runtime_error const&) {
      // no CLUE3D conditions module configured: the reference's defaults (EventSetup::get is the
      // reference framework's only accessor, so the module builds there unchanged)
    }
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
    const size_t col = ClueStaging::column_bytes(n);
    staging_.reserve(ctx_, ctx, kColumns * col);
    const void* in[] = {pc.z.data(), pc.eta.data(), pc.phi.data(), pc.r_over_absz.data(),
                        pc.radius.data(), pc.energy.data(), pc.layer.data(), pc.isSilicon.data()};
    for (int c = kZ; c <= kSilicon; ++c) std::memcpy(staging_.host(c, n), in[c], n * 4);
    auto f = [&](Column c) { return reinterpret_cast<float*>(staging_.device(c, n)); };
    auto i = [&](Column c) { return reinterpret_cast<int32_t*>(staging_.device(c, n)); };
    ctx_.check(clue_memcpy_htod(ctx, staging_.device(kZ, n), staging_.host(kZ, n), kRho * col, nullptr),
               "CLUEHIPTracksterizer: copy in");
    ctx_.check(clue3d_run_device(ctx, &p, static_cast<int32_t>(n), f(kZ), f(kEta), f(kPhi), f(kRoz), f(kRadius),
                                 f(kEnergy), i(kLayer), i(kSilicon), f(kRho), f(kDeltaF), i(kDeltaL), i(kNhLayer),
                                 i(kNhIdx), i(kSeed), i(kTrackster), nullptr),
               "CLUEHIPTracksterizer");
    ctx_.check(clue_memcpy_dtoh(ctx, staging_.host(kRho, n), staging_.device(kRho, n), (kColumns - kRho) * col, nullptr),
               "CLUEHIPTracksterizer: copy out");
    auto* h = new edm::WaitingTaskWithArenaHolder(std::move(holder));
    const int rc = clue_stream_callback(ctx, nullptr, &give_back, h);
    if (rc != CLUE_OK) {
      delete h;
      ctx_.check(rc, "CLUEHIPTracksterizer: completion callback");
    }
  }

  void produce(edm::Event& event, edm::EventSetup const&) override {
    auto const& pc = event.get(clusterCollectionToken_);
    const size_t n = pc.layer.size();
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
    if (n > 0) {
      ctx_.check(clue_check(ctx_.get(event.streamID()), nullptr), "CLUEHIPTracksterizer");
      auto col = [&](Column c) { return staging_.host(c, n); };
      std::memcpy(out.rho.data(), col(kRho), n * 4);
      std::memcpy(out.isSeed.data(), col(kSeed), n * 4);
      std::memcpy(out.tracksterIndex.data(), col(kTrackster), n * 4);
      auto const* df = reinterpret_cast<float const*>(col(kDeltaF));
      auto const* dl = reinterpret_cast<int32_t const*>(col(kDeltaL));
      auto const* nhl = reinterpret_cast<int32_t const*>(col(kNhLayer));
      auto const* nhi = reinterpret_cast<int32_t const*>(col(kNhIdx));
      for (size_t k = 0; k < n; ++k) {
        out.delta[k] = {df[k], dl[k]};
        out.nearestHigher[k] = {nhl[k], nhi[k]};
      }
    }
    event.emplace(tracksterToken_, std::move(out));
  }

  edm::EDGetTokenT<ClusterCollection> clusterCollectionToken_;
  edm::EDPutTokenT<ClusterCollectionSerial> tracksterToken_;
  ClueContext ctx_;
  ClueStaging staging_;
};

DEFINE_FWK_MODULE(CLUEHIPTracksterizer);
