// CLUE2D on the HIP backend behind the serial backend's module interface: consumes PointsCloud,
// produces PointsCloudSerial (src/serial/plugin-CLUEClusterizer/CLUESerialClusterizer.cc:10-38),
// with bit-identical outputs. It is an external-work module (src/serial/Framework/EDProducer.h:30-50):
// acquire() packs the event into pinned memory and enqueues the copy in, the clustering and the copy
// out on the module's HIP stream, then returns; a stream callback hands the holder back when the
// device is done (the CUDA backend's ScopedContextAcquire pattern, src/cuda/CUDACore/
// ScopedContext.cc:94-109); produce() checks the device's error flags and fills the product.
#include <cstring>

#include "ClueContext.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/PointsCloud.h"
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

class CLUEHIPClusterizer : public edm::EDProducerExternalWork {
public:
  explicit CLUEHIPClusterizer(edm::ProductRegistry& reg)
      : pointsCloudToken_{reg.consumes<PointsCloud>()}, clusterToken_{reg.produces<PointsCloudSerial>()} {}

private:
  // staging columns: inputs x, y, layer, weight; outputs rho, delta, nearestHigher, isSeed, clusterIndex
  enum Column { kX, kY, kLayer, kWeight, kRho, kDelta, kNh, kSeed, kCluster, kColumns };

  void acquire(edm::Event const& event, edm::EventSetup const& eventSetup,
               edm::WaitingTaskWithArenaHolder holder) override {
    auto const& pc = event.get(pointsCloudToken_);
    const size_t n = pc.x.size();
    if (pc.y.size() != n || pc.layer.size() != n || pc.weight.size() != n)
      throw std::runtime_error("CLUEHIPClusterizer: PointsCloud columns differ in length");
    if (n == 0) return;
    Parameters const& par = eventSetup.get<Parameters>();
    clue_ctx* ctx = ctx_.get(event.streamID());
    staging_.reserve(ctx_, ctx, kColumns * ClueStaging::column_bytes(n));
    std::memcpy(staging_.host(kX, n), pc.x.data(), n * 4);
    std::memcpy(staging_.host(kY, n), pc.y.data(), n * 4);
    std::memcpy(staging_.host(kLayer, n), pc.layer.data(), n * 4);
    std::memcpy(staging_.host(kWeight, n), pc.weight.data(), n * 4);
    auto d = [&](Column c) { return staging_.device(c, n); };
    ctx_.check(clue_memcpy_htod(ctx, d(kX), staging_.host(kX, n), kRho * ClueStaging::column_bytes(n), nullptr),
               "CLUEHIPClusterizer: copy in");
    const clue2d_params p{par.dc, par.rhoc, par.outlierDeltaFactor};
    ctx_.check(clue2d_run_device(ctx, &p, static_cast<int32_t>(n), reinterpret_cast<float*>(d(kX)),
                                 reinterpret_cast<float*>(d(kY)), reinterpret_cast<int32_t*>(d(kLayer)),
                                 reinterpret_cast<float*>(d(kWeight)), reinterpret_cast<float*>(d(kRho)),
                                 reinterpret_cast<float*>(d(kDelta)), reinterpret_cast<int32_t*>(d(kNh)),
                                 reinterpret_cast<int32_t*>(d(kSeed)), reinterpret_cast<int32_t*>(d(kCluster)), nullptr),
               "CLUEHIPClusterizer");
    ctx_.check(clue_memcpy_dtoh(ctx, staging_.host(kRho, n), d(kRho), (kColumns - kRho) * ClueStaging::column_bytes(n),
                                nullptr),
               "CLUEHIPClusterizer: copy out");
    auto* h = new edm::WaitingTaskWithArenaHolder(std::move(holder));
    const int rc = clue_stream_callback(ctx, nullptr, &give_back, h);
    if (rc != CLUE_OK) {
      delete h;
      ctx_.check(rc, "CLUEHIPClusterizer: completion callback");
    }
  }

  void produce(edm::Event& event, edm::EventSetup const&) override {
    auto const& pc = event.get(pointsCloudToken_);
    const size_t n = pc.x.size();
    PointsCloudSerial out;
    out.x = pc.x;
    out.y = pc.y;
    out.layer = pc.layer;
    out.weight = pc.weight;
    out.outResize();
    if (n > 0) {
      ctx_.check(clue_check(ctx_.get(event.streamID()), nullptr), "CLUEHIPClusterizer");
      std::memcpy(out.rho.data(), staging_.host(kRho, n), n * 4);
      std::memcpy(out.delta.data(), staging_.host(kDelta, n), n * 4);
      std::memcpy(out.nearestHigher.data(), staging_.host(kNh, n), n * 4);
      std::memcpy(out.isSeed.data(), staging_.host(kSeed, n), n * 4);
      std::memcpy(out.clusterIndex.data(), staging_.host(kCluster, n), n * 4);
    }
    event.emplace(clusterToken_, std::move(out));
  }

  edm::EDGetTokenT<PointsCloud> pointsCloudToken_;
  edm::EDPutTokenT<PointsCloudSerial> clusterToken_;
  ClueContext ctx_;
  ClueStaging staging_;
};

DEFINE_FWK_MODULE(CLUEHIPClusterizer);
