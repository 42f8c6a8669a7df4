// Bit-exact dumps of the clusterizer outputs, one file per event (Event_<id>.bin in the directory
// given to CLUEDumpESProducer): uint32 n, then the output columns as raw little-endian arrays.
//   CLUE2D: rho f32, delta f32, nearestHigher i32, isSeed i32, clusterIndex i32
//   CLUE3D: rho f32, delta.first f32, delta.second i32, nearestHigher.first i32,
//           nearestHigher.second i32, isSeed i32, tracksterIndex i32
// No reference counterpart: the reference's CSV prints rounded floats; these dumps let tests
// compare every bit.
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "DataFormats/ClusterCollection.h"
#include "DataFormats/PointsCloud.h"
#include "Framework/EDProducer.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/ESProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"

struct DumpDirectory {
  std::filesystem::path dir;
};

class CLUEDumpESProducer : public edm::ESProducer {
public:
  explicit CLUEDumpESProducer(std::filesystem::path const& dir) : dir_{dir} {}
  void produce(edm::EventSetup& eventSetup) override {
    std::filesystem::create_directories(dir_);
    eventSetup.put(std::make_unique<DumpDirectory>(DumpDirectory{dir_}));
  }

private:
  std::filesystem::path dir_;
};
DEFINE_FWK_EVENTSETUP_MODULE(CLUEDumpESProducer);

namespace {
  template <typename T>
  void put(std::ofstream& out, std::vector<T> const& v) {
    out.write(reinterpret_cast<char const*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
  }
  std::ofstream open_dump(edm::EventSetup const& es, int eventId, uint32_t n) {
    auto const& d = es.get<DumpDirectory>();
    std::ofstream out(d.dir / ("Event_" + std::to_string(eventId) + ".bin"), std::ios::binary);
    out.write(reinterpret_cast<char const*>(&n), sizeof(n));
    return out;
  }
}  // namespace

class CLUEBinaryDump : public edm::EDProducer {
public:
  explicit CLUEBinaryDump(edm::ProductRegistry& reg) : token_(reg.consumes<PointsCloudSerial>()) {}

private:
  void produce(edm::Event& event, edm::EventSetup const& es) override {
    auto const& r = event.get(token_);
    auto out = open_dump(es, event.eventID(), static_cast<uint32_t>(r.rho.size()));
    put(out, r.rho);
    put(out, r.delta);
    put(out, r.nearestHigher);
    put(out, r.isSeed);
    put(out, r.clusterIndex);
  }
  edm::EDGetTokenT<PointsCloudSerial> token_;
};
DEFINE_FWK_MODULE(CLUEBinaryDump);

class CLUE3DBinaryDump : public edm::EDProducer {
public:
  explicit CLUE3DBinaryDump(edm::ProductRegistry& reg) : token_(reg.consumes<ClusterCollectionSerial>()) {}

private:
  void produce(edm::Event& event, edm::EventSetup const& es) override {
    auto const& r = event.get(token_);
    const size_t n = r.rho.size();
    std::vector<float> df(n);
    std::vector<int32_t> dl(n), nhl(n), nhi(n);
    for (size_t i = 0; i < n; ++i) {
      df[i] = r.delta[i].first;
      dl[i] = r.delta[i].second;
      nhl[i] = r.nearestHigher[i].first;
      nhi[i] = r.nearestHigher[i].second;
    }
    auto out = open_dump(es, event.eventID(), static_cast<uint32_t>(n));
    put(out, r.rho);
    put(out, df);
    put(out, dl);
    put(out, nhl);
    put(out, nhi);
    put(out, r.isSeed);
    put(out, r.tracksterIndex);
  }
  edm::EDGetTokenT<ClusterCollectionSerial> token_;
};
DEFINE_FWK_MODULE(CLUE3DBinaryDump);
