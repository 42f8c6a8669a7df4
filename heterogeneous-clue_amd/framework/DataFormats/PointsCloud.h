// CLUE2D event data (members of src/serial/DataFormats/PointsCloud.h:6-43): the input SoA and the
// serial backend's output product. The HIP backend fills the same output product, so downstream
// modules (output writers, validators) see no difference. `followers` is left empty: the device
// assignment does not build follower lists.
#ifndef CLUE_DF_PointsCloud_h
#define CLUE_DF_PointsCloud_h

#include <vector>

struct PointsCloud {
  std::vector<float> x;
  std::vector<float> y;
  std::vector<int> layer;
  std::vector<float> weight;
};

struct PointsCloudSerial {
  void outResize() {
    const auto n = x.size();
    rho.resize(n);
    delta.resize(n);
    nearestHigher.resize(n);
    clusterIndex.resize(n);
    isSeed.resize(n);
  }

  std::vector<float> x;
  std::vector<float> y;
  std::vector<int> layer;
  std::vector<float> weight;

  std::vector<float> rho;
  std::vector<float> delta;
  std::vector<int> nearestHigher;
  std::vector<std::vector<int>> followers;
  std::vector<int> isSeed;
  std::vector<int> clusterIndex;
};

#endif
