// CLUE3D event data (members of src/serial/DataFormats/ClusterCollection.h:6-53): input layer
// clusters and the serial backend's output product (delta = {transverse distance, |layer difference|},
// nearestHigher = {layer, index}); `followers` is left empty.
#ifndef CLUE_DF_ClusterCollection_h
#define CLUE_DF_ClusterCollection_h

#include <utility>
#include <vector>

struct ClusterCollection {
  std::vector<float> x;
  std::vector<float> y;
  std::vector<float> z;
  std::vector<float> eta;
  std::vector<float> phi;
  std::vector<float> r_over_absz;
  std::vector<float> radius;
  std::vector<int> layer;
  std::vector<float> energy;
  std::vector<int> isSilicon;
};

struct ClusterCollectionSerial {
  void outResize() {
    const auto n = x.size();
    rho.resize(n);
    delta.resize(n);
    nearestHigher.resize(n);
    tracksterIndex.resize(n);
    isSeed.resize(n);
  }

  std::vector<float> x;
  std::vector<float> y;
  std::vector<float> z;
  std::vector<float> eta;
  std::vector<float> phi;
  std::vector<float> r_over_absz;
  std::vector<float> radius;
  std::vector<int> layer;
  std::vector<float> energy;
  std::vector<int> isSilicon;

  std::vector<float> rho;
  std::vector<std::pair<float, int>> delta;
  std::vector<std::pair<int, int>> nearestHigher;
  std::vector<std::vector<std::pair<int, int>>> followers;
  std::vector<int> isSeed;
  std::vector<int> tracksterIndex;
};

#endif
