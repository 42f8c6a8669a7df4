// One row of the CLUE2D result CSV, as the reference's writer prints it
// (src/serial/plugin-CLUEOutputProducer/CLUEOutputProducer.cc:36-42): default stream formatting,
// delta clamped at 999, no trailing newline. Shared by the CSV writer and the strict validator so
// that both print a point identically.
#ifndef CLUE_DF_CLUECsvRow_h
#define CLUE_DF_CLUECsvRow_h

#include <cstddef>
#include <ostream>

#include "DataFormats/PointsCloud.h"

inline constexpr char kClue2DCsvHeader[] = "index,x,y,layer,weight,rho,delta,nh,isSeed,clusterId";

inline void write_clue2d_row(std::ostream& out, PointsCloudSerial const& r, std::size_t i) {
  out << i << "," << r.x[i] << "," << r.y[i] << "," << r.layer[i] << "," << r.weight[i] << "," << r.rho[i] << ","
      << (r.delta[i] > 999 ? 999 : r.delta[i]) << "," << r.nearestHigher[i] << "," << r.isSeed[i] << ","
      << r.clusterIndex[i];
}

#endif
