// CLUE2D parameters (src/serial/DataFormats/CLUE_config.h:8-13) and the output file naming of the
// reference's CSV writer (:15-39): Event_<id>_<dc>_<rhoc>_<odf>.csv with two decimals.
#ifndef CLUE_DF_CLUE_config_h
#define CLUE_DF_CLUE_config_h

#include <iomanip>
#include <sstream>
#include <string>

struct Parameters {
  float dc = 20.;
  float rhoc = 25.;
  float outlierDeltaFactor = 2.;
  bool produceOutput = false;
};

// CLUE3D kernel parameters with the reference's compiled-in defaults
// (src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:26-32, 286-290, 509-515).
struct Parameters3D {
  int densitySiblingLayers = 3;
  float densityXYDistanceSqr = 3.24f;
  float kernelDensityFactor = 0.2f;
  bool densityOnSameLayer = false;
  bool nearestHigherOnSameLayer = false;
  float criticalXYDistance = 1.8f;
  float criticalZDistanceLyr = 5.f;
  float criticalDensity = 0.6f;
  float criticalSelfDensity = 0.15f;
  float outlierMultiplier = 2.f;
};

template <typename T>
std::string to_string_with_precision(const T value, const int n = 6) {
  std::ostringstream out;
  out.precision(n);
  out << std::fixed << value;
  return out.str();
}

inline std::string create_outputfileName(int eventId, float dc, float rhoc, float outlierDeltaFactor) {
  return "Event_" + to_string_with_precision(eventId, 0) + "_" + to_string_with_precision(dc, 2) + "_" +
         to_string_with_precision(rhoc, 2) + "_" + to_string_with_precision(outlierDeltaFactor, 2) + ".csv";
}

#endif
