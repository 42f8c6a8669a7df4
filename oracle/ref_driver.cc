// C entry points around the reference's own serial CLUE / CLUE3D kernels, compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`). TEST INFRASTRUCTURE:
// used to validate the C restatement, to write golden fixtures and as the CPU baseline.
//
// CLUE2D: CLUEAlgoSerial::makeClusters (src/serial/plugin-CLUEClusterizer/CLUEAlgoSerial.cc:18-35),
//         exactly as CLUESerialClusterizer::produce calls it (CLUESerialClusterizer.cc:29-36).
// CLUE3D: the SoA sequence of CLUE3DAlgoSerial::makeTrackstersSoA
//         (src/serial/plugin-CLUETracksterizer/CLUE3DAlgoSerial.cc:42-50), with the kernel
//         parameters passed explicitly instead of taking the defaults.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "DataFormats/LayerTilesSerial.h"
#include "DataFormats/Math/deltaPhi.h"
#include "DataFormats/PointsCloud.h"
#include "DataFormats/ClusterCollection.h"
#include "DataFormats/TICLLayerTile.h"
#include "plugin-CLUEClusterizer/CLUEAlgoSerial.h"
#include "plugin-CLUETracksterizer/CLUE3DAlgoKernels.h"

extern "C" {

struct ref_clue3d_params {
  int32_t densitySiblingLayers;
  float densityXYDistanceSqr;
  float kernelDensityFactor;
  int32_t densityOnSameLayer;
  int32_t nearestHigherOnSameLayer;
  float criticalXYDistance;
  float criticalZDistanceLyr;
  float criticalDensity;
  float criticalSelfDensity;
  float outlierMultiplier;
};

// One algorithm object per caller thread, as one module instance per EDM stream holds one.
static CLUEAlgoSerial& algo2d() {
  thread_local std::unique_ptr<CLUEAlgoSerial> a{new CLUEAlgoSerial()};
  return *a;
}
static TICLLayerTiles& tiles3d() {
  thread_local std::unique_ptr<TICLLayerTiles> t{new TICLLayerTiles()};
  return *t;
}

int ref_clue2d(int32_t n, const float* x, const float* y, const int32_t* layer, const float* w, float dc, float rhoc,
               float odf, float* rho, float* delta, int32_t* nh, int32_t* isSeed, int32_t* clusterIndex) {
  for (int32_t i = 0; i < n; ++i)
    if (layer[i] < 0 || layer[i] >= NLAYERS) return -1;
  PointsCloud pc;
  pc.x.assign(x, x + n);
  pc.y.assign(y, y + n);
  pc.layer.assign(layer, layer + n);
  pc.weight.assign(w, w + n);
  PointsCloudSerial out;
  algo2d().makeClusters(pc, out, dc, rhoc, odf);
  std::memcpy(rho, out.rho.data(), sizeof(float) * n);
  std::memcpy(delta, out.delta.data(), sizeof(float) * n);
  std::memcpy(nh, out.nearestHigher.data(), sizeof(int32_t) * n);
  std::memcpy(isSeed, out.isSeed.data(), sizeof(int32_t) * n);
  std::memcpy(clusterIndex, out.clusterIndex.data(), sizeof(int32_t) * n);
  return 0;
}

int ref_clue3d(int32_t n, const float* z, const float* eta, const float* phi, const float* roz, const float* radius,
               const float* energy, const int32_t* layer, const int32_t* isSi, const ref_clue3d_params* p, float* rho,
               float* delta_f, int32_t* delta_l, int32_t* nh_layer, int32_t* nh_idx, int32_t* isSeed,
               int32_t* tracksterIndex) {
  for (int32_t i = 0; i < n; ++i)
    if (layer[i] < 0 || layer[i] >= ticl::TileConstants::nLayers) return -1;
  // CLUE3DAlgoSerial::setupSoA (CLUE3DAlgoSerial.cc:7-19); x and y are not read by the kernels.
  ClusterCollectionSerial c;
  c.x.assign(n, 0.f);
  c.y.assign(n, 0.f);
  c.z.assign(z, z + n);
  c.eta.assign(eta, eta + n);
  c.phi.assign(phi, phi + n);
  c.r_over_absz.assign(roz, roz + n);
  c.radius.assign(radius, radius + n);
  c.layer.assign(layer, layer + n);
  c.energy.assign(energy, energy + n);
  c.isSilicon.assign(isSi, isSi + n);
  c.outResize();
  TICLLayerTiles& h = tiles3d();
  KernelComputeHistogramSoA(h, c);
  KernelCalculateDensitySoA(h, c, 0, p->densitySiblingLayers, p->densityXYDistanceSqr, p->kernelDensityFactor,
                            p->densityOnSameLayer != 0);
  KernelComputeDistanceToHigherSoA(h, c, 0, p->densitySiblingLayers, p->nearestHigherOnSameLayer != 0);
  KernelFindAndAssignClustersSoA(c, 0, p->criticalXYDistance, p->criticalZDistanceLyr, p->criticalDensity,
                                 p->criticalSelfDensity, p->outlierMultiplier);
  h.clear();
  for (int32_t i = 0; i < n; ++i) {
    rho[i] = c.rho[i];
    delta_f[i] = c.delta[i].first;
    delta_l[i] = c.delta[i].second;
    nh_layer[i] = c.nearestHigher[i].first;
    nh_idx[i] = c.nearestHigher[i].second;
    isSeed[i] = c.isSeed[i];
    tracksterIndex[i] = c.tracksterIndex[i];
  }
  return 0;
}

// The reference's own scalar binning helpers, evaluated on n values
// (LayerTilesSerial::getXBin/getYBin, LayerTilesSerial.h:24-40; TICLLayerTileT::etaBin/phiBin,
// TICLLayerTile.h:15-30; reco::reduceRange<float>, Math/deltaPhi.h:17-24).
int ref_binning(int32_t n, const float* v, int32_t* xbin, int32_t* ybin, int32_t* etaBin, int32_t* phiBin,
                float* reduced) {
  static LayerTilesSerial lt;
  static ticl::TICLLayerTile tt;
  for (int32_t i = 0; i < n; ++i) {
    xbin[i] = lt.getXBin(v[i]);
    ybin[i] = lt.getYBin(v[i]);
    etaBin[i] = tt.etaBin(v[i]);
    phiBin[i] = tt.phiBin(v[i]);
    reduced[i] = reco::reduceRange(v[i]);
  }
  return 0;
}

// Event-parallel CLUE2D over a packed batch (offsets[e]..offsets[e+1]) on `threads` std::threads,
// one algorithm object per thread, like EDM streams each owning a module instance.
int ref_clue2d_batch(int32_t nEvents, const int64_t* offsets, const float* x, const float* y, const int32_t* layer,
                     const float* w, float dc, float rhoc, float odf, float* rho, float* delta, int32_t* nh,
                     int32_t* isSeed, int32_t* clusterIndex, int32_t threads) {
  if (threads < 1) threads = 1;
  std::vector<int> rcs(nEvents, 0);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (int e = t; e < nEvents; e += threads) {
        const int64_t o = offsets[e];
        const int32_t m = static_cast<int32_t>(offsets[e + 1] - o);
        rcs[e] = ref_clue2d(m, x + o, y + o, layer + o, w + o, dc, rhoc, odf, rho + o, delta + o, nh + o, isSeed + o,
                            clusterIndex + o);
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int rc : rcs)
    if (rc) return rc;
  return 0;
}

}  // extern "C"
