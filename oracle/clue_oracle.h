/*
 * CPU restatement of the reference serial CLUE (2D) and CLUE3D (SoA) clusterizers.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is compared
 * against; it is loaded by tests/, by the smoke check and by the CPU-baseline leg of the
 * benchmark, never by the product path.
 *
 * Semantics follow (paths relative to the reference repository):
 *   src/serial/plugin-CLUEClusterizer/CLUEAlgoKernels.h:7-153      (CLUE2D)
 *   src/serial/DataFormats/LayerTilesSerial.h:22-52, LayerTilesConstants.h:7-27
 *   src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:19-24,26-150,286-392,509-580 (CLUE3D)
 *   src/serial/DataFormats/TICLLayerTile.h:15-54, Common.h:9-16, Math/deltaPhi.h:17-32
 * compiled with -ffp-contract=off (no FMA contraction), IEEE fp32, correctly rounded sqrt/div.
 */
#ifndef CLUE_ORACLE_H
#define CLUE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar helpers (exported so tests can pin them against closed forms) ---- */
int oracle_f2i(float v);                 /* x86 cvttss2si semantics: INT_MIN when out of range/NaN */
int oracle_xbin(float x);                /* LayerTilesSerial::getXBin (LayerTilesSerial.h:24-31) */
int oracle_ybin(float y);                /* LayerTilesSerial::getYBin (LayerTilesSerial.h:33-40) */
float oracle_reduce_range(float x);      /* reco::reduceRange<float> (Math/deltaPhi.h:17-24) */
int oracle_eta_bin(float eta);           /* TICLLayerTileT::etaBin (TICLLayerTile.h:15-22) */
int oracle_phi_bin(float phi);           /* TICLLayerTileT::phiBin (TICLLayerTile.h:24-30) */

/* ---- CLUE2D ---- returns 0, or -1 for a layer outside [0,100) */
int oracle_clue2d(int32_t n, const float* x, const float* y, const int32_t* layer, const float* weight,
                  float dc, float rhoc, float outlierDeltaFactor,
                  float* rho, float* delta, int32_t* nearestHigher, int32_t* isSeed, int32_t* clusterIndex);

/* ---- CLUE3D ---- */
typedef struct {
  int32_t densitySiblingLayers;   /* reference default 3 */
  float densityXYDistanceSqr;     /* 3.24 */
  float kernelDensityFactor;      /* 0.2 */
  int32_t densityOnSameLayer;     /* false */
  int32_t nearestHigherOnSameLayer; /* false */
  float criticalXYDistance;       /* 1.8 cm */
  float criticalZDistanceLyr;     /* 5 */
  float criticalDensity;          /* 0.6 GeV */
  float criticalSelfDensity;      /* 0.15 */
  float outlierMultiplier;        /* 2 */
} oracle_clue3d_params;

/* returns 0, or -1 for a layer outside [0,94) */
int oracle_clue3d(int32_t n, const float* z, const float* eta, const float* phi, const float* r_over_absz,
                  const float* radius, const float* energy, const int32_t* layer, const int32_t* isSilicon,
                  const oracle_clue3d_params* p,
                  float* rho, float* delta_f, int32_t* delta_l, int32_t* nh_layer, int32_t* nh_idx,
                  int32_t* isSeed, int32_t* tracksterIndex);

#ifdef __cplusplus
}
#endif
#endif
