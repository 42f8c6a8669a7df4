// CLUE3D (SoA path) launch sequence (device buffers and parameters).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tiles.h"

namespace clue {

// Kernel parameters of KernelCalculateDensitySoA / KernelComputeDistanceToHigherSoA /
// KernelFindAndAssignClustersSoA (src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:26-32,
// 286-290, 509-515). The reference's compiled-in defaults are listed in clue3d_default_params().
struct Clue3DParams {
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

struct Clue3DBuffers {
  int E = 0;
  int64_t N = 0;
  int nBlocks = 0;
  const BlockDesc* blocks = nullptr;
  const int* evOff = nullptr;
  uint32_t* k1 = nullptr;
  int* i1 = nullptr;
  uint32_t* k2 = nullptr;
  int* i2 = nullptr;
  float4* cre = nullptr;  // sorted {r_over_absz, phi, energy, z}
  float4* crp = nullptr;  // sorted {r_over_absz, phi, rho, z}
  int* hist = nullptr;
  int* part = nullptr;
  int* table = nullptr;
  int* tileStart = nullptr;
  int64_t tableStride = 0;
  SeedRank seeds{};
  int* err = nullptr;  // device error flags: err[event * errStride]
  int errStride = 0;
  int densityVariant = 0;  // kernel shapes (points per thread, staged points) of the density and
  int nhVariant = 0;       // nearest-higher steps; see clue3d_variant_shape
};

struct Clue3DIO {
  const float* z;
  const float* eta;
  const float* phi;
  const float* r_over_absz;
  const float* radius;
  const float* energy;
  const int32_t* layer;
  const int32_t* isSilicon;
  float* rho;
  float* delta_f;         // delta.first
  int32_t* delta_l;       // delta.second (|layer difference|)
  int32_t* nh_layer;      // nearestHigher.first
  int32_t* nh_idx;        // nearestHigher.second (event-local)
  int32_t* isSeed;
  int32_t* tracksterIndex;
};

constexpr int64_t clue3d_table_stride() { return ((int64_t)k3dKeys + 1 + 63) / 64 * 64; }

void clue3d_enqueue(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& p, hipStream_t s);
// Re-runs one step of an enqueued batch with a given kernel variant: which = 0 density, 1 nearest
// higher (and seed flags). Used by the plan autotuner, after a full clue3d_enqueue.
void clue3d_enqueue_step(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& p, int which, int variant,
                         hipStream_t s);
constexpr int kClue3DVariants = 4;
void clue3d_variant_shape(int variant, int& items, int& stage);
size_t clue3d_workspace_bytes(int64_t N, int nEvents, int nBlocks);
void clue3d_carve(void* base, int64_t N, int E, int nBlocks, Clue3DBuffers& b);

}  // namespace clue
