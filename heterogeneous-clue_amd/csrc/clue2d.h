// CLUE2D launch sequence (device buffers and parameters).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tiles.h"

namespace clue {

struct Clue2DParams {
  float dc;                  // critical distance (cm)
  float rhoc;                // critical density
  float outlierDeltaFactor;  // dm = outlierDeltaFactor * dc
  // Squared-distance thresholds: sqrtf(v) <= dc  <=>  v <= dc2 (and the same for dm), exactly,
  // because a correctly rounded sqrt is monotone; see sqrt_le_threshold().
  float dc2;
  float dm2;
  float dm2up;  // the float after dm2 (the nearest-higher scan's initial running minimum)
};

// Largest float v >= 0 with sqrtf(v) <= d (sqrtf correctly rounded, hence non-decreasing), so that
// for every v >= +0 or NaN: (sqrtf(v) <= d) == (v <= sqrt_le_threshold(d)).
float sqrt_le_threshold(float d);
Clue2DParams make_clue2d_params(float dc, float rhoc, float outlierDeltaFactor);

// Device workspace of one batch plan. Sizes: N points, E events, nBlocks sort blocks.
struct Clue2DBuffers {
  int E = 0;
  int64_t N = 0;
  int nBlocks = 0;
  const BlockDesc* blocks = nullptr;  // device
  const int* evOff = nullptr;         // device, E + 1 (int32 offsets)
  uint32_t* keyS = nullptr;  // tile order: bin keys (written at each 256-point tile's first and last position)
  int* scr = nullptr;        // slots of layer segments too large for LDS (k_segsort2d); later the
                             // tile-order parent links of the index-pass output form
  float4* lg = nullptr;      // layer order: {x, y, weight, point index as int bits}; later the parent
                             // array, or the index-order records of the index-pass output form
  int* segB = nullptr;       // per event, the first position of each layer (E * (k2dLayers + 1))
  int* segList = nullptr;    // the layers of the largest bin-sort class (up to E * k2dLayers)
  int* segCnt = nullptr;     // their count
  // Workgroups launched for the largest class: at most that many layers can be in it (from the
  // event sizes, clue2d_segment_bounds), or -1 for E * k2dLayers.
  int segBoundH = -1;
  float4* xyw = nullptr;   // tile order: {x, y, weight, point index as int bits}
  uint32_t* rhoKey = nullptr;  // tile order: rho as an order-preserving unsigned key (rho_key)
  int2* lab = nullptr;         // tile order, index-pass output form: {label, rho bits}
  int4* reachD = nullptr;  // per 256-point tile: staging / table reach of the density boxes
  int4* reachN = nullptr;  // the same for the nearest-higher boxes
  int* hist = nullptr;
  int* part = nullptr;
  uint16_t* table16 = nullptr;  // per event: bin-table entries as offsets from their layer's start
  int* table = nullptr;         // per event: the entries of layers too large for 16-bit offsets
  int64_t tableStride = 0;
  SeedRank seeds{};
  int* err = nullptr;  // device error flags: err[event * errStride]
  int errStride = 0;
  int densityVariant = 0;  // (points per thread, staging) variant of the density kernel (kVariants2D)
  int nhVariant = 0;       // the same for the nearest-higher kernel
  // Output form of the neighbour steps: 0 writes rho (density), delta, nearestHigher, isSeed and the
  // links (nearest higher) at each point's index; 1 keeps rho and the links in tile order and writes
  // one index-order record per point (k_nh2d's INDEX_PASS), and the labels step writes all five
  // outputs in index order.
  int nhIndexPass = 0;
};

// Caller-owned device inputs/outputs, packed over the batch (nearestHigher and clusterIndex are
// event-local indices, as each event's PointsCloudSerial holds them).
struct Clue2DIO {
  const float* x;
  const float* y;
  const int32_t* layer;
  const float* weight;
  float* rho;
  float* delta;
  int32_t* nearestHigher;
  int32_t* isSeed;
  int32_t* clusterIndex;
};

constexpr int64_t clue2d_table_stride() { return ((int64_t)k2dKeys + 1 + 63) / 64 * 64; }

void clue2d_enqueue(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& p, hipStream_t s);
// The density (which = 0) or nearest-higher (which = 1) kernel alone with the given variant, or the
// density, nearest-higher and labels steps (variants b.densityVariant, b.nhVariant) in output form
// `variant` (which = 2, see nhIndexPass), on a batch clue2d_enqueue has run on: each reads only
// what earlier steps left and rewrites its outputs with the same values (the seed bits are OR-ed in
// again), so the autotuner can time them in place.
void clue2d_enqueue_step(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& p, int which, int variant,
                         hipStream_t s);
constexpr int kClue2DVariants = 6;
// Points per block and staged points of a variant.
void clue2d_variant_shape(int variant, int& points, int& stage);
size_t clue2d_workspace_bytes(int64_t N, int nEvents, int nBlocks);
// Tight launch size of the largest bin-sort class for a batch with these offsets (E + 1 entries):
// an event of n points has at most n / 10241 layers above 10240 points. Call after clue2d_carve
// (which resets it to the whole grid).
void clue2d_segment_bounds(Clue2DBuffers& b, const int64_t* offsets, int E);
void clue2d_carve(void* base, int64_t N, int E, int nBlocks, Clue2DBuffers& b);

}  // namespace clue
