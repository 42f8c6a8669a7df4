// CLUE2D on gfx950: tile order (layer pass, per-layer bin sort with the bin table), density,
// nearest-higher (+ seed/outlier flags) and labels for a packed batch of events. A single event is
// the batch of one.
//
// Reference semantics: src/serial/plugin-CLUEClusterizer/CLUEAlgoKernels.h:7-153.
// Bit-exactness rests on three facts:
//  * the sorted array lists each (layer, xBin, yBin) bin in ascending point index, keyed x-major,
//    so the bins (xBin, yLo..yHi) of one search-box column are one contiguous run, and walking the
//    columns in ascending xBin visits candidates in exactly the serial order (xBin outer, yBin
//    inner, index ascending) -> identical fp32 summation order for rho and the same "first
//    minimum wins" order for delta;
//  * every expression keeps the reference's operands and order, without FMA contraction;
//  * sqrtf is correctly rounded (the library is built with correctly rounded fp32 sqrt/div).
#include "clue2d_kernels.h"

#include <float.h>

#include <algorithm>
#include <math.h>
#include <string.h>

namespace clue {

float sqrt_le_threshold(float d) {
  if (d != d) return d;                    // NaN: no distance passes
  if (d < 0.f) return -1.f;                // sqrt(v) >= +0 > d for every v >= +0
  if (d == INFINITY) return INFINITY;
  uint32_t lo = 0, hi = 0x7f800000u;        // bit patterns of +0 .. +inf, ordered like the floats
  while (lo < hi) {                         // find the last pattern whose sqrt is <= d
    const uint32_t mid = lo + (hi - lo + 1) / 2;
    float v;
    memcpy(&v, &mid, 4);
    if (sqrtf(v) <= d)
      lo = mid;
    else
      hi = mid - 1;
  }
  float t;
  memcpy(&t, &lo, 4);
  return t;
}

Clue2DParams make_clue2d_params(float dc, float rhoc, float odf) {
  Clue2DParams p;
  p.dc = dc;
  p.rhoc = rhoc;
  p.outlierDeltaFactor = odf;
  p.dc2 = sqrt_le_threshold(dc);
  p.dm2 = sqrt_le_threshold(odf * dc);
  p.dm2up = nextafterf(p.dm2, INFINITY);
  return p;
}

using namespace g2d;

namespace {

template <int V, class... A>
void density_v(bool indexPass, int nBlocks, hipStream_t s, A... a) {
  constexpr Variant2D v = kVariants2D[V];
  const int grid = CompShape<v.threads, v.items>::kPerSort * nBlocks;
  if (indexPass)
    k_density2d<v.threads, v.items, v.stage, false><<<grid, v.threads, 0, s>>>(a...);
  else
    k_density2d<v.threads, v.items, v.stage, true><<<grid, v.threads, 0, s>>>(a...);
}
template <int V, class... A>
void nh_v(bool indexPass, int nBlocks, hipStream_t s, A... a) {
  constexpr Variant2D v = kVariants2D[V];
  const int grid = CompShape<v.threads, v.items>::kPerSort * nBlocks;
  if (indexPass)
    k_nh2d<v.threads, v.items, v.stage, true><<<grid, v.threads, 0, s>>>(a...);
  else
    k_nh2d<v.threads, v.items, v.stage, false><<<grid, v.threads, 0, s>>>(a...);
}

}  // namespace

template <class... A>
void launch_density2d(int variant, bool indexPass, int nBlocks, hipStream_t s, A... a) {
  switch (variant) {
    case 1: density_v<1>(indexPass, nBlocks, s, a...); break;
    case 2: density_v<2>(indexPass, nBlocks, s, a...); break;
    case 3: density_v<3>(indexPass, nBlocks, s, a...); break;
    case 4: density_v<4>(indexPass, nBlocks, s, a...); break;
    case 5: density_v<5>(indexPass, nBlocks, s, a...); break;
    default: density_v<0>(indexPass, nBlocks, s, a...); break;
  }
}
template <class... A>
void launch_nh2d(int variant, bool indexPass, int nBlocks, hipStream_t s, A... a) {
  switch (variant) {
    case 1: nh_v<1>(indexPass, nBlocks, s, a...); break;
    case 2: nh_v<2>(indexPass, nBlocks, s, a...); break;
    case 3: nh_v<3>(indexPass, nBlocks, s, a...); break;
    case 4: nh_v<4>(indexPass, nBlocks, s, a...); break;
    case 5: nh_v<5>(indexPass, nBlocks, s, a...); break;
    default: nh_v<0>(indexPass, nBlocks, s, a...); break;
  }
}
static_assert(kNumVariants2D == 6, "launch_density2d / launch_nh2d dispatch six variants");

namespace {

// The neighbour steps and labels in one of the two output forms (Clue2DBuffers::nhIndexPass). The
// links and the index-order records of the index-pass form live in buffers the tile-order steps no
// longer need: `lg` (16 bytes per point: the records, or the direct form's links) and `scr` (4
// bytes per point: the tile-order links).
void density_step(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& prm, int variant, bool indexPass,
                  hipStream_t s) {
  const BinTables2D tabs{b.table16, b.table, b.segB, b.tableStride};
  launch_density2d(variant, indexPass, b.nBlocks, s, b.blocks, (const float4*)b.xyw, tabs, (const int4*)b.reachD,
                   prm.dc, prm.dc2, b.rhoKey, io.rho, b.lab);
}
void nh_step(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& prm, int variant, bool indexPass,
             hipStream_t s) {
  const BinTables2D tabs{b.table16, b.table, b.segB, b.tableStride};
  int* parent = indexPass ? b.scr : reinterpret_cast<int*>(b.lg);
  launch_nh2d(variant, indexPass, b.nBlocks, s, b.blocks, (const float4*)b.xyw, (const uint32_t*)b.rhoKey, tabs,
              (const int4*)b.reachN, b.evOff, prm.dc, prm.rhoc, prm.outlierDeltaFactor, prm.dm2, prm.dm2up, io.delta,
              io.nearestHigher, io.isSeed, parent, reinterpret_cast<int4*>(b.lg), b.seeds.word);
}
void labels_step(const Clue2DBuffers& b, const Clue2DIO& io, bool indexPass, hipStream_t s) {
  if (!indexPass) {
    launch_labels(b.blocks, b.nBlocks, reinterpret_cast<const int*>(b.lg), nullptr, b.N, b.evOff, io.clusterIndex,
                  b.seeds, b.part, s);
    return;
  }
  launch_seed_ranks(b.N, b.seeds, b.part, s);
  k_labels_tile2d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, b.scr, b.seeds, b.evOff, b.lab);
  k_outputs2d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, reinterpret_cast<const int4*>(b.lg), b.lab, io.rho, io.delta,
                                             io.nearestHigher, io.isSeed, io.clusterIndex);
}

}  // namespace

void clue2d_enqueue(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& prm, hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  const int nHist = kRadix * b.nBlocks;
  // tile order: a stable pass by layer into {x, y, w, index} records (its histogram kernel also
  // clears the seed words the nearest-higher kernel sets and the largest class's list count; its
  // scatter writes each event's layer starts and that list), then each (event, layer) sorted by bin
  // in its own workgroup, which also writes that layer's bin-table entries
  const LayerOf2D layerOf{io.layer, b.err, b.errStride};
  const Prologue2D pro{b.seeds.word, (int64_t)seed_tiles(b.N) * 32, b.N, b.segCnt};
  k_radix_hist<LayerOf2D, DigitLayer2D, Prologue2D><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, layerOf, b.hist, pro);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<LayerOf2D, DigitLayer2D, EmitSeg2D><<<b.nBlocks, kThreads, 0, s>>>(
      b.blocks, layerOf, b.hist, EmitSeg2D{io.x, io.y, io.weight, b.lg, b.segB, b.segList, b.segCnt});
  const int nSeg = b.E * k2dLayers;
  k_segsort2d<kSegThreadsS, 0, kSegSmall><<<nSeg, kSegThreadsS, 0, s>>>(
      b.evOff, b.segB, b.lg, b.scr, b.keyS, b.xyw, b.table16, b.table, b.tableStride, nullptr, nullptr);
  k_segsort2d<kSegThreadsL, kSegSmall, kSegCap><<<nSeg, kSegThreadsL, 0, s>>>(
      b.evOff, b.segB, b.lg, b.scr, b.keyS, b.xyw, b.table16, b.table, b.tableStride, nullptr, nullptr);
  // the largest class: as many workgroups as the event sizes allow such layers (the list holds them)
  const int gridH = b.segBoundH >= 0 ? b.segBoundH : nSeg;
  if (gridH > 0)
    k_segsort2d<kSegThreadsH, kSegCap, kSegNone><<<gridH, kSegThreadsH, 0, s>>>(
        b.evOff, b.segB, b.lg, b.scr, b.keyS, b.xyw, b.table16, b.table, b.tableStride, b.segList, b.segCnt);
  const BinTables2D tabs{b.table16, b.table, b.segB, b.tableStride};
  const int nTiles = b.nBlocks * kTilesPerSort;
  k_reach2d<<<(nTiles + kReachThreads - 1) / kReachThreads, kReachThreads, 0, s>>>(
      b.blocks, nTiles, b.keyS, tabs, prm.dc, prm.outlierDeltaFactor, b.reachD, b.reachN);
  const bool ip = b.nhIndexPass != 0;
  density_step(b, io, prm, b.densityVariant, ip, s);
  nh_step(b, io, prm, b.nhVariant, ip, s);
  labels_step(b, io, ip, s);
}

void clue2d_enqueue_step(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& prm, int which, int variant,
                         hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  const bool ip = b.nhIndexPass != 0;
  if (which == 0) {
    density_step(b, io, prm, variant, ip, s);
  } else if (which == 1) {
    nh_step(b, io, prm, variant, ip, s);
  } else {
    density_step(b, io, prm, b.densityVariant, variant != 0, s);
    nh_step(b, io, prm, b.nhVariant, variant != 0, s);
    labels_step(b, io, variant != 0, s);
  }
}

void clue2d_segment_bounds(Clue2DBuffers& b, const int64_t* offsets, int E) {
  int64_t g = 0;
  for (int e = 0; e < E; ++e) g += std::min<int64_t>(k2dLayers, (offsets[e + 1] - offsets[e]) / (kSegCap + 1));
  b.segBoundH = (int)g;
}

void clue2d_variant_shape(int variant, int& points, int& stage) {
  const Variant2D v = kVariants2D[variant >= 0 && variant < kNumVariants2D ? variant : 0];
  points = v.threads * v.items;
  stage = v.stage;
}

// Carve a workspace of clue2d_workspace_bytes(N, E, nBlocks) bytes into the buffers of a batch.
void clue2d_carve(void* base, int64_t N, int E, int nBlocks, Clue2DBuffers& b) {
  Carver cv{reinterpret_cast<char*>(base)};
  const size_t n = (size_t)(N > 0 ? N : 1);
  b.E = E;
  b.N = N;
  b.nBlocks = nBlocks;
  b.keyS = cv.take<uint32_t>(n);
  b.scr = cv.take<int>(n);
  b.lg = cv.take<float4>(n);
  b.segB = cv.take<int>((size_t)(E > 0 ? E : 1) * (k2dLayers + 1));
  b.segList = cv.take<int>((size_t)(E > 0 ? E : 1) * k2dLayers);
  b.segCnt = cv.take<int>(1);
  b.segBoundH = -1;
  b.xyw = cv.take<float4>(n);
  b.rhoKey = cv.take<uint32_t>(n);
  b.lab = cv.take<int2>(n);
  b.reachD = cv.take<int4>((size_t)(nBlocks > 0 ? nBlocks : 1) * kTilesPerSort);
  b.reachN = cv.take<int4>((size_t)(nBlocks > 0 ? nBlocks : 1) * kTilesPerSort);
  b.hist = cv.take<int>((size_t)kRadix * (nBlocks > 0 ? nBlocks : 1));
  const int64_t maxScan = (int64_t)kRadix * nBlocks > N ? (int64_t)kRadix * nBlocks : N;
  b.part = cv.take<int>((size_t)scan_partials(maxScan > 0 ? maxScan : 1));
  b.tableStride = clue2d_table_stride();
  b.table16 = cv.take<uint16_t>((size_t)b.tableStride * (E > 0 ? E : 1));
  b.table = cv.take<int>((size_t)b.tableStride * (E > 0 ? E : 1));
  b.evOff = cv.take<int>((size_t)E + 1);
  b.blocks = cv.take<BlockDesc>((size_t)(nBlocks > 0 ? nBlocks : 1));
  b.seeds = seed_rank_carve(cv, N);
}

size_t clue2d_workspace_bytes(int64_t N, int nEvents, int nBlocks) {
  const size_t n = (size_t)(N > 0 ? N : 1);
  size_t bytes = 0;
  auto add = [&](size_t sz) { bytes += (sz + 255) & ~(size_t)255; };
  add(n * 4);                                               // keyS
  add(n * 4);                                               // scr
  add(n * 16);                                              // lg
  add((size_t)(nEvents > 0 ? nEvents : 1) * (k2dLayers + 1) * 4);  // segB
  add((size_t)(nEvents > 0 ? nEvents : 1) * k2dLayers * 4);        // segList
  add(4);                                                           // segCnt
  add(n * 16);                                              // xyw
  add(n * 4);                                               // rhoKey
  add(n * 8);                                               // lab
  for (int k = 0; k < 2; ++k) add((size_t)(nBlocks > 0 ? nBlocks : 1) * kTilesPerSort * sizeof(int4));  // reach
  add((size_t)kRadix * (size_t)(nBlocks > 0 ? nBlocks : 1) * 4);  // hist
  const int64_t maxScan = (int64_t)kRadix * nBlocks > N ? (int64_t)kRadix * nBlocks : N;
  add((size_t)scan_partials(maxScan > 0 ? maxScan : 1) * 4);  // partials
  add((size_t)clue2d_table_stride() * (size_t)(nEvents > 0 ? nEvents : 1) * 2);  // table16
  add((size_t)clue2d_table_stride() * (size_t)(nEvents > 0 ? nEvents : 1) * 4);  // table
  add((size_t)(nEvents + 1) * 4);                                   // evOff
  add((size_t)(nBlocks > 0 ? nBlocks : 1) * sizeof(BlockDesc));     // blocks
  bytes += seed_rank_bytes(N);                                      // seed words
  add(256);                                                         // error flags
  return bytes;
}

}  // namespace clue
