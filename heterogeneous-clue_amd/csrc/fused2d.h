// CLUE2D fused segment path: every (event, layer) of a batch is clustered by one workgroup with
// its points, bin table and intermediate results held in LDS.
//
// Layers never interact in CLUE2D (src/serial/plugin-CLUEClusterizer/CLUEAlgoKernels.h:28,72:
// the tiles are per layer), so a layer of an event is a closed unit of work. The path is
//   (1) a stable counting sort of the batch by (event, layer): one radix pass whose payload is the
//       point itself {x, y, weight, index}, which leaves every (event, layer) segment contiguous
//       and in ascending point index;
//   (2) a segment list, one size class per capacity (k_segments2d);
//   (3) one persistent kernel per class: each workgroup takes segments from a queue and, in LDS,
//       sorts the segment by (xBin, yBin, index), builds the bin table, computes rho, delta and
//       nearestHigher in the serial order, and walks the follower tree to its root (followers and
//       their nearest higher are always on the same layer);
//   (4) the seed numbering, which is per event and in point-index order across layers: an
//       exclusive scan of isSeed over the batch and one gather clusterIndex = rank(root).
// A layer with more than kFusedCap points does not fit; batches with such a layer run on the
// global-sort path (clue2d_enqueue), which has no size limit.
#pragma once

#include "clue2d.h"

namespace clue {

constexpr int kFusedCap = 10112;  // points of one (event, layer) segment that fit in 160 KB of LDS
constexpr int kFusedClasses = 3;  // segment size classes: <= 1024, <= 4096, <= kFusedCap points
constexpr int kErrCapacity = 4;   // device error bit: a segment above kFusedCap reached the fused path

struct Fused2DBuffers {
  int E = 0;
  int64_t N = 0;
  int nBlocks = 0;
  int cus = 256;                      // compute units of the device (persistent grid sizes)
  const BlockDesc* blocks = nullptr;  // device, the batch's sort blocks
  const int* evOff = nullptr;         // device, E + 1 int32 offsets
  const int* evBlk = nullptr;         // device, E + 1: first sort block of each event
  float4* lg = nullptr;               // points grouped by (event, layer): {x, y, weight, index bits}
  int* root = nullptr;                // per point: batch index of its seed, or -1
  int* rank = nullptr;                // exclusive scan of isSeed over the batch
  int* hist = nullptr;
  int* part = nullptr;
  int4* segs = nullptr;  // kFusedClasses lists of E * 100 entries {start, count, event base, 0}
  int* ctl = nullptr;    // [0, 3) class sizes, [4, 7) queue heads, [8] largest segment
  int* err = nullptr;
};

size_t fused2d_workspace_bytes(int64_t N, int E, int nBlocks);
void fused2d_carve(void* base, int64_t N, int E, int nBlocks, Fused2DBuffers& b);
// Host: first sort block of each event (E + 1 entries), for the segment list.
std::vector<int> make_event_blocks(const int64_t* offsets, int E);
// Steps (1) and (2); ctl[8] then holds the largest segment of the batch. strict: a segment above
// kFusedCap raises kErrCapacity (otherwise the caller reads ctl[8] and takes the global path).
void fused2d_enqueue_segments(const Fused2DBuffers& b, const Clue2DIO& io, bool strict, hipStream_t s);
// Steps (3) and (4).
void fused2d_enqueue_cluster(const Fused2DBuffers& b, const Clue2DIO& io, const Clue2DParams& p, hipStream_t s);

}  // namespace clue
