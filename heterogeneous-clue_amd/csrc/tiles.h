// Tile construction shared by CLUE2D and CLUE3D.
//
// The reference fills per-bin vectors in ascending point index (LayerTilesSerial.h:22,
// TICLLayerTile.h:15-17). Here the tiles of a whole batch of events are one array of point
// indices sorted by (event, bin key) and, inside a bin, by index: a stable LSD counting sort
// with one pass per bin coordinate (digits < 128). Each pass is
//   (1) a per-block digit histogram, written event-major as hist[event][digit][block],
//   (2) one exclusive scan of that array, which turns every (event, digit, block) count into
//       the global output position of that block's first item of that digit,
//   (3) a scatter in which each block ranks its items stably (wave ballot match + per-wave
//       counts in LDS) and writes them at base + rank.
// Blocks never straddle an event, so every event keeps its own range of the sorted array.
//
// A dense table then gives, per event, the sorted position where each bin key starts
// (the searches read contiguous runs of bins through it).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "clue_device.h"

namespace clue {

constexpr int kSortItems = 8;
constexpr int kSortBlock = kThreads * kSortItems;  // points per sort block
constexpr int kRadix = 128;                         // histogram stride per block (digits < 128)
constexpr int kTableTile = 4096;                    // bins per table-build block

struct BlockDesc {
  int32_t event;     // event id within the batch
  int32_t begin;     // first point (global position) of the block
  int32_t end;       // one past the last point
  int32_t histBase;  // kRadix * (first block of the event)
  int32_t nbEvent;   // number of blocks of the event
  int32_t bLocal;    // block index within the event
};

// Host: one BlockDesc per kSortBlock points of each event (none for an empty event).
std::vector<BlockDesc> make_blocks(const int64_t* offsets, int nEvents, int blockSize = kSortBlock);

// ---------------------------------------------------------------- device-wide exclusive scan
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;
inline int scan_partials(int64_t n) { return (int)((n + kScanTile - 1) / kScanTile); }
// Exclusive scan of n int32 (in may equal out). `part` needs scan_partials(n) ints.
void scan_exclusive(const int* in, int* out, int64_t n, int* part, hipStream_t s);

// ---------------------------------------------------------------- radix pass kernels

template <int DIV, int MOD>
struct Digit {
  __device__ __forceinline__ static int get(uint32_t k) { return (int)((k / (uint32_t)DIV) % (uint32_t)MOD); }
};

// Reads the key of input position i (and may record it as a side effect).
struct KeyArray {
  const uint32_t* key;
  __device__ __forceinline__ uint32_t operator()(int i) const { return key[i]; }
};

template <class Src, class D>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const BlockDesc* __restrict__ blocks, Src src,
                                                         int* __restrict__ hist) {
  __shared__ int s_h[kRadix];
  const BlockDesc bd = blocks[blockIdx.x];
  if (threadIdx.x < kRadix) s_h[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    const bool valid = i < bd.end;
    const int d = valid ? D::get(src(i)) : kRadix - 1;
    const uint64_t m = match_digit7(d);
    const int before = __popcll(m & ((1ull << lane_id()) - 1ull));
    if (valid && before == 0) atomicAdd(&s_h[d], (int)__popcll(m));
  }
  __syncthreads();
  if (threadIdx.x < kRadix) hist[bd.histBase + (int)threadIdx.x * bd.nbEvent + bd.bLocal] = s_h[threadIdx.x];
}

// Emit(pos, i, key): write the item read from input position i at sorted position pos.
template <class Src, class D, class Emit>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const BlockDesc* __restrict__ blocks, Src src,
                                                            const int* __restrict__ histScanned, Emit emit) {
  __shared__ int s_base[kRadix];
  __shared__ int s_run[kRadix];
  __shared__ int s_cnt[kWaves][kRadix];
  __shared__ int s_off[kWaves][kRadix];
  const BlockDesc bd = blocks[blockIdx.x];
  const int tid = (int)threadIdx.x, lane = lane_id(), wave = tid / kWave;
  if (tid < kRadix) {
    s_base[tid] = histScanned[bd.histBase + tid * bd.nbEvent + bd.bLocal];
    s_run[tid] = 0;
  }
  for (int k = lane; k < kRadix; k += kWave) s_cnt[wave][k] = 0;
  __syncthreads();
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + tid;
    const bool valid = i < bd.end;
    const uint32_t key = valid ? src(i) : 0u;
    const int d = valid ? D::get(key) : kRadix - 1;
    const uint64_t m = match_digit7(d);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (valid && before == 0) s_cnt[wave][d] = (int)__popcll(m);
    __syncthreads();
    if (tid < kRadix) {
      int run = s_run[tid];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        s_off[w][tid] = run;
        run += s_cnt[w][tid];
      }
      s_run[tid] = run;
    }
    __syncthreads();
    if (valid) emit(s_base[d] + s_off[wave][d] + before, i, key);
    for (int k = lane; k < kRadix; k += kWave) s_cnt[wave][k] = 0;
  }
}

// Dense bin table: table[e * stride + b] = first sorted position of event e whose key is >= b,
// for b in [0, nKeys]. A first pass over the sorted keys records where each tile of kTableTile
// bins starts (tileStart, nEvents * (table_tiles(nKeys) + 1) ints); then one block per
// (event, tile) fills the tile's bins from the keys of that range.
inline int table_tiles(int nKeys) { return (nKeys + 1 + kTableTile - 1) / kTableTile; }
void build_table(const uint32_t* keySorted, const BlockDesc* blocksDev, int nBlocks, const int* evOffDev,
                 int nEvents, int nKeys, int64_t stride, int* table, int* tileStart, hipStream_t s);

// ---------------------------------------------------------------- seed labels
// clusterIndex[i] = isSeed[root(i)] ? (#seeds of the event before root) : -1, where root(i) is
// reached by following parent[] (parent[i] == i at roots). rankAll is the exclusive scan of
// isSeed over the batch. parent[] is shortened in place while walking.
void launch_labels(const BlockDesc* blocksDev, int nBlocks, int* parent, const int* isSeed, const int* rankAll,
                   const int* evOffDev, int32_t* label, hipStream_t s);

// Sequential carving of one allocation into 256-byte aligned sub-buffers.
struct Carver {
  char* p;
  template <class T>
  T* take(size_t count) {
    T* r = reinterpret_cast<T*>(p);
    p += (count * sizeof(T) + 255) & ~(size_t)255;
    return r;
  }
};

// Fill an int buffer on the stream with a kernel (graph-capture friendly, no memset nodes).
void fill_i32(int* p, int value, int64_t n, hipStream_t s);

}  // namespace clue
