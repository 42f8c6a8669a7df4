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

// Reads the key of input position i of event ev (and may record it as a side effect).
struct KeyArray {
  const uint32_t* key;
  __device__ __forceinline__ uint32_t operator()(int i, int) const { return key[i]; }
};

// Work a histogram pass does once per block beside its counts (clearing what later kernels of the
// same launch sequence accumulate into), so that it needs no launch of its own.
struct NoPrologue {
  __device__ __forceinline__ void operator()(const BlockDesc&) const {}
};

template <class Src, class D, class Pro = NoPrologue>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const BlockDesc* __restrict__ blocks, Src src,
                                                         int* __restrict__ hist, Pro pro = Pro{}) {
  __shared__ int s_h[kRadix];
  const BlockDesc bd = blocks[xcd_block((int)blockIdx.x, (int)gridDim.x)];
  pro(bd);
  if (threadIdx.x < kRadix) s_h[threadIdx.x] = 0;
  // all of the thread's keys in flight at once
  int d[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    d[r] = i < bd.end ? D::get(src(i, bd.event)) : kRadix - 1;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = bd.begin + r * kThreads + (int)threadIdx.x < bd.end;
    const uint64_t m = match_digit7(d[r]);
    const int before = __popcll(m & ((1ull << lane_id()) - 1ull));
    if (valid && before == 0) atomicAdd(&s_h[d[r]], (int)__popcll(m));
  }
  __syncthreads();
  if (threadIdx.x < kRadix) hist[bd.histBase + (int)threadIdx.x * bd.nbEvent + bd.bLocal] = s_h[threadIdx.x];
}

// Emit::load(i) reads what item i carries (issued for all of a thread's items before ranking);
// Emit::fetch(payload) turns it into what is written (it may gather from global memory: only the
// small payload then passes through LDS) and Emit::store(pos, key, out) writes that at the item's
// sorted position. The block first ranks its items stably by digit (each wave its own consecutive
// items: wave ballot match and a wave-private count per digit, then one prefix over the waves),
// places them in LDS in (digit, rank) order, and then
// writes each digit's run to its global position with consecutive threads on consecutive positions,
// all of a thread's fetches issued before its stores.
template <class Src, class D, class Emit>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const BlockDesc* __restrict__ blocks, Src src,
                                                            const int* __restrict__ histScanned, Emit emit) {
  using Payload = typename Emit::Payload;
  __shared__ int s_base[kRadix];
  __shared__ int s_run[kRadix];
  __shared__ int s_start[kRadix];
  __shared__ int s_cnt[kWaves][kRadix];
  __shared__ int s_off[kWaves][kRadix];
  __shared__ uint32_t s_key[kSortBlock];
  __shared__ Payload s_pay[kSortBlock];
  const BlockDesc bd = blocks[xcd_block((int)blockIdx.x, (int)gridDim.x)];
  const int tid = (int)threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const int nItems = bd.end - bd.begin;
  // wave w holds items [w * kWaveItems, (w + 1) * kWaveItems) of the block, round r its lanes'
  // consecutive items r * kWave + lane: input order is (wave, round, lane)
  constexpr int kWaveItems = kSortItems * kWave;
  const int first = bd.begin + wave * kWaveItems + lane;
  uint32_t key[kSortItems];
  Payload pay[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = first + r * kWave;
    key[r] = i < bd.end ? src(i, bd.event) : 0u;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = first + r * kWave;
    pay[r] = emit.load(i < bd.end ? i : bd.begin);
  }
  if (tid < kRadix) s_base[tid] = histScanned[bd.histBase + tid * bd.nbEvent + bd.bLocal];
  for (int k = lane; k < kRadix; k += kWave) s_cnt[wave][k] = 0;
  // (s_base is read by emit.block_starts after the ranking's barrier)
  // Stable rank inside the wave: per round, the lanes of one digit (ballot match) take consecutive
  // ranks after that digit's running count in the wave; the digit's lowest lane advances the count
  // and broadcasts the old value (the shuffle orders every update of a round before the next
  // round's reads). Invalid items carry digit kRadix - 1, which no key has, and are never placed.
  int rank[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = first + r * kWave < bd.end;
    const int d = valid ? D::get(key[r]) : kRadix - 1;
    const uint64_t m = match_digit7(d);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    int old = 0;
    if (before == 0) {
      old = s_cnt[wave][d];
      s_cnt[wave][d] = old + (int)__popcll(m);
    }
    old = __shfl(old, __builtin_ctzll(m), kWave);  // m holds this lane, so it is not 0
    rank[r] = old + before;
  }
  __syncthreads();
  // the block's digit bases are in LDS now: an Emit may use them (uniform over the block)
  emit.block_starts(bd, s_base);
  // the waves' counts of each digit, in wave order: offsets and the block total
  if (tid < kRadix) {
    int run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      s_off[w][tid] = run;
      run += s_cnt[w][tid];
    }
    s_run[tid] = run;
  }
  __syncthreads();
  // start of each digit's run within the block (exclusive scan of the block's digit counts)
  if (tid < kWave) {
    int v0 = s_run[2 * tid], v1 = s_run[2 * tid + 1];
    const int inc = wave_inclusive_sum(v0 + v1);
    s_start[2 * tid] = inc - v0 - v1;
    s_start[2 * tid + 1] = inc - v1;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (first + r * kWave < bd.end) {
      const int d = D::get(key[r]);
      const int pos = s_start[d] + s_off[wave][d] + rank[r];
      s_key[pos] = key[r];
      s_pay[pos] = pay[r];
    }
  }
  __syncthreads();
  typename Emit::Out out[kSortItems];
  uint32_t okey[kSortItems];
  int opos[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int j = r * kThreads + tid;
    okey[r] = 0u;
    opos[r] = 0;
    if (j < nItems) {
      okey[r] = s_key[j];
      const int d = D::get(okey[r]);
      opos[r] = s_base[d] + (j - s_start[d]);
      out[r] = emit.fetch(s_pay[j]);
    }
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r)
    if (r * kThreads + tid < nItems) emit.store(opos[r], okey[r], out[r]);
}

// First pass: the item is its input position.
struct EmitFirst {
  uint32_t* kOut;
  int* iOut;
  using Payload = int;
  using Out = int;
  __device__ __forceinline__ void block_starts(const BlockDesc&, const int*) const {}
  __device__ __forceinline__ Payload load(int i) const { return i; }
  __device__ __forceinline__ Out fetch(Payload v) const { return v; }
  __device__ __forceinline__ void store(int pos, uint32_t k, Out v) const {
    kOut[pos] = k;
    iOut[pos] = v;
  }
};

// Middle passes: the item carries the point index it was given by the previous pass.
struct EmitMid {
  const int* iIn;
  uint32_t* kOut;
  int* iOut;
  using Payload = int;
  using Out = int;
  __device__ __forceinline__ void block_starts(const BlockDesc&, const int*) const {}
  __device__ __forceinline__ Payload load(int i) const { return iIn[i]; }
  __device__ __forceinline__ Out fetch(Payload v) const { return v; }
  __device__ __forceinline__ void store(int pos, uint32_t k, Out v) const {
    kOut[pos] = k;
    iOut[pos] = v;
  }
};

// Dense bin table: table[e * stride + b] = first sorted position of event e whose key is >= b,
// for b in [0, nKeys]. A first pass over the sorted keys records where each tile of kTableTile
// bins starts (tileStart, nEvents * (table_tiles(nKeys) + 1) ints); then one block per
// (event, tile) fills the tile's bins from the keys of that range.
inline int table_tiles(int nKeys) { return (nKeys + 1 + kTableTile - 1) / kTableTile; }
// Entries of an event beyond layerReach layers past its last occupied layer (binsPerLayer bins each) are
// not written: callers guarantee that no lookup reaches them.
void build_table(const uint32_t* keySorted, const BlockDesc* blocksDev, int nBlocks, const int* evOffDev,
                 int nEvents, int nKeys, int binsPerLayer, int layerReach, int64_t stride, int* table,
                 int* tileStart, hipStream_t s);

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

// ---------------------------------------------------------------- seed labels
// Seed ranks in point-index order without a scan of the whole isSeed array: one pass packs isSeed
// into 64-bit words (one per 64 points) with the exclusive prefix of each word inside its tile of
// kSeedTile points (u16) and the tile's count; a scan of the tile counts completes the ranks:
//   rank(i) = tilePre[i / kSeedTile] + wordPre[i / 64] + popcount(word[i / 64] below bit i % 64).
constexpr int kSeedTile = kScanTile;  // 2048 points, 32 words
struct SeedRank {
  uint64_t* word;
  uint16_t* wordPre;
  int* tilePre;
};
inline int seed_tiles(int64_t N) { return (int)((N + kSeedTile - 1) / kSeedTile); }
inline size_t seed_rank_bytes(int64_t N) {
  const size_t t = (size_t)(seed_tiles(N) > 0 ? seed_tiles(N) : 1);
  auto a = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return a(t * 32 * sizeof(uint64_t)) + a(t * 32 * sizeof(uint16_t)) + a((t + 1) * sizeof(int));
}
inline SeedRank seed_rank_carve(Carver& cv, int64_t N) {
  const size_t t = (size_t)(seed_tiles(N) > 0 ? seed_tiles(N) : 1);
  SeedRank r;
  r.word = cv.take<uint64_t>(t * 32);
  r.wordPre = cv.take<uint16_t>(t * 32);
  r.tilePre = cv.take<int>(t + 1);
  return r;
}

// Seed bit and seed rank (seeds of the batch before point i, in index order) of point i.
__device__ __forceinline__ bool seed_of(const SeedRank& sr, int i) {
  return (sr.word[i >> 6] >> (i & 63)) & 1ull;
}
__device__ __forceinline__ int rank_of(const SeedRank& sr, int i) {
  return sr.tilePre[i / kSeedTile] + sr.wordPre[i >> 6] + __popcll(sr.word[i >> 6] & ((1ull << (i & 63)) - 1ull));
}

// clusterIndex[i] = isSeed[root(i)] ? (#seeds of the event before root) : -1, where root(i) is
// reached by following parent[] (parent[i] == i at roots). `part` needs scan_partials(seed_tiles(N))
// ints. isSeed == nullptr: the seed words are already set (cleared, then one bit per seed set by the
// caller's kernels).
void launch_labels(const BlockDesc* blocksDev, int nBlocks, const int* parent, const int* isSeed, int64_t N,
                   const int* evOffDev, int32_t* label, const SeedRank& sr, int* part, hipStream_t s);
// The seed words' prefixes and scanned tile counts alone (the words already set), for callers with
// their own labels kernel.
void launch_seed_ranks(int64_t N, const SeedRank& sr, int* part, hipStream_t s);

// Fill an int buffer on the stream with a kernel (graph-capture friendly, no memset nodes).
void fill_i32(int* p, int value, int64_t n, hipStream_t s);

}  // namespace clue
