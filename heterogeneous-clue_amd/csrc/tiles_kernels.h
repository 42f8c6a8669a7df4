// Device code of the tile construction shared by CLUE2D and CLUE3D (device-wide scan, bin table,
// label propagation, fill). Header-only so that the kernels also build for the CPU emulation of the
// HIP execution model used by the tests; the launch functions are in tiles.hip.
#pragma once

#include "tiles.h"

namespace clue {

// ---------------------------------------------------------------- scan

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const int* __restrict__ in, int64_t n,
                                                          int* __restrict__ part) {
  __shared__ int s_tmp[kWaves];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) v += in[base + k];
  int total;
  block_exclusive_sum(v, s_tmp, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Single block: exclusive scan of the partials, chunk by chunk with a carry.
__global__ __launch_bounds__(kThreads) void k_scan_partials(int* __restrict__ part, int np) {
  __shared__ int s_tmp[kWaves];
  int carry = 0;
  for (int c = 0; c < np; c += kScanTile) {
    const int base = c + (int)threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = (base + k < np) ? part[base + k] : 0;
      sum += v[k];
    }
    int total;
    int run = carry + block_exclusive_sum(sum, s_tmp, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < np) part[base + k] = run;
      run += v[k];
    }
    carry += total;
  }
}

// Exclusive scan of the kScanTile items of this block, starting from `offset`.
__device__ __forceinline__ void scan_tile(const int* in, int* out, int64_t n, int offset, int* s_tmp) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    sum += v[k];
  }
  int total;
  int run = offset + block_exclusive_sum(sum, s_tmp, total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

// `part` holds the exclusive scan of the block totals (k_scan_partials).
__global__ __launch_bounds__(kThreads) void k_scan_down(const int* in, int* out, int64_t n,
                                                        const int* __restrict__ part) {
  __shared__ int s_tmp[kWaves];
  scan_tile(in, out, n, part[blockIdx.x], s_tmp);
}

// `part` holds the block totals themselves (at most kScanTile of them): each block adds up those of
// the blocks before it, one coalesced round of kScanItems loads per thread, so that no single-block
// pass over the partials sits between the reduce and this kernel.
__global__ __launch_bounds__(kThreads) void k_scan_down_direct(const int* in, int* out, int64_t n,
                                                               const int* __restrict__ part) {
  __shared__ int s_tmp[kWaves];
  int pre = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int j = k * kThreads + (int)threadIdx.x;
    pre += j < (int)blockIdx.x ? part[j] : 0;
  }
  int offset;
  block_exclusive_sum(pre, s_tmp, offset);
  scan_tile(in, out, n, offset, s_tmp);
}


// ---------------------------------------------------------------- dense bin table

constexpr int kTablePer = kTableTile / kThreads;  // bins per thread

// tileStart[e][k] = first sorted position of event e whose key is >= k * kTableTile, k in [0, tiles].
__global__ __launch_bounds__(kThreads) void k_tile_starts(const BlockDesc* __restrict__ blocks,
                                                          const uint32_t* __restrict__ key,
                                                          const int* __restrict__ evOff, int tiles,
                                                          int* __restrict__ tileStart) {
  const BlockDesc bd = blocks[xcd_block((int)blockIdx.x, (int)gridDim.x)];
  const int s = evOff[bd.event], t = evOff[bd.event + 1];
  int* TS = tileStart + (int64_t)bd.event * (tiles + 1);
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int p = bd.begin + r * kThreads + (int)threadIdx.x;
    if (p >= bd.end) break;
    const int tl = (int)(key[p] / (uint32_t)kTableTile);
    const int prev = (p == s) ? -1 : (int)(key[p - 1] / (uint32_t)kTableTile);
    for (int k = prev + 1; k <= tl; ++k) TS[k] = p;
    if (p == t - 1)
      for (int k = tl + 1; k <= tiles; ++k) TS[k] = t;
  }
}

// LDS index of bin j of a tile, padded by one word per kTablePer so that the per-thread chunks
// [16t, 16t+16) fall into distinct banks.
__device__ __forceinline__ int tpad(int j) { return j + j / kTablePer; }

// Only the entries [0, lim) of an event are written, lim = (the event's last key's layer + layerReach +
// 1) * binsPerLayer + 1 (at most nKeys + 1): no search window of a point reaches further, so the
// table of an event that occupies few layers is written only as far as it is read.
__global__ __launch_bounds__(kThreads) void k_table(const uint32_t* __restrict__ key, const int* __restrict__ evOff,
                                                    const int* __restrict__ tileStart, int tiles, int nKeys,
                                                    int binsPerLayer, int layerReach, int64_t stride,
                                                    int* __restrict__ table) {
  __shared__ int s_first[kTableTile + kTableTile / kTablePer];
  __shared__ int s_tmp[kWaves];
  const int e = blockIdx.x / tiles;
  const int tile = blockIdx.x % tiles;
  const int b0 = tile * kTableTile;
  const int b1 = (b0 + kTableTile) < (nKeys + 1) ? (b0 + kTableTile) : (nKeys + 1);
  const int s = evOff[e], t = evOff[e + 1];
  if (t <= s) return;  // no point of an empty event looks up the table (uniform over the block)
  const int64_t limL = ((int64_t)(key[t - 1] / (uint32_t)binsPerLayer) + layerReach + 1) * binsPerLayer + 1;
  const int lim = limL < (int64_t)nKeys + 1 ? (int)limL : nKeys + 1;
  if (b0 >= lim) return;  // uniform over the block
  const int* TS = tileStart + (int64_t)e * (tiles + 1);
  const int lo = TS[tile], hi = TS[tile + 1];
  const int bEnd = b1 < lim ? b1 : lim;
  for (int j = threadIdx.x; j < kTableTile; j += kThreads) s_first[tpad(j)] = 0x7fffffff;
  __syncthreads();
  // first position of every non-empty bin of the tile (one writer per bin: the run start); four
  // rounds of key loads in flight together
  constexpr int kU = 4;
  for (int p0 = lo; p0 < hi; p0 += kU * kThreads) {
    uint32_t k[kU], kp[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int p = p0 + u * kThreads + (int)threadIdx.x;
      k[u] = p < hi ? key[p] : 0u;
      kp[u] = p < hi && p > lo ? key[p - 1] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int p = p0 + u * kThreads + (int)threadIdx.x;
      if (p < hi && (p == lo || kp[u] != k[u])) s_first[tpad((int)(k[u] - (uint32_t)b0))] = p;
    }
  }
  __syncthreads();
  // table[b] = min(first of bins >= b, hi): a suffix minimum. Thread t owns bins
  // [(kThreads-1-t)*kTablePer, +kTablePer) of the tile, so thread order runs from high bins to low.
  const int j0 = (kThreads - 1 - (int)threadIdx.x) * kTablePer;
  int m = 0x7fffffff;
#pragma unroll
  for (int k = 0; k < kTablePer; ++k) {
    const int v = s_first[tpad(j0 + k)];
    m = v < m ? v : m;
  }
  int run = block_exclusive_min(m, s_tmp);
  run = run < hi ? run : hi;
#pragma unroll
  for (int k = kTablePer - 1; k >= 0; --k) {
    const int v = s_first[tpad(j0 + k)];
    run = v < run ? v : run;
    s_first[tpad(j0 + k)] = run;
  }
  __syncthreads();
  int* out = table + (int64_t)e * stride + b0;
  for (int j = threadIdx.x; j < bEnd - b0; j += kThreads) __builtin_nontemporal_store(s_first[tpad(j)], &out[j]);
}


// ---------------------------------------------------------------- labels

// isSeed -> seed words, in-tile word prefixes and tile counts (tilePre, scanned afterwards).
__global__ __launch_bounds__(kThreads) void k_seed_bits(const int* __restrict__ isSeed, int64_t n,
                                                        SeedRank sr) {
  __shared__ int s_cnt[kSeedTile / kWave];
  const int tid = (int)threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const int64_t base = (int64_t)blockIdx.x * kSeedTile;
  const int64_t w0 = (int64_t)blockIdx.x * (kSeedTile / kWave);
#pragma unroll
  for (int k = 0; k < kSeedTile / kThreads; ++k) {
    const int64_t i = base + k * kThreads + tid;
    const uint64_t m = __ballot(i < n && isSeed[i] != 0);
    if (lane == 0) {
      sr.word[w0 + k * kWaves + wave] = m;
      s_cnt[k * kWaves + wave] = __popcll(m);
    }
  }
  __syncthreads();
  if (tid < kWave) {
    const int v = tid < kSeedTile / kWave ? s_cnt[tid] : 0;
    const int inc = wave_inclusive_sum(v);
    if (tid < kSeedTile / kWave) sr.wordPre[w0 + tid] = (uint16_t)(inc - v);
    if (tid == kSeedTile / kWave - 1) sr.tilePre[blockIdx.x] = inc;
  }
}

// The same prefixes from seed words a kernel has already set (bit i % 64 of word i / 64 for each
// seed i, over words cleared beforehand): one thread per word, two tiles of 32 words per wave.
__global__ __launch_bounds__(kThreads) void k_seed_prefix(int64_t nWords, SeedRank sr) {
  static_assert(kSeedTile / kWave == 32, "two seed tiles per wave");
  const int lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int v = w < nWords ? __popcll(sr.word[w]) : 0;
  int inc = wave_inclusive_sum(v);
  // restart at the wave's second tile; the shuffle runs on every lane (a lane it reads must be active)
  const int firstTile = __shfl(inc, 31, kWave);
  inc -= lane >= 32 ? firstTile : 0;
  if (w < nWords) {
    sr.wordPre[w] = (uint16_t)(inc - v);
    if ((lane & 31) == 31) sr.tilePre[w / 32] = inc;
  }
}


__global__ __launch_bounds__(kThreads) void k_labels(const BlockDesc* __restrict__ blocks,
                                                     const int* __restrict__ parent,
                                                     SeedRank sr, const int* __restrict__ evOff,
                                                     int32_t* __restrict__ label) {
  const BlockDesc bd = blocks[xcd_block((int)blockIdx.x, (int)gridDim.x)];
  const int rank0 = rank_of(sr, evOff[bd.event]);
  // The thread's kSortItems points walk to their roots together: each round issues the next parent
  // read of every unfinished walk before using any, so the dependent loads of different walks
  // overlap. parent[] is only read: every follower's parent is strictly higher in the (rho, index)
  // order and roots point at themselves, so each walk ends. (Writing the roots back into parent[]
  // to shorten later walks costs more in writes than it saves.)
  int cur[kSortItems];
  bool walking[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    walking[r] = i < bd.end;
    cur[r] = walking[r] ? parent[i] : 0;
  }
  bool any = true;
  while (any) {
    int nxt[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) nxt[r] = walking[r] ? parent[cur[r]] : cur[r];
    any = false;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      walking[r] = nxt[r] != cur[r];
      cur[r] = nxt[r];
      any |= walking[r];
    }
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    if (i < bd.end) label[i] = seed_of(sr, cur[r]) ? rank_of(sr, cur[r]) - rank0 : -1;
  }
}

__global__ void k_fill(int* p, int v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace clue
