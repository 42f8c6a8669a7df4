// Device code of the CLUE2D global-sort path: the layer pass functors, the per-layer bin sort that
// writes the bin table, and the density / nearest-higher kernels. Header-only so that the kernels
// also build for the CPU emulation of the HIP execution model used by the tests; the launch sequence
// is in clue2d.hip.
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
#pragma once

#include <float.h>
#include <math.h>

#include "clue2d.h"

namespace clue {
namespace g2d {

// ---------------------------------------------------------------- tile order
//
// The sorted array lists every (event, layer, xBin, yBin) bin in ascending point index. It is built
// in two steps:
//  (1) a stable counting-sort pass by layer (the radix-pass kernels of tiles.h with the layer as the
//      digit), whose output is each point as {x, y, w, index bits}, grouped by (event, layer) and in
//      index order inside a layer (the first block of each event also writes where each layer
//      begins);
//  (2) k_segsort2d: one workgroup per (event, layer) orders that layer's points by bin, keeping index
//      order inside a bin, writes the sorted {x, y, w, index} records (and the bin keys of each
//      256-point tile's first and last position), and writes the layer's 10^4 entries of the
//      event's bin table from its own bin counts.

// Layer of point i of event ev, the digit of step (1); an out-of-range layer sets kErrLayerRange in
// err[ev * errStride] (errStride 0: one flag word for the whole batch) and is clamped.
struct LayerOf2D {
  const int32_t* layer;
  int* err;
  int errStride;
  __device__ __forceinline__ uint32_t operator()(int i, int ev) const {
    int L = layer[i];
    if (L < 0 || L >= k2dLayers) {
      atomicOr(err + ev * errStride, kErrLayerRange);
      L = L < 0 ? 0 : k2dLayers - 1;
    }
    return (uint32_t)L;
  }
};
using DigitLayer2D = Digit<1, kRadix>;

// Output of step (1): only the point index passes through the block's LDS reorder; x, y and w are
// then read from the block's own input range (consecutive points of one event) and written as one
// 16-byte record at the point's position in layer order.
struct EmitSeg2D {
  const float* x;
  const float* y;
  const float* w;
  float4* lg;
  int* segB;     // per event: the start of each layer and the event's end (k2dLayers + 1 entries)
  int* segList;  // the layers above kSegCap points, as e * k2dLayers + L
  int* segCnt;   // their count (zero before the pass)
  using Payload = int;
  using Out = float4;
  // The first block of each event (bLocal 0) holds the event's digit bases: base[L] is where layer L
  // of the event begins and base[k2dLayers] (a digit no point has) where the event ends. Such a
  // block writes them as segB and lists the layers above kSegCap points (one atomic per wave that
  // has any); called by every thread of the block.
  __device__ __forceinline__ void block_starts(const BlockDesc& bd, const int* base) const;
  __device__ __forceinline__ Payload load(int i) const { return i; }
  __device__ __forceinline__ Out fetch(int idx) const { return make_float4(x[idx], y[idx], w[idx], __int_as_float(idx)); }
  __device__ __forceinline__ void store(int pos, uint32_t, const Out& v) const { lg[pos] = v; }
};

// Segment size classes of k_segsort2d: up to kSegSmall points (kSegThreadsS threads, 28 KB of LDS,
// 5 workgroups per CU), up to kSegCap points (kSegThreadsL threads, 72 KB, 2 per CU), and larger
// ones (kSegThreadsH threads, slots in global memory). The two LDS classes run over all (event,
// layer) segments and leave the others at once; the largest class, rare, takes its segments from a
// list (written by the layer pass, EmitSeg2D::block_starts), so that its launch holds few empty
// workgroups.
constexpr int kSegSmall = 2048;
constexpr int kSegCap = 10240;
constexpr int kSegThreadsS = 256;
constexpr int kSegThreadsL = 1024;
constexpr int kSegThreadsH = 512;
constexpr int kSegNone = 0;  // "no upper bound": the class of the segments above kSegCap

__device__ __forceinline__ void EmitSeg2D::block_starts(const BlockDesc& bd, const int* base) const {
  if (bd.bLocal != 0) return;  // uniform over the block
  const int L = (int)threadIdx.x;
  if (L <= k2dLayers) segB[bd.event * (k2dLayers + 1) + L] = base[L];
  const bool big = L < k2dLayers && base[L + 1] - base[L] > kSegCap;
  const uint64_t m = __ballot(big);  // every lane of the wave is here
  if (m != 0ull) {
    const int lane = lane_id(), leader = __builtin_ctzll(m);
    int first = 0;
    if (lane == leader) first = atomicAdd(segCnt, __popcll(m));
    first = __shfl(first, leader, kWave);
    if (big) segList[first + __popcll(m & ((1ull << lane) - 1ull))] = bd.event * k2dLayers + L;
  }
}

// Once per block of the layer pass's histogram kernel: the block's words of the seed bits (which the
// nearest-higher kernel sets with atomics; a word shared with the next block is cleared by both), the
// padding words after the batch's last point (the block that ends the batch), and the count of the
// largest bin-sort class's list (the block that starts the batch).
struct Prologue2D {
  uint64_t* word;
  int64_t nWords;
  int64_t N;
  int* segCnt;
  __device__ __forceinline__ void operator()(const BlockDesc& bd) const {
    const int64_t w0 = (int64_t)bd.begin >> 6;
    const int64_t w1 = bd.end == N ? nWords : ((int64_t)bd.end + 63) >> 6;
    for (int64_t k = w0 + (int64_t)threadIdx.x; k < w1; k += kThreads) word[k] = 0ull;
    if (bd.begin == 0 && threadIdx.x == 0) *segCnt = 0;
  }
};

constexpr int kSegIlp = 4;   // points of a thread taken together in each loop over a segment
// Points of a thread (positions tid + r * NT, r < kSegHold) whose records the kSegCap class keeps in
// registers from the count pass to the output pass, so that they are read once; with the LDS of two
// workgroups per CU the registers of 8 waves per SIMD (64 each) hold five records.
constexpr int kSegHold = 5;
template <int CAP>
constexpr int kSegHoldOf = CAP == kSegCap ? kSegHold : 0;

__device__ __forceinline__ int bin2d(const float4& v) { return xbin2d(v.x) * k2dRows + ybin2d(v.y); }

// The bin table of an event: entry b (key b = layer * 10^4 + bin, b <= k2dKeys) is the first sorted
// position whose key is >= b. Stored as a 16-bit offset from the start of b's layer (segB) when the
// layer has at most kTab16Max points, else as a 32-bit position in t32 (written only for such
// layers); the closing entry b = k2dKeys is the event's end (offset 0 from segB[k2dLayers]).
constexpr int kTab16Max = 65535;
// The entries of one layer L: j in [0, 10^4] is entry L * 10^4 + j (j = 10^4, the next layer's first
// entry, is where layer L ends).
struct LayerTab2D {
  const uint16_t* t16;
  const int* t32;
  int off;  // L * 10^4
  int base, next;
  bool wide;
  __device__ __forceinline__ int operator[](int j) const {
    const int v = wide ? t32[off + j] : base + (int)t16[off + j];
    return j < k2dBinsPerLayer ? v : next;
  }
  // the same for a layer of 16-bit entries: one load whose value is always used, so that a point's
  // loads are issued together (entry j = 10^4 is the next layer's bin 0, offset 0 from `next`)
  __device__ __forceinline__ int narrow(int j) const {
    return (j < k2dBinsPerLayer ? base : next) + (int)t16[off + j];
  }
};
struct BinTable2D {
  const uint16_t* t16;
  const int* t32;
  const int* segB;  // the event's k2dLayers + 1 layer starts
  // any entry: the 16-bit word is loaded together with the layer starts; the 32-bit one only for
  // a layer that uses it
  __device__ __forceinline__ int operator[](int b) const {
    const int L = b / k2dBinsPerLayer;
    const int s = segB[L], e = segB[L < k2dLayers ? L + 1 : L];
    const int v16 = s + (int)t16[b];
    if (e - s <= kTab16Max) return v16;
    return t32[b];
  }
  __device__ __forceinline__ LayerTab2D layer(int L) const {
    const int s = segB[L], e = segB[L + 1];
    return LayerTab2D{t16 + L * k2dBinsPerLayer, t32 + L * k2dBinsPerLayer, 0, s, e, e - s > kTab16Max};
  }
};
// The bin tables of a batch (one row of `stride` entries per event in each array).
struct BinTables2D {
  const uint16_t* t16;
  const int* t32;
  const int* segB;
  int64_t stride;
  __device__ __forceinline__ BinTable2D event(int e) const {
    return BinTable2D{t16 + (int64_t)e * stride, t32 + (int64_t)e * stride, segB + e * (k2dLayers + 1)};
  }
};

// Exclusive prefix sum over the NT threads of a block (`s_tmp`: NT / kWave ints); two barriers, call
// from uniform control flow.
template <int NT>
__device__ __forceinline__ int block_exclusive_sum_nt(int v, int* s_tmp) {
  constexpr int W = NT / kWave;
  const int lane = lane_id(), wave = (int)threadIdx.x / kWave;
  const int inc = wave_inclusive_sum(v);
  if (lane == kWave - 1) s_tmp[wave] = inc;
  __syncthreads();
  int base = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) base += w < wave ? s_tmp[w] : 0;
  __syncthreads();
  return base + inc - v;
}

using v4i = int __attribute__((ext_vector_type(4)));

// Bin counts of a segment in LDS. LDS class (CAP > 0): two 16-bit counts per 32-bit word (a segment
// has at most CAP < 65536 points), each thread owning SCAN consecutive bins (a multiple of 8, so its
// bins are whole 16-byte words); the 32-bit atomics add 1 << 16 for an odd bin, which never carries
// into the neighbouring count. Global class: one 32-bit count per bin.
template <int NT, int CAP>
struct SegCounts {
  static constexpr bool kPacked = CAP != kSegNone;
  static constexpr int kScan = kPacked ? (k2dBinsPerLayer + NT * 8 - 1) / (NT * 8) * 8 : 0;  // bins per thread
  static constexpr int kWords = kPacked ? NT * kScan / 2 : k2dBinsPerLayer;                 // 32-bit words
  static_assert(!kPacked || CAP < 65536, "16-bit counts");
  static_assert(kWords % 4 == 0, "whole 16-byte words");
  uint32_t* w;
  __device__ __forceinline__ int get(int b) const {
    if constexpr (kPacked)
      return (int)((w[b >> 1] >> (16 * (b & 1))) & 0xFFFFu);
    else
      return (int)w[b];
  }
  // adds one to bin b and returns its previous value
  __device__ __forceinline__ int take(int b) const {
    if constexpr (kPacked) {
      const int sh = 16 * (b & 1);
      return (int)((atomicAdd(&w[b >> 1], 1u << sh) >> sh) & 0xFFFFu);
    } else {
      return (int)atomicAdd(&w[b], 1u);
    }
  }
};

// Step (2) for the layer segment lg[start, end) of event e and layer L, id = e * 100 + L: id =
// blockIdx.x, or the blockIdx.x-th entry of the class's list when one is given (workgroups past the
// list's count leave at once), if its size n is in this class, (LO, CAP] (CAP = kSegNone: above LO). LayerTilesSerial::fill
// (LayerTilesSerial.h:22) appends in index order, and the segment arrives in index order:
//   counts of the 10^4 bins (LDS atomics); their exclusive scan is each bin's first position, written
//   as the layer's bin-table entries; every point then takes the next free slot of its bin (LDS
//   atomic, so the order inside a bin is arbitrary); finally each slot's point counts the points of
//   its bin with a smaller position, which is its rank in index order, and is written at bin start +
//   rank. The ranking costs the sum over bins of (points in the bin)^2, the order of the density
//   step's own work on that bin. The LDS classes keep each point's bin and each slot's point in LDS
//   (u16); the class above kSegCap keeps the slots in `scr` (the segment's own range of a global
//   array) and recomputes bins. Empty segments take the smallest class: their table entries are
//   written too. The kSegCap class ranks point by point instead, its first kSegHold * NT points from
//   the records it kept in registers since the count pass (the rest read again, in index order):
//   its segments (up to 160 KB) are evicted from L2 between the two passes.
template <int NT, int LO, int CAP>
__global__ __launch_bounds__(NT) CLUE_WAVES_PER_EU(kSegHoldOf<CAP> > 0 ? 8 : 1) void k_segsort2d(const int* __restrict__ evOff, const int* __restrict__ segB,
                                                  const float4* __restrict__ lg, int* __restrict__ scr,
                                                  uint32_t* __restrict__ keyOut, float4* __restrict__ xyw,
                                                  uint16_t* __restrict__ t16, int* __restrict__ t32, int64_t stride,
                                                  const int* __restrict__ segList, const int* __restrict__ segCnt) {
  using C = SegCounts<NT, CAP>;
  constexpr bool kLds = CAP != kSegNone;
  constexpr int kPts = kLds ? CAP : 1;
  __shared__ __align__(16) uint32_t s_w[C::kWords];
  __shared__ uint16_t s_bin[kPts];
  __shared__ uint16_t s_slot[kPts];
  __shared__ int s_tmp[NT / kWave];
  int id = (int)blockIdx.x;
  if (segList) {
    if (id >= *segCnt) return;  // uniform over the block
    id = segList[id];
  }
  const int e = id / k2dLayers, L = id % k2dLayers;
  const int evBegin = evOff[e], evEnd = evOff[e + 1];
  if (evBegin == evEnd) return;  // an empty event: nothing reads its table (uniform over the block)
  const int start = segB[e * (k2dLayers + 1) + L];
  const int n = segB[e * (k2dLayers + 1) + L + 1] - start;
  if (n < LO || (LO > 0 && n == LO) || (kLds && n > CAP)) return;  // another class (uniform)
  const C cnt{s_w};
  const int tid = (int)threadIdx.x;
  {
    uint4* w4 = reinterpret_cast<uint4*>(s_w);
    for (int k = tid; k < C::kWords / 4; k += NT) w4[k] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();
  // bin counts; every loop over the points takes kSegIlp of a thread's points at once, all their
  // loads issued before any is used (the held points first, all at once)
  constexpr int kHold = kSegHoldOf<CAP>;
  float4 held[kHold > 0 ? kHold : 1];
  if constexpr (kHold > 0) {
#pragma unroll
    for (int r = 0; r < kHold; ++r) {
      const int p = tid + r * NT;
      held[r] = lg[start + (p < n ? p : 0)];
    }
#pragma unroll
    for (int r = 0; r < kHold; ++r) {
      const int p = tid + r * NT;
      if (p < n) {
        const int b = bin2d(held[r]);
        s_bin[p] = (uint16_t)b;
        cnt.take(b);
      }
    }
  }
  for (int p0 = tid + kHold * NT; p0 < n; p0 += kSegIlp * NT) {
    float4 v[kSegIlp];
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) {
      const int p = p0 + r * NT;
      v[r] = lg[start + (p < n ? p : 0)];
    }
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) {
      const int p = p0 + r * NT;
      if (p < n) {
        const int b = bin2d(v[r]);
        if constexpr (kLds) s_bin[p] = (uint16_t)b;
        cnt.take(b);
      }
    }
  }
  __syncthreads();
  // exclusive scan of the counts
  if constexpr (kLds) {
    // thread t owns bins [t * kScan, (t + 1) * kScan): kScan / 8 whole 16-byte words
    constexpr int kV = C::kScan / 8;
    uint4* w4 = reinterpret_cast<uint4*>(s_w) + tid * kV;
    uint4 v[kV];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kV; ++k) {
      v[k] = w4[k];
      const uint32_t q[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) sum += (int)(q[j] & 0xFFFFu) + (int)(q[j] >> 16);
    }
    uint32_t run = (uint32_t)block_exclusive_sum_nt<NT>(sum, s_tmp);
#pragma unroll
    for (int k = 0; k < kV; ++k) {
      uint32_t q[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t lo = run, hi = run + (q[j] & 0xFFFFu);
        run = hi + (q[j] >> 16);
        q[j] = lo | (hi << 16);
      }
      w4[k] = make_uint4(q[0], q[1], q[2], q[3]);
    }
  } else {
    constexpr int kScan = (k2dBinsPerLayer + NT - 1) / NT;
    const int b0 = tid * kScan;
    int sum = 0;
    for (int k = 0; k < kScan; ++k)
      if (b0 + k < k2dBinsPerLayer) sum += (int)s_w[b0 + k];
    uint32_t run = (uint32_t)block_exclusive_sum_nt<NT>(sum, s_tmp);
    for (int k = 0; k < kScan; ++k)
      if (b0 + k < k2dBinsPerLayer) {
        const uint32_t c = s_w[b0 + k];
        s_w[b0 + k] = run;
        run += c;
      }
  }
  __syncthreads();
  // the layer's bin-table entries: each bin's first position as an offset from the layer's start
  // (in the LDS classes the scanned count words already hold them in that layer-row format, two
  // 16-bit entries per word); the entry after the last layer closes the table. Full-line writes
  // read back sparsely, kept out of L2.
  {
    const int64_t row = (int64_t)e * stride + (int64_t)L * k2dBinsPerLayer;  // 16-byte aligned
    static_assert(k2dBinsPerLayer % 8 == 0, "whole 16-byte table words per layer");
    if constexpr (C::kPacked) {
      using v4u = unsigned __attribute__((ext_vector_type(4)));
      const v4u* w4 = reinterpret_cast<const v4u*>(s_w);
      v4u* D = reinterpret_cast<v4u*>(t16 + row);
      for (int k = tid; k < k2dBinsPerLayer / 8; k += NT) __builtin_nontemporal_store(w4[k], D + k);
    } else if (n <= kTab16Max) {
      for (int b = tid; b < k2dBinsPerLayer; b += NT) __builtin_nontemporal_store((uint16_t)cnt.get(b), t16 + row + b);
    } else {
      for (int b = tid; b < k2dBinsPerLayer; b += NT) __builtin_nontemporal_store(start + cnt.get(b), t32 + row + b);
      if (tid == 0) t16[row] = 0;  // bin 0 starts the layer: read as the previous layer's end entry
    }
    if (L == k2dLayers - 1 && tid == 0) {
      t16[(int64_t)e * stride + k2dKeys] = 0;
      t32[(int64_t)e * stride + k2dKeys] = evEnd;
    }
  }
  __syncthreads();  // the bin starts are read above before the slots advance them
  // slots
  for (int p0 = tid; p0 < n; p0 += kSegIlp * NT) {
    int b[kSegIlp];
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) {
      const int p = p0 + r * NT;
      if constexpr (kLds)
        b[r] = p < n ? (int)s_bin[p] : 0;
      else
        b[r] = bin2d(lg[start + (p < n ? p : 0)]);
    }
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) {
      const int p = p0 + r * NT;
      if (p < n) {
        const int slot = cnt.take(b[r]);
        if constexpr (kLds)
          s_slot[slot] = (uint16_t)p;
        else
          scr[start + slot] = p;
      }
    }
  }
  __syncthreads();
  // the count words now hold each bin's end; rank each point inside its bin by position
  const auto emit = [&](int b, int bs, int rank, const float4& v) {
    const int out = start + bs + rank;
    // keys only where k_reach2d reads them: the first and last position of every 256-point tile
    // (blocks start every kSortBlock points from the event's start, make_blocks) and the event's
    // last position
    const int rel = (out - evBegin) & (kThreads - 1);
    if (rel == 0 || rel == kThreads - 1 || out == evEnd - 1) keyOut[out] = (uint32_t)(L * k2dBinsPerLayer + b);
    xyw[out] = v;
  };
  if constexpr (kHold > 0) {
    // by point: the rank of p is the number of its bin's slots holding a smaller position
    const auto emit_point = [&](int p, const float4& v) {
      const int b = (int)s_bin[p];
      const int bs = b > 0 ? cnt.get(b - 1) : 0, be = cnt.get(b);
      int rank = 0;
      for (int f = bs; f < be; ++f) rank += (int)s_slot[f] < p ? 1 : 0;
      emit(b, bs, rank, v);
    };
#pragma unroll
    for (int r = 0; r < kHold; ++r) {
      const int p = tid + r * NT;
      if (p < n) emit_point(p, held[r]);
    }
    for (int p0 = tid + kHold * NT; p0 < n; p0 += kSegIlp * NT) {
      float4 v[kSegIlp];
#pragma unroll
      for (int r = 0; r < kSegIlp; ++r) {
        const int p = p0 + r * NT;
        v[r] = lg[start + (p < n ? p : 0)];
      }
#pragma unroll
      for (int r = 0; r < kSegIlp; ++r) {
        const int p = p0 + r * NT;
        if (p < n) emit_point(p, v[r]);
      }
    }
    return;
  }
  // by slot: each slot's point counts the points of its bin with a smaller position
  for (int j0 = tid; j0 < n; j0 += kSegIlp * NT) {
    int p[kSegIlp];
    float4 v[kSegIlp];
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) {
      const int j = j0 + r * NT < n ? j0 + r * NT : 0;
      if constexpr (kLds)
        p[r] = (int)s_slot[j];
      else
        p[r] = scr[start + j];
    }
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) v[r] = lg[start + p[r]];
#pragma unroll
    for (int r = 0; r < kSegIlp; ++r) {
      if (j0 + r * NT >= n) continue;
      int b;
      if constexpr (kLds)
        b = (int)s_bin[p[r]];
      else
        b = bin2d(v[r]);
      const int bs = b > 0 ? cnt.get(b - 1) : 0, be = cnt.get(b);
      int rank = 0;
      for (int f = bs; f < be; ++f) {
        int q;
        if constexpr (kLds)
          q = (int)s_slot[f];
        else
          q = scr[start + f];
        rank += q < p[r] ? 1 : 0;
      }
      emit(b, bs, rank, v[r]);
    }
  }
}

// The density and nearest-higher kernels give each block of NT threads NT * ITEMS consecutive sorted
// points of one event (ITEMS points per thread). In x-major tile order everything their search boxes
// can reach lies in one contiguous range of the sorted array, whole columns from the lowest box's
// first column to the highest box's last (reach_columns, computed from the block's first and last
// keys, so staging starts without waiting for the points' own bin-table lookups). When that range
// has at most STAGE points it is staged into LDS once and every candidate is read from there;
// otherwise candidates are read from global memory. Either way the candidates are visited in the
// serial order, so the results do not depend on NT, ITEMS, STAGE or which path ran; the plan
// autotuner picks (NT, ITEMS, STAGE) per kernel among kVariants2D by timing. Larger blocks stage
// fewer candidates per point (the reach beyond the block's own columns is shared by more points).
template <int NT, int ITEMS>
struct CompShape {
  static constexpr int kHits = NT * ITEMS;               // points per compute block
  static constexpr int kPerSort = kSortBlock / kHits;    // compute blocks per sort block
  static constexpr int kTiles = kHits / kThreads;        // 256-point reach tiles per compute block
  static_assert(kSortBlock % kHits == 0 && kHits % kThreads == 0, "compute blocks tile each sort block");
};
constexpr int kMaxCols = 5;  // column runs kept in registers per point

// Waves per SIMD the compiler must allow: one point per thread with the 1024-point stage (or any
// larger block) admits 8 by LDS, two points per thread with it 7; registers are held to 64 and 72.
template <int NT, int ITEMS, int STAGE>
constexpr int kWavesPerEu2D = ITEMS == 1 ? 8 : (ITEMS == 2 && STAGE == 1024) ? 7 : 1;

struct Variant2D {
  int threads;  // threads per block
  int items;    // points per thread
  int stage;    // staged points
};
constexpr int kNumVariants2D = 6;
constexpr Variant2D kVariants2D[kNumVariants2D] = {{256, 2, 1024}, {256, 1, 1024}, {256, 2, 2048},
                                                   {256, 4, 2048}, {512, 1, 1536}, {1024, 1, 2048}};

// A search box of the reference (LayerTilesSerial::searchBox) as column runs: run k covers column
// xlo + k, rows ylo..yhi, which are the sorted positions [s[k], t[k]). Columns beyond kMaxCols
// (only for large distances) are read from the table when they are scanned.
struct Runs {
  int s[kMaxCols], t[kMaxCols];
  int n, L, xlo, xhi, ylo, yhi;
};

// The event's bin table as a block reads it: when its window fits (staged, uniform over the block),
// the raw 16-bit entries [base, base + n) and the starts of the window's layers from L0 on are in the
// block's LDS, else everything comes from global memory. A point of a layer with 32-bit entries
// (more than kTab16Max points) reads those from global memory either way.
struct TableView {
  BinTable2D g;
  const uint16_t* s;      // staged entry b at s[b - base]
  const int* layerStart;  // staged start of layer L at layerStart[L - L0]
  int base, L0;
  bool staged;
};

__device__ __forceinline__ void make_runs(const TableView& T, int L, float xi, float yi, float d, Runs& r) {
  r.L = L;
  r.xlo = xbin2d(xi - d);
  r.xhi = xbin2d(xi + d);
  r.ylo = ybin2d(yi - d);
  r.yhi = ybin2d(yi + d);
  r.n = r.xhi - r.xlo + 1;
  const int lb = T.staged ? T.layerStart[r.L - T.L0] : 0, nb = T.staged ? T.layerStart[r.L + 1 - T.L0] : 0;
  if (T.staged && nb - lb <= kTab16Max) {
    // entry j of the layer is its start + the 16-bit offset; j = 10^4 is the next layer's bin 0
    const uint16_t* S = T.s + (r.L * k2dBinsPerLayer - T.base);
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) {
      const int xb = r.xlo + k < r.xhi ? r.xlo + k : r.xhi;
      const int j = xb * k2dRows + r.yhi + 1;
      r.s[k] = lb + (int)S[xb * k2dRows + r.ylo];
      r.t[k] = (j < k2dBinsPerLayer ? lb : nb) + (int)S[j];
    }
  } else {
    // every lookup of a box stays in the point's own layer: the ten 16-bit loads are issued together
    // with the layer's starts, and only a wave with a lane in a layer of 32-bit entries reads those
    const LayerTab2D LT = T.g.layer(r.L);
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) {
      const int xb = r.xlo + k < r.xhi ? r.xlo + k : r.xhi;
      r.s[k] = LT.narrow(xb * k2dRows + r.ylo);
      r.t[k] = LT.narrow(xb * k2dRows + r.yhi + 1);
    }
    if (__ballot(LT.wide) != 0) {
#pragma unroll
      for (int k = 0; k < kMaxCols; ++k) {
        const int xb = r.xlo + k < r.xhi ? r.xlo + k : r.xhi;
        r.s[k] = LT[xb * k2dRows + r.ylo];
        r.t[k] = LT[xb * k2dRows + r.yhi + 1];
      }
    }
  }
}

// The sorted range every search box (distance d) of the block's points [b0, b1) can reach, from the
// block's first and last keys alone. A point in column xb has xbin(x - d) >= xb - m and
// xbin(x + d) <= xb + m with m = ceil(d * rX + 0.01): the fp32 bin arithmetic errs by far less than
// 0.01 of a bin, and clamped or NaN coordinates only narrow the box. The sorted array holds whole
// columns in key order, so [first position of column max(0, xb0 - m) of layer L0, first position
// after column min(99, xb1 + m) of layer L1) holds every candidate of every point of the block, and
// the bin-table entries [tlo, thi] hold every lookup of their boxes (the largest is the entry after
// the last reachable column's last row). reach_columns gives tlo and thi; the range is T[tlo], T[thi].
__device__ __forceinline__ void reach_columns(uint32_t k0, uint32_t k1, float d, int& tlo, int& thi) {
  const float rX = (float)k2dCols / (250.f - (-250.f));
  const float md = d * rX + 0.01f;
  const int m = md < 98.f ? (md > 0.f ? (int)ceilf(md) : 0) : k2dCols - 1;  // NaN: whole layers
  const int L0 = (int)(k0 / (uint32_t)k2dBinsPerLayer), x0 = (int)(k0 / (uint32_t)k2dRows) % k2dCols;
  const int L1 = (int)(k1 / (uint32_t)k2dBinsPerLayer), x1 = (int)(k1 / (uint32_t)k2dRows) % k2dCols;
  const int c0 = x0 - m > 0 ? x0 - m : 0;
  const int c1 = x1 + m < k2dCols - 1 ? x1 + m : k2dCols - 1;
  tlo = L0 * k2dBinsPerLayer + c0 * k2dRows;
  thi = L1 * k2dBinsPerLayer + (c1 + 1) * k2dRows;
}

// The reach of every 256-point tile of the sorted batch for the density (dc) and nearest-higher
// (dm) distances, one thread per tile, as {first sorted position, end position, first table entry,
// last table entry}. A compute block of several consecutive tiles of one event reaches from its
// first tile's starts to its last tile's ends (both grow with the sorted keys), so the neighbour
// kernels read two records instead of walking the keys and the table before they can stage.
constexpr int kTilesPerSort = kSortBlock / kThreads;
constexpr int kReachThreads = 64;  // one wave per block: the few hundred blocks spread over all CUs
__global__ __launch_bounds__(kReachThreads) void k_reach2d(const BlockDesc* __restrict__ blocks, int nTiles,
                                                           const uint32_t* __restrict__ keyS, BinTables2D tabs,
                                                           float dc, float odf, int4* __restrict__ reachD,
                                                           int4* __restrict__ reachN) {
  const int tile = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (tile >= nTiles) return;
  const BlockDesc bd = blocks[tile / kTilesPerSort];
  const int b0 = bd.begin + (tile % kTilesPerSort) * kThreads;
  if (b0 >= bd.end) return;  // past the event's last point: no compute block reads it
  const int b1 = b0 + kThreads < bd.end ? b0 + kThreads : bd.end;
  const BinTable2D T = tabs.event(bd.event);
  const uint32_t k0 = keyS[b0], k1 = keyS[b1 - 1];
  int tloD, thiD, tloN, thiN;
  reach_columns(k0, k1, dc, tloD, thiD);
  reach_columns(k0, k1, odf * dc, tloN, thiN);
  const int loD = T[tloD], hiD = T[thiD], loN = T[tloN], hiN = T[thiN];
  reachD[tile] = make_int4(loD, hiD, tloD, thiD);
  reachN[tile] = make_int4(loN, hiN, tloN, thiN);
}

// The reach of compute block `blk` (Shape::kTiles tiles) covering sorted points [b0, b1).
template <class Shape>
__device__ __forceinline__ void compute_reach(const int4* __restrict__ reach, int blk, int b0, int b1, int& lo,
                                              int& hi, int& tlo, int& tn) {
  const int t0 = (blk / Shape::kPerSort) * kTilesPerSort + (blk % Shape::kPerSort) * Shape::kTiles;
  // a one-tile block reads one record, whose load then need not wait for the block's range
  const int t1 = Shape::kTiles == 1 ? t0 : t0 + (b1 - b0 - 1) / kThreads;
  const int4 r0 = reach[t0], r1 = reach[t1];
  lo = r0.x;
  hi = r1.y;
  tlo = r0.z;
  tn = r1.w - r0.z + 1;
}

// The table window [tlo, tlo + tn) of a block: its raw 16-bit entries from the 16-byte word that
// holds tlo (the rows of an event's table and each layer's 10^4 entries start on 16-byte words), as
// one 16-byte load per thread, staged when it spans at most kTab16Stage entries (3 KB of LDS), and
// the starts of its layers (up to 101, one load per thread), staged always. The window's first entry
// is in the layer of the block's first point (L0) and its last entry is in the layer after the last
// point's (L1, from thi = L1 * 10^4 + (last column + 1) * 100, reach_columns).
constexpr int kTab16Stage = 1536;
static_assert(kTab16Stage % 8 == 0 && kTab16Stage / 8 <= kThreads && k2dLayers + 1 <= kThreads,
              "one load per thread of the smallest block");
struct TabWindow {
  int base, n, L0, L1, nStarts;
  bool fits;
};
__device__ __forceinline__ TabWindow table_window(int tlo, int tn) {
  TabWindow w;
  w.base = tlo & ~7;
  w.n = tlo + tn - w.base;
  w.fits = w.n <= kTab16Stage;
  w.L0 = tlo / k2dBinsPerLayer;
  w.L1 = (tlo + tn - 2) / k2dBinsPerLayer;
  const int Lend = (tlo + tn - 1) / k2dBinsPerLayer + 1;  // one past the last entry's layer (> L1)
  w.nStarts = (Lend < k2dLayers ? Lend : k2dLayers) - w.L0 + 1;
  return w;
}
__device__ __forceinline__ void stage_table(const BinTable2D& T, const TabWindow& w, uint16_t* st, int* layerStart) {
  const int tid = (int)threadIdx.x;
  if (w.fits && tid < (w.n + 7) / 8)
    reinterpret_cast<uint4*>(st)[tid] = reinterpret_cast<const uint4*>(T.t16 + w.base)[tid];
  if (tid < w.nStarts) layerStart[tid] = T.segB[w.L0 + tid];
}
// The layer of the block's sorted position p (layers L0..L1, starts staged by stage_table): a point
// of layer L has start(L) <= p < start(L + 1) <= start(L'') for every L'' > L, so the number of
// starts of layers L0 + 1..L1 that are <= p is L - L0. Most blocks lie in one layer (no step).
__device__ __forceinline__ int layer_at(const TabWindow& w, const int* layerStart, int p) {
  int L = w.L0;
  for (int j = 1; j <= w.L1 - w.L0; ++j) L += p >= layerStart[j] ? 1 : 0;
  return L;
}

template <bool STAGED>
__device__ __forceinline__ float4 cand(const float4* __restrict__ g, const float4* st, int q, int lo) {
  if constexpr (STAGED)
    return st[q - lo];
  else
    return g[q];
}

// rho as an unsigned key in the fp32 order of rho: -0 is taken as +0, and NaN maps to 0, below every
// number (a point of NaN rho is higher than no other point). The nearest-higher step compares
// (rho key << 32 | point index) as one 64-bit integer, which is the reference's "higher" relation
// rho_j > rho_i || (rho_j == rho_i && j > i) (CLUEAlgoKernels.h:84-86) for every pair of points.
__device__ __forceinline__ uint32_t rho_key(float r) {
  uint32_t u = (uint32_t)__float_as_int(r);
  if (u == 0x80000000u) u = 0u;
  const uint32_t k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return r != r ? 0u : k;
}
__device__ __forceinline__ float rho_of_key(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return k == 0u ? __int_as_float(0x7fc00000) : __int_as_float((int)u);
}
// The query side of the relation: a point of NaN rho has no higher point at all.
__device__ __forceinline__ uint64_t query_key(uint32_t rk, int i) {
  return rk == 0u ? ~0ull : (((uint64_t)rk << 32) | (uint32_t)i);
}
__device__ __forceinline__ uint64_t cand_key(const float4& c) {
  return ((uint64_t)(uint32_t)__float_as_int(c.w) << 32) | (uint32_t)__float_as_int(c.z);
}

// Nearest-higher candidates {x, y, index bits, rho key bits}: x, y and the index from the tile-order
// {x, y, w, index} records, the rho key from the tile-order key array the density step writes.
struct NhSrc {
  const float4* xyw;
  const uint32_t* rhoKey;
  __device__ __forceinline__ float4 at(int q) const {
    const float4 a = xyw[q];
    return make_float4(a.x, a.y, a.w, __int_as_float((int)rhoKey[q]));
  }
};
template <bool STAGED>
__device__ __forceinline__ float4 cand(const NhSrc& g, const float4* st, int q, int lo) {
  if constexpr (STAGED)
    return st[q - lo];
  else
    return g.at(q);
}

// Copy of [lo, hi) (1 <= hi - lo <= STAGE points) into LDS for the nearest-higher candidates, all
// loads of a thread in flight together (indices past the end re-read the last point, not stored).
template <int NT, int STAGE>
__device__ __forceinline__ void stage_range_nh(const NhSrc& g, float4* st, int lo, int hi) {
  static_assert(STAGE % NT == 0, "staging splits evenly over the block");
  constexpr int kPer = STAGE / NT;
  const int n = hi - lo;
  float4 v[kPer];
  uint32_t r[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = u * NT + (int)threadIdx.x;
    const int q = lo + (k < n ? k : n - 1);
    v[u] = g.xyw[q];
    r[u] = g.rhoKey[q];
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = u * NT + (int)threadIdx.x;
    if (k < n) st[k] = make_float4(v[u].x, v[u].y, v[u].w, __int_as_float((int)r[u]));
  }
}

// The same for the density candidates: {x, y, 0.5f * w, index} (see DensityStaged).
template <int NT, int STAGE>
__device__ __forceinline__ void stage_range_half(const float4* __restrict__ g, float4* st, int lo, int hi) {
  static_assert(STAGE % NT == 0, "staging splits evenly over the block");
  constexpr int kPer = STAGE / NT;
  const int n = hi - lo;
  float4 v[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = u * NT + (int)threadIdx.x;
    v[u] = g[lo + (k < n ? k : n - 1)];
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = u * NT + (int)threadIdx.x;
    if (k < n) st[k] = make_float4(v[u].x, v[u].y, 0.5f * v[u].z, v[u].w);
  }
}

// Run k >= kMaxCols of a box: sorted positions [s, t) of column xlo + k, read from the table.
__device__ __forceinline__ void run_bounds(const BinTable2D& T, const Runs& r, int k, int& s, int& t) {
  const LayerTab2D LT = T.layer(r.L);
  const int j = (r.xlo + k) * k2dRows;
  s = LT[j + r.ylo];
  t = LT[j + r.yhi + 1];
}

// Visits the candidates of the register runs [s[k], t[k]), k < min(n, kMaxCols), in column order,
// calling f(q) once per candidate. One loop iteration per candidate of this lane, moving to the next
// non-empty run only when the current one ends: a wave then iterates the largest candidate count of
// its lanes, instead of the sum over columns of each column's largest run.
// Candidates are taken two at a time inside a run (the second masked off at an odd run end), so
// the end-of-run test and the loop branch are paid once per pair and two LDS reads are in flight.
static_assert(kMaxCols == 5, "walk_runs keeps the pending runs in five register pairs");
template <class F>
__device__ __forceinline__ void walk_runs(const Runs& r, F& f) {
  int pend = (r.n < kMaxCols ? r.n : kMaxCols) - 1;
  int q = r.s[0], t = r.t[0];
  int a1 = r.s[1], b1 = r.t[1], a2 = r.s[2], b2 = r.t[2], a3 = r.s[3], b3 = r.t[3], a4 = r.s[4], b4 = r.t[4];
  while (q >= t && pend > 0) {
    q = a1; t = b1; a1 = a2; b1 = b2; a2 = a3; b2 = b3; a3 = a4; b3 = b4; --pend;
  }
  while (q < t) {
    const bool two = q + 1 < t;
    f(q, two ? q + 1 : q, two);
    q += two ? 2 : 1;
    if (q >= t) {
      while (q >= t && pend > 0) {
        q = a1; t = b1; a1 = a2; b1 = b2; a2 = a3; b2 = b3; a3 = a4; b3 = b4; --pend;
      }
    }
  }
}

// The same walk over staged candidates, with less work per step. A run is packed into one word as
// LDS byte offsets from the start of the staging array (low 16 bits: first candidate, high 16 bits:
// one past the last; a stage holds at most 2048 points of 16 bytes, so both fit); 0 is the empty
// run. The runs are compacted once per point (empty ones dropped, order kept), so that a run switch
// is a single shift of the pending words; a step reads the candidates at qa and qa + 16 (the second
// masked off at an odd run end: the staging array has one spare element after its last), and the
// lane leaves the loop when the shift brings in an empty run.
constexpr int kStageMax = 2048;
struct StagedRuns {
  uint32_t r0, r1, r2, r3, r4;
};
__device__ __forceinline__ uint32_t pack_run(int s, int t, int lo) {
  return t > s ? ((uint32_t)(s - lo) << 4) | ((uint32_t)(t - lo) << 20) : 0u;
}
__device__ __forceinline__ StagedRuns pack_runs(const Runs& r, int lo) {
  uint32_t v[kMaxCols];
#pragma unroll
  for (int k = 0; k < kMaxCols; ++k) v[k] = k < r.n ? pack_run(r.s[k], r.t[k], lo) : 0u;
  // stable compaction: each pass moves every empty run one place towards the end
#pragma unroll
  for (int i = 0; i < kMaxCols - 1; ++i) {
#pragma unroll
    for (int j = 0; j < kMaxCols - 1 - i; ++j) {
      const bool sw = v[j] == 0u;
      const uint32_t a = v[j], b = v[j + 1];
      v[j] = sw ? b : a;
      v[j + 1] = sw ? a : b;
    }
  }
  return StagedRuns{v[0], v[1], v[2], v[3], v[4]};
}
template <class F>
__device__ __forceinline__ void walk_staged(const StagedRuns& R, F& f) {
  int qa = (int)(R.r0 & 0xFFFFu), ta = (int)(R.r0 >> 16);
  uint32_t p1 = R.r1, p2 = R.r2, p3 = R.r3, p4 = R.r4;
  while (qa < ta) {
    f(qa, qa + 16 < ta);
    qa += 32;
    if (qa >= ta) {
      qa = (int)(p1 & 0xFFFFu);
      ta = (int)(p1 >> 16);
      p1 = p2;
      p2 = p3;
      p3 = p4;
      p4 = 0u;
    }
  }
}
__device__ __forceinline__ float4 staged_at(const unsigned char* st, int qa) {
  return *reinterpret_cast<const float4*>(st + qa);
}

// Density candidates at staged byte offsets qa and (if two) qa + 16. The density staging holds
// {x, y, 0.5f * w, index}: the reference's weight factor (i == j ? 1.f : 0.5f) * w[j] is that
// product for every other point and w itself for the point (wself, at byte offset self).
struct DensityStaged {
  float xi, yi, dc2, wself;
  int self, selfBefore;  // the point's byte offset, and that minus 16
  const unsigned char* st;
  float rho;
  __device__ __forceinline__ void operator()(int qa, bool two) {
    const float4 c0 = staged_at(st, qa);
    const float4 c1 = staged_at(st, qa + 16);
    const float dx0 = xi - c0.x, dy0 = yi - c0.y;
    const float dx1 = xi - c1.x, dy1 = yi - c1.y;
    const float d20 = dx0 * dx0 + dy0 * dy0;
    const float d21 = dx1 * dx1 + dy1 * dy1;
    const float w0 = qa == self ? wself : c0.z;
    const float w1 = qa == selfBefore ? wself : c1.z;
    rho += d20 <= dc2 ? w0 : 0.f;
    rho += (two & (d21 <= dc2)) ? w1 : 0.f;
  }
};

// Candidates q0 and (if two) q1 = q0 + 1, in that order: their rho contributions
// (CLUEAlgoKernels.h:36-52); see density_run. A masked or failing term adds +0.f.
template <bool STAGED>
struct DensityStep {
  int p;
  float xi, yi, dc2;
  const float4* g;
  const float4* st;
  int lo;
  float rho;
  __device__ __forceinline__ void operator()(int q0, int q1, bool two) {
    const float4 c0 = cand<STAGED>(g, st, q0, lo);
    const float4 c1 = cand<STAGED>(g, st, q1, lo);
    const float dx0 = xi - c0.x, dy0 = yi - c0.y;
    const float dx1 = xi - c1.x, dy1 = yi - c1.y;
    const float d20 = dx0 * dx0 + dy0 * dy0;
    const float d21 = dx1 * dx1 + dy1 * dy1;
    const float w0 = (q0 == p ? 1.f : 0.5f) * c0.z;
    const float w1 = (q1 == p ? 1.f : 0.5f) * c1.z;
    rho += d20 <= dc2 ? w0 : 0.f;
    rho += (two & (d21 <= dc2)) ? w1 : 0.f;
  }
};

// rho of one point (kernel_calculate_density, CLUEAlgoKernels.h:25-55), summed in the serial order.
// A contribution that fails the distance test adds +0.f, which leaves rho unchanged: rho starts at
// +0 and a round-to-nearest sum started from +0 is never -0.
template <bool STAGED>
__device__ __forceinline__ float density_run(int p, float xi, float yi, int s, int t,
                                             const float4* __restrict__ xyw, const float4* st, int lo, float dc2,
                                             float rho) {
  int q = s;
  for (; q + 1 < t; q += 2) {
    const float4 c0 = cand<STAGED>(xyw, st, q, lo);
    const float4 c1 = cand<STAGED>(xyw, st, q + 1, lo);
    const float dx0 = xi - c0.x, dy0 = yi - c0.y;
    const float dx1 = xi - c1.x, dy1 = yi - c1.y;
    const float d20 = dx0 * dx0 + dy0 * dy0;
    const float d21 = dx1 * dx1 + dy1 * dy1;
    const float w0 = (q == p ? 1.f : 0.5f) * c0.z;
    const float w1 = (q + 1 == p ? 1.f : 0.5f) * c1.z;
    rho += d20 <= dc2 ? w0 : 0.f;
    rho += d21 <= dc2 ? w1 : 0.f;
  }
  if (q < t) {
    const float4 c0 = cand<STAGED>(xyw, st, q, lo);
    const float dx0 = xi - c0.x, dy0 = yi - c0.y;
    const float d20 = dx0 * dx0 + dy0 * dy0;
    const float w0 = (q == p ? 1.f : 0.5f) * c0.z;
    rho += d20 <= dc2 ? w0 : 0.f;
  }
  return rho;
}

// rho of every point: rhoKey[p] at its sorted position p (what the nearest-higher step reads) and
// rho itself, with RHO_OUT at the point's index, else (the index-pass output form) as lab[p].y, from
// where the output pass copies it to the index in index order.
template <int NT, int ITEMS, int STAGE, bool RHO_OUT>
__global__ __launch_bounds__(NT) CLUE_WAVES_PER_EU(kWavesPerEu2D<NT, ITEMS, STAGE>) void k_density2d(const BlockDesc* __restrict__ blocks,
                                                        const float4* __restrict__ xyw, BinTables2D tabs,
                                                        const int4* __restrict__ reach, float dc, float dc2,
                                                        uint32_t* __restrict__ rhoKey, float* __restrict__ rhoOut,
                                                        int2* __restrict__ lab) {
  using Shape = CompShape<NT, ITEMS>;
  static_assert(STAGE <= kStageMax, "staged byte offsets take 16 bits");
  __shared__ float4 s_st[STAGE + 1];  // + the spare element a masked second candidate may read
  __shared__ __align__(16) uint16_t s_tab[kTab16Stage];
  __shared__ int s_layerStart[k2dLayers + 1];
  // consecutive compute blocks (overlapping staged ranges) run on one XCD and share its L2
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const BlockDesc bd = blocks[blk / Shape::kPerSort];
  const int b0 = bd.begin + (blk % Shape::kPerSort) * Shape::kHits;
  const int b1 = (b0 + Shape::kHits) < bd.end ? (b0 + Shape::kHits) : bd.end;
  if (b0 >= b1) return;  // uniform over the block
  const BinTable2D T = tabs.event(bd.event);
  int rlo, rhi, tlo, tn;
  compute_reach<Shape>(reach, blk, b0, b1, rlo, rhi, tlo, tn);
  const bool staged = rhi - rlo <= STAGE;
  if (staged) stage_range_half<NT, STAGE>(xyw, s_st, rlo, rhi);
  const TabWindow win = table_window(tlo, tn);
  const TableView tab{T, s_tab, s_layerStart, win.base, win.L0, win.fits};
  stage_table(T, win, s_tab, s_layerStart);
  float4 me[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * NT + (int)threadIdx.x;
    me[r] = xyw[p < b1 ? p : b0];
  }
  __syncthreads();
  Runs runs[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * NT + (int)threadIdx.x;
    make_runs(tab, layer_at(win, s_layerStart, p < b1 ? p : b0), me[r].x, me[r].y, dc, runs[r]);
  }
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * NT + (int)threadIdx.x;
    if (p < b1) {
      const float xi = me[r].x, yi = me[r].y;
      float rho = 0.f;
      if (staged) {
        const int self = (p - rlo) * 16;
        DensityStaged f{xi, yi, dc2, me[r].z, self, self - 16, reinterpret_cast<const unsigned char*>(s_st), 0.f};
        walk_staged(pack_runs(runs[r], rlo), f);
        for (int k = kMaxCols; k < runs[r].n; ++k) {
          int s, t;
          run_bounds(T, runs[r], k, s, t);
          walk_staged(StagedRuns{pack_run(s, t, rlo), 0u, 0u, 0u, 0u}, f);
        }
        rho = f.rho;
      } else {
        DensityStep<false> f{p, xi, yi, dc2, xyw, s_st, rlo, 0.f};
        walk_runs(runs[r], f);
        rho = f.rho;
        for (int k = kMaxCols; k < runs[r].n; ++k) {
          int s, t;
          run_bounds(T, runs[r], k, s, t);
          rho = density_run<false>(p, xi, yi, s, t, xyw, s_st, rlo, dc2, rho);
        }
      }
      rhoKey[p] = rho_key(rho);
      if constexpr (RHO_OUT)
        rhoOut[__float_as_int(me[r].w)] = rho;
      else
        lab[p].y = __float_as_int(rho);
    }
  }
}

// Nearest higher (kernel_calculate_distanceToHigher, CLUEAlgoKernels.h:57-106): the reference keeps
// the first candidate, in serial order, with the strictly smallest sqrtf(d2) among the higher ones
// with sqrtf(d2) <= dm. The scan keeps the running minimum of d2 instead (strict improvements,
// so the first of equal d2 values) and the running minimum before it. Because sqrtf is monotone
// (non-decreasing), every candidate before the final minimum has d2 >= that previous minimum; so
// unless the previous minimum has the same sqrtf as the final one, no earlier candidate can tie
// with it and the final minimum is the reference's answer. In the tie case the scan is redone
// exactly as the reference does it.
// The running minimum starts at dm2up, the float after dm2, so that "d2 < best2" also enforces
// d2 <= dm2 until the first update and is implied by it afterwards; a previous minimum equal to
// dm2up never ties (sqrtf(dm2up) > dm >= sqrtf of any accepted d2).
template <bool STAGED>
__device__ __forceinline__ void nh_run(uint64_t qkey, float xi, float yi, int s, int t, const NhSrc& src,
                                       const float4* st, int lo, float& best2, int& best, float& prev2) {
  for (int q = s; q < t; ++q) {
    const float4 c = cand<STAGED>(src, st, q, lo);
    const float dx = xi - c.x;
    const float dy = yi - c.y;
    const float d2 = dx * dx + dy * dy;
    const bool upd = (cand_key(c) > qkey) & (d2 < best2);
    prev2 = upd ? best2 : prev2;
    best2 = upd ? d2 : best2;
    best = upd ? q : best;
  }
}

// Candidates q0 and (if two) q1 = q0 + 1, in that order, of the deferred nearest-higher scan; see
// nh_run. `best` is the sorted position of the running minimum.
template <bool STAGED>
struct NhStep {
  uint64_t qkey;
  float xi, yi;
  NhSrc g;
  const float4* st;
  int lo;
  float best2, prev2;
  int best;
  __device__ __forceinline__ void take(const float4& c, int q, bool valid) {
    const float dx = xi - c.x;
    const float dy = yi - c.y;
    const float d2 = dx * dx + dy * dy;
    // bitwise forms: lane masks combined without branches
    const bool upd = valid & (cand_key(c) > qkey) & (d2 < best2);
    prev2 = upd ? best2 : prev2;
    best2 = upd ? d2 : best2;
    best = upd ? q : best;
  }
  __device__ __forceinline__ void operator()(int q0, int q1, bool two) {
    const float4 c0 = cand<STAGED>(g, st, q0, lo);
    const float4 c1 = cand<STAGED>(g, st, q1, lo);
    take(c0, q0, true);
    take(c1, q1, two);
  }
};

// NhStep over staged candidates at byte offsets qa and (if two) qa + 16 (see walk_staged); `best` is
// the staged byte offset of the running minimum.
struct NhStaged {
  uint64_t qkey;
  float xi, yi;
  const unsigned char* st;
  float best2, prev2;
  int best;
  __device__ __forceinline__ void take(const float4& c, int qa, bool valid) {
    const float dx = xi - c.x;
    const float dy = yi - c.y;
    const float d2 = dx * dx + dy * dy;
    const bool upd = valid & (cand_key(c) > qkey) & (d2 < best2);
    prev2 = upd ? best2 : prev2;
    best2 = upd ? d2 : best2;
    best = upd ? qa : best;
  }
  __device__ __forceinline__ void operator()(int qa, bool two) {
    const float4 c0 = staged_at(st, qa);
    const float4 c1 = staged_at(st, qa + 16);
    take(c0, qa, true);
    take(c1, qa + 16, two);
  }
};

template <bool STAGED>
__device__ __forceinline__ void nh_run_exact(uint64_t qkey, float xi, float yi, int s, int t, const NhSrc& src,
                                             const float4* st, int lo, float dm2, float& delta, int& nh) {
  for (int q = s; q < t; ++q) {
    const float4 c = cand<STAGED>(src, st, q, lo);
    const float dx = xi - c.x;
    const float dy = yi - c.y;
    const float d2 = dx * dx + dy * dy;
    if (cand_key(c) > qkey && d2 <= dm2) {
      const float d = sqrtf(d2);
      if (d < delta) {
        delta = d;
        nh = q;
      }
    }
  }
}

// The nearest higher of a point as its sorted position (-1 if it has none), with delta.
template <bool STAGED>
__device__ __forceinline__ int nearest_higher_of(uint64_t qkey, float xi, float yi, const BinTable2D& T,
                                                 const Runs& runs, const NhSrc& src, const float4* st, int lo,
                                                 float dm2, float dm2up, float& delta) {
  float best2, prev2;
  int best;
  if constexpr (STAGED) {
    NhStaged f{qkey, xi, yi, reinterpret_cast<const unsigned char*>(st), dm2up, INFINITY, -1};
    walk_staged(pack_runs(runs, lo), f);
    for (int k = kMaxCols; k < runs.n; ++k) {
      int s, t;
      run_bounds(T, runs, k, s, t);
      walk_staged(StagedRuns{pack_run(s, t, lo), 0u, 0u, 0u, 0u}, f);
    }
    best2 = f.best2;
    prev2 = f.prev2;
    best = f.best >= 0 ? lo + (f.best >> 4) : -1;
  } else {
    NhStep<STAGED> f{qkey, xi, yi, src, st, lo, dm2up, INFINITY, -1};
    walk_runs(runs, f);
    best2 = f.best2;
    prev2 = f.prev2;
    best = f.best;
    for (int k = kMaxCols; k < runs.n; ++k) {
      int s, t;
      run_bounds(T, runs, k, s, t);
      nh_run<STAGED>(qkey, xi, yi, s, t, src, st, lo, best2, best, prev2);
    }
  }
  delta = FLT_MAX;
  if (best < 0) return -1;
  delta = sqrtf(best2);
  if (sqrtf(prev2) != delta) return best;
  delta = FLT_MAX;
  int nh = -1;
  for (int k = 0; k < runs.n; ++k) {
    int s, t;
    run_bounds(T, runs, k, s, t);
    nh_run_exact<STAGED>(qkey, xi, yi, s, t, src, st, lo, dm2, delta, nh);
  }
  return nh;
}

// delta / nearestHigher plus the seed/outlier decision of kernel_findAndAssign_clusters
// (CLUEAlgoKernels.h:108-137); each seed also sets its bit in the batch's seed words (cleared before
// the launch), from which the labels step ranks the seeds. Two output forms:
//  * INDEX_PASS = false: delta, nearestHigher and isSeed at the point's index, and `parent` (index
//    space: the nearest higher of a follower, the point itself for a root);
//  * INDEX_PASS = true: one 16-byte record per point at its index, rec[i] = {delta bits,
//    nearestHigher, isSeed, the point's sorted position}, and `parent` in tile order, written
//    coalesced: a follower's link is the sorted position of its nearest higher, a root's is ~i (its
//    own index, negative). The labels step then walks the links in tile order (k_labels_tile2d) and
//    writes every output in index order (k_outputs2d). One scattered 16-byte write per point instead
//    of four 4-byte ones: where hit indices are unrelated to position, each scattered write costs a
//    memory burst.
template <int NT, int ITEMS, int STAGE, bool INDEX_PASS>
__global__ __launch_bounds__(NT) CLUE_WAVES_PER_EU(kWavesPerEu2D<NT, ITEMS, STAGE>) void k_nh2d(const BlockDesc* __restrict__ blocks,
                                                   const float4* __restrict__ xyw, const uint32_t* __restrict__ rhoKey,
                                                   BinTables2D tabs,
                                                   const int4* __restrict__ reach, const int* __restrict__ evOff,
                                                   float dc, float rhoc,
                                                   float odf, float dm2, float dm2up, float* __restrict__ deltaOut,
                                                   int32_t* __restrict__ nhOut, int32_t* __restrict__ isSeedOut,
                                                   int* __restrict__ parent, int4* __restrict__ rec,
                                                   uint64_t* __restrict__ seedWord) {
  using Shape = CompShape<NT, ITEMS>;
  static_assert(STAGE <= kStageMax, "staged byte offsets take 16 bits");
  __shared__ float4 s_st[STAGE + 1];  // + the spare element a masked second candidate may read
  __shared__ __align__(16) uint16_t s_tab[kTab16Stage];
  __shared__ int s_layerStart[k2dLayers + 1];
  // consecutive compute blocks (overlapping staged ranges) run on one XCD and share its L2
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const BlockDesc bd = blocks[blk / Shape::kPerSort];
  const int b0 = bd.begin + (blk % Shape::kPerSort) * Shape::kHits;
  const int b1 = (b0 + Shape::kHits) < bd.end ? (b0 + Shape::kHits) : bd.end;
  if (b0 >= b1) return;
  const BinTable2D T = tabs.event(bd.event);
  const int evBase = evOff[bd.event];
  const float dm = odf * dc;
  const NhSrc src{xyw, rhoKey};
  int rlo, rhi, tlo, tn;
  compute_reach<Shape>(reach, blk, b0, b1, rlo, rhi, tlo, tn);
  const bool staged = rhi - rlo <= STAGE;
  if (staged) stage_range_nh<NT, STAGE>(src, s_st, rlo, rhi);
  const TabWindow win = table_window(tlo, tn);
  const TableView tab{T, s_tab, s_layerStart, win.base, win.L0, win.fits};
  stage_table(T, win, s_tab, s_layerStart);
  float4 me[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * NT + (int)threadIdx.x;
    me[r] = src.at(p < b1 ? p : b0);
  }
  __syncthreads();
  Runs runs[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * NT + (int)threadIdx.x;
    make_runs(tab, layer_at(win, s_layerStart, p < b1 ? p : b0), me[r].x, me[r].y, dm, runs[r]);
  }
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * NT + (int)threadIdx.x;
    if (p < b1) {
      const int i = __float_as_int(me[r].z);
      const uint32_t rk = (uint32_t)__float_as_int(me[r].w);
      const float rhoi = rho_of_key(rk);
      const uint64_t qkey = query_key(rk, i);
      float delta;
      const int nhp = staged ? nearest_higher_of<true>(qkey, me[r].x, me[r].y, T, runs[r], src, s_st, rlo, dm2, dm2up,
                                                       delta)
                             : nearest_higher_of<false>(qkey, me[r].x, me[r].y, T, runs[r], src, s_st, rlo, dm2,
                                                        dm2up, delta);
      // the nearest higher's point index, from the staged candidate or its sorted record
      const int nhg = nhp < 0 ? -1 : staged ? __float_as_int(s_st[nhp - rlo].z) : __float_as_int(xyw[nhp].w);
      const bool seed = (delta > dc) && (rhoi >= rhoc);
      const bool outlier = (delta > dm) && (rhoi < rhoc);
      const bool root = seed || outlier || nhg < 0;
      if (seed) atomicOr(reinterpret_cast<unsigned long long*>(seedWord + (i >> 6)), 1ull << (i & 63));
      // Followers point at their nearest higher; seeds, outliers (and a point without a nearest
      // higher, which the reference cannot register) are roots.
      if constexpr (INDEX_PASS) {
        rec[i] = make_int4(__float_as_int(delta), nhg >= 0 ? nhg - evBase : -1, seed ? 1 : 0, p);
        parent[p] = root ? ~i : nhp;
      } else {
        deltaOut[i] = delta;
        nhOut[i] = nhg >= 0 ? nhg - evBase : -1;
        isSeedOut[i] = seed ? 1 : 0;
        parent[i] = root ? i : nhg;
      }
    }
  }
}

// Labels of the index-pass output form, in tile order: each sorted position walks the links
// (parentS, see k_nh2d) until it reads a root's ~index; the lanes of a wave start at consecutive
// sorted positions, whose links lie in the same few search boxes, so their loads share cache lines.
// Writes lab[p].x = the label of p (lab[p].y holds its rho: the pair the output pass gathers with
// one load).
__global__ __launch_bounds__(kThreads) void k_labels_tile2d(const BlockDesc* __restrict__ blocks,
                                                            const int* __restrict__ parentS, SeedRank sr,
                                                            const int* __restrict__ evOff, int2* __restrict__ lab) {
  // the block's own links, staged: links inside the block's range are read from LDS
  __shared__ int s_par[kSortBlock];
  const BlockDesc bd = blocks[xcd_block((int)blockIdx.x, (int)gridDim.x)];
  const int rank0 = rank_of(sr, evOff[bd.event]);
  const int n = bd.end - bd.begin;
  int v[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int k = r * kThreads + (int)threadIdx.x;
    v[r] = k < n ? parentS[bd.begin + k] : -1;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int k = r * kThreads + (int)threadIdx.x;
    if (k < n) s_par[k] = v[r];
  }
  __syncthreads();
  // every link leads to a strictly higher point (or is a root's ~index), so each walk ends
  bool any = true;
  while (any) {
    any = false;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int k = v[r] - bd.begin;
      const bool local = (unsigned)k < (unsigned)n;
      int a = v[r], g = v[r];
      if (v[r] >= 0 && local) a = s_par[k];
      if (v[r] >= 0 && !local) g = parentS[v[r]];
      v[r] = local ? a : g;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) any |= v[r] >= 0;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int p = bd.begin + r * kThreads + (int)threadIdx.x;
    if (p < bd.end) {
      const int root = ~v[r];
      lab[p].x = seed_of(sr, root) ? rank_of(sr, root) - rank0 : -1;
    }
  }
}

// The outputs of the index-pass form in index order: rec[i] = {delta bits, nearestHigher, isSeed,
// sorted position p of i} (coalesced) and lab[p] (the one scattered load per point) give rho,
// delta, nearestHigher, isSeed and the label of point i, all written coalesced.
__global__ __launch_bounds__(kThreads) void k_outputs2d(const BlockDesc* __restrict__ blocks,
                                                        const int4* __restrict__ rec, const int2* __restrict__ lab,
                                                        float* __restrict__ rhoOut, float* __restrict__ deltaOut,
                                                        int32_t* __restrict__ nhOut, int32_t* __restrict__ isSeedOut,
                                                        int32_t* __restrict__ label) {
  const BlockDesc bd = blocks[xcd_block((int)blockIdx.x, (int)gridDim.x)];
  int4 v[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    v[r] = rec[i < bd.end ? i : bd.begin];
  }
  int2 l[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) l[r] = lab[v[r].w];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    if (i < bd.end) {
      rhoOut[i] = __int_as_float(l[r].y);
      deltaOut[i] = __int_as_float(v[r].x);
      nhOut[i] = v[r].y;
      isSeedOut[i] = v[r].z;
      label[i] = l[r].x;
    }
  }
}

}  // namespace g2d
}  // namespace clue
