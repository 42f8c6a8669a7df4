// CLUE2D on gfx950: tile build, density, nearest-higher (+ seed/outlier flags) and labels for a
// packed batch of events. A single event is the batch of one.
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
#include "clue2d.h"

#include <float.h>
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
  return p;
}

namespace {

struct Keys2D {
  const float* x;
  const float* y;
  const int32_t* layer;
  uint32_t* keyOut;
  int* err;
  __device__ __forceinline__ uint32_t operator()(int i) const {
    int L = layer[i];
    if (L < 0 || L >= k2dLayers) {
      atomicOr(err, kErrLayerRange);
      L = L < 0 ? 0 : k2dLayers - 1;
    }
    const uint32_t k = (uint32_t)(L * k2dBinsPerLayer + xbin2d(x[i]) * k2dRows + ybin2d(y[i]));
    keyOut[i] = k;
    return k;
  }
};

struct EmitFirst {
  uint32_t* kOut;
  int* iOut;
  __device__ __forceinline__ void operator()(int pos, int i, uint32_t k) const {
    kOut[pos] = k;
    iOut[pos] = i;
  }
};

struct EmitMid {
  const int* iIn;
  uint32_t* kOut;
  int* iOut;
  __device__ __forceinline__ void operator()(int pos, int i, uint32_t k) const {
    kOut[pos] = k;
    iOut[pos] = iIn[i];
  }
};

struct EmitLast2D {
  const int* iIn;
  uint32_t* kOut;
  int* iOut;
  const float* x;
  const float* y;
  const float* w;
  float4* xyw;
  __device__ __forceinline__ void operator()(int pos, int i, uint32_t k) const {
    const int idx = iIn[i];
    kOut[pos] = k;
    iOut[pos] = idx;
    xyw[pos] = make_float4(x[idx], y[idx], w[idx], 0.f);
  }
};

using DigitY = Digit<1, k2dRows>;
using DigitX = Digit<k2dRows, k2dCols>;
using DigitL = Digit<k2dBinsPerLayer, kRadix>;

// The density and nearest-higher kernels give each block kCompHits consecutive sorted points of one
// event. In x-major tile order everything their search boxes can reach is one contiguous range of
// the sorted array: from the first bin of the lowest box to the end of the highest box. When that
// range has at most kStage points it is staged into LDS once and every candidate is read from
// there; otherwise candidates are read from global memory. Either way the candidates are visited
// in the serial order, so the results do not depend on which path ran.
constexpr int kCompItems = 4;
constexpr int kCompHits = kThreads * kCompItems;  // points per block (two blocks per sort block)
constexpr int kStage = 2048;                      // staged points (32 KB of float4)

static_assert(kSortBlock == 2 * kCompHits, "compute blocks split each sort block in two");

struct Box {
  int L, xlo, xhi, ylo, yhi;
};

__device__ __forceinline__ Box box2d(uint32_t key, float xi, float yi, float d) {
  Box b;
  b.L = (int)(key / (uint32_t)k2dBinsPerLayer);
  b.xlo = xbin2d(xi - d);
  b.xhi = xbin2d(xi + d);
  b.ylo = ybin2d(yi - d);
  b.yhi = ybin2d(yi + d);
  return b;
}

// Block-wide [min lo, max hi) of the sorted positions the block's boxes reach.
__device__ __forceinline__ void block_range(int lo, int hi, int* s_red, int& outLo, int& outHi) {
  const int lane = lane_id(), wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const int ol = __shfl_xor(lo, off, kWave);
    const int oh = __shfl_xor(hi, off, kWave);
    lo = ol < lo ? ol : lo;
    hi = oh > hi ? oh : hi;
  }
  if (lane == 0) {
    s_red[wave] = lo;
    s_red[kWaves + wave] = hi;
  }
  __syncthreads();
  outLo = s_red[0];
  outHi = s_red[kWaves];
#pragma unroll
  for (int w = 1; w < kWaves; ++w) {
    outLo = s_red[w] < outLo ? s_red[w] : outLo;
    outHi = s_red[kWaves + w] > outHi ? s_red[kWaves + w] : outHi;
  }
}

template <bool STAGED>
__device__ __forceinline__ float4 cand(const float4* __restrict__ g, const float4* st, int q, int lo) {
  if constexpr (STAGED)
    return st[q - lo];
  else
    return g[q];
}

// rho of one point (kernel_calculate_density, CLUEAlgoKernels.h:25-55).
template <bool STAGED>
__device__ __forceinline__ float density_of(int p, float xi, float yi, const Box& b, const int* __restrict__ T,
                                            const float4* __restrict__ xyw, const float4* st, int lo, float dc2) {
  const int* TL = T + b.L * k2dBinsPerLayer;
  float rho = 0.f;
  for (int xb = b.xlo; xb <= b.xhi; ++xb) {
    const int s = TL[xb * k2dRows + b.ylo];
    const int t = TL[xb * k2dRows + b.yhi + 1];
    for (int q = s; q < t; ++q) {
      const float4 c = cand<STAGED>(xyw, st, q, lo);
      const float dx = xi - c.x;
      const float dy = yi - c.y;
      // sqrtf(dx*dx + dy*dy) <= dc, decided on the squared distance (same outcome, see dc2)
      if (dx * dx + dy * dy <= dc2) rho += (q == p ? 1.f : 0.5f) * c.z;
    }
  }
  return rho;
}

__global__ __launch_bounds__(kThreads) void k_density2d(const BlockDesc* __restrict__ blocks,
                                                        const uint32_t* __restrict__ keyS,
                                                        const float4* __restrict__ xyw, const int* __restrict__ perm,
                                                        const int* __restrict__ table, int64_t stride, float dc,
                                                        float dc2, float4* __restrict__ xyrp,
                                                        float* __restrict__ rhoOut) {
  __shared__ float4 s_st[kStage];
  __shared__ int s_red[2 * kWaves];
  const BlockDesc bd = blocks[blockIdx.x >> 1];
  const int b0 = bd.begin + (int)(blockIdx.x & 1) * kCompHits;
  const int b1 = (b0 + kCompHits) < bd.end ? (b0 + kCompHits) : bd.end;
  if (b0 >= b1) return;  // uniform over the block
  const int* T = table + (int64_t)bd.event * stride;
  int lo = 0x7fffffff, hi = -1;
#pragma unroll
  for (int r = 0; r < kCompItems; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p < b1) {
      const float4 c = xyw[p];
      const Box b = box2d(keyS[p], c.x, c.y, dc);
      const int* TL = T + b.L * k2dBinsPerLayer;
      const int l = TL[b.xlo * k2dRows + b.ylo], h = TL[b.xhi * k2dRows + b.yhi + 1];
      lo = l < lo ? l : lo;
      hi = h > hi ? h : hi;
    }
  }
  int rlo, rhi;
  block_range(lo, hi, s_red, rlo, rhi);
  const bool staged = rhi - rlo <= kStage;
  if (staged)
    for (int k = rlo + (int)threadIdx.x; k < rhi; k += kThreads) s_st[k - rlo] = xyw[k];
  __syncthreads();
#pragma unroll 1
  for (int r = 0; r < kCompItems; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p >= b1) break;
    const float4 c = staged ? s_st[p - rlo] : xyw[p];
    const Box b = box2d(keyS[p], c.x, c.y, dc);
    const float rho = staged ? density_of<true>(p, c.x, c.y, b, T, xyw, s_st, rlo, dc2)
                             : density_of<false>(p, c.x, c.y, b, T, xyw, s_st, rlo, dc2);
    const int i = perm[p];
    xyrp[p] = make_float4(c.x, c.y, rho, __int_as_float(i));
    rhoOut[i] = rho;
  }
}

// delta and nearest higher of one point (kernel_calculate_distanceToHigher, CLUEAlgoKernels.h:57-106).
template <bool STAGED>
__device__ __forceinline__ int nearest_higher_of(int i, float xi, float yi, float rhoi, const Box& b,
                                                 const int* __restrict__ T, const float4* __restrict__ xyrp,
                                                 const float4* st, int lo, float dm2, float& delta) {
  const int* TL = T + b.L * k2dBinsPerLayer;
  delta = FLT_MAX;
  int nh = -1;
  for (int xb = b.xlo; xb <= b.xhi; ++xb) {
    const int s = TL[xb * k2dRows + b.ylo];
    const int t = TL[xb * k2dRows + b.yhi + 1];
    for (int q = s; q < t; ++q) {
      const float4 c = cand<STAGED>(xyrp, st, q, lo);
      const float dx = xi - c.x;
      const float dy = yi - c.y;
      const float d2 = dx * dx + dy * dy;
      const int j = __float_as_int(c.w);
      // higher && sqrtf(d2) <= dm, then the strict "first minimum wins" test on sqrtf(d2)
      if (d2 <= dm2 && ((c.z > rhoi) || ((c.z == rhoi) && (j > i)))) {
        const float d = sqrtf(d2);
        if (d < delta) {
          delta = d;
          nh = j;
        }
      }
    }
  }
  return nh;
}

// delta / nearestHigher plus the seed/outlier decision of kernel_findAndAssign_clusters
// (CLUEAlgoKernels.h:108-137).
__global__ __launch_bounds__(kThreads) void k_nh2d(const BlockDesc* __restrict__ blocks,
                                                   const uint32_t* __restrict__ keyS,
                                                   const float4* __restrict__ xyrp, const int* __restrict__ table,
                                                   int64_t stride, const int* __restrict__ evOff, float dc, float rhoc,
                                                   float odf, float dm2, float* __restrict__ deltaOut,
                                                   int32_t* __restrict__ nhOut, int32_t* __restrict__ isSeedOut,
                                                   int* __restrict__ parent) {
  __shared__ float4 s_st[kStage];
  __shared__ int s_red[2 * kWaves];
  const BlockDesc bd = blocks[blockIdx.x >> 1];
  const int b0 = bd.begin + (int)(blockIdx.x & 1) * kCompHits;
  const int b1 = (b0 + kCompHits) < bd.end ? (b0 + kCompHits) : bd.end;
  if (b0 >= b1) return;
  const int* T = table + (int64_t)bd.event * stride;
  const int evBase = evOff[bd.event];
  const float dm = odf * dc;
  int lo = 0x7fffffff, hi = -1;
#pragma unroll
  for (int r = 0; r < kCompItems; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p < b1) {
      const float4 c = xyrp[p];
      const Box b = box2d(keyS[p], c.x, c.y, dm);
      const int* TL = T + b.L * k2dBinsPerLayer;
      const int l = TL[b.xlo * k2dRows + b.ylo], h = TL[b.xhi * k2dRows + b.yhi + 1];
      lo = l < lo ? l : lo;
      hi = h > hi ? h : hi;
    }
  }
  int rlo, rhi;
  block_range(lo, hi, s_red, rlo, rhi);
  const bool staged = rhi - rlo <= kStage;
  if (staged)
    for (int k = rlo + (int)threadIdx.x; k < rhi; k += kThreads) s_st[k - rlo] = xyrp[k];
  __syncthreads();
#pragma unroll 1
  for (int r = 0; r < kCompItems; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p >= b1) break;
    const float4 c = staged ? s_st[p - rlo] : xyrp[p];
    const int i = __float_as_int(c.w);
    const float rhoi = c.z;
    const Box b = box2d(keyS[p], c.x, c.y, dm);
    float delta;
    const int nhg = staged ? nearest_higher_of<true>(i, c.x, c.y, rhoi, b, T, xyrp, s_st, rlo, dm2, delta)
                           : nearest_higher_of<false>(i, c.x, c.y, rhoi, b, T, xyrp, s_st, rlo, dm2, delta);
    deltaOut[i] = delta;
    nhOut[i] = nhg >= 0 ? nhg - evBase : -1;
    const bool seed = (delta > dc) && (rhoi >= rhoc);
    const bool outlier = (delta > dm) && (rhoi < rhoc);
    isSeedOut[i] = seed ? 1 : 0;
    // Followers point at their nearest higher; seeds, outliers (and a point without a nearest
    // higher, which the reference cannot register) are roots.
    parent[i] = (seed || outlier || nhg < 0) ? i : nhg;
  }
}

}  // namespace

void clue2d_enqueue(const Clue2DBuffers& b, const Clue2DIO& io, const Clue2DParams& prm, hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  const int nHist = kRadix * b.nBlocks;
  // pass 1: yBin (keys computed and recorded in k2)
  Keys2D keys{io.x, io.y, io.layer, b.k2, b.err};
  k_radix_hist<Keys2D, DigitY><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, keys, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<KeyArray, DigitY, EmitFirst>
      <<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k2}, b.hist, EmitFirst{b.k1, b.i1});
  // pass 2: xBin
  k_radix_hist<KeyArray, DigitX><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k1}, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<KeyArray, DigitX, EmitMid>
      <<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k1}, b.hist, EmitMid{b.i1, b.k2, b.i2});
  // pass 3: layer; gathers the SoA into tile order
  k_radix_hist<KeyArray, DigitL><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k2}, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<KeyArray, DigitL, EmitLast2D><<<b.nBlocks, kThreads, 0, s>>>(
      b.blocks, KeyArray{b.k2}, b.hist,
      EmitLast2D{b.i2, b.k1, b.i1, io.x, io.y, io.weight, b.xyw});
  const uint32_t* keyS = b.k1;
  const int* perm = b.i1;
  build_table(keyS, b.blocks, b.nBlocks, b.evOff, b.E, k2dKeys, b.tableStride, b.table, b.tileStart, s);
  k_density2d<<<2 * b.nBlocks, kThreads, 0, s>>>(b.blocks, keyS, b.xyw, perm, b.table, b.tableStride, prm.dc,
                                                 prm.dc2, b.xyrp, io.rho);
  int* parent = reinterpret_cast<int*>(b.k2);
  int* rankAll = b.i2;
  k_nh2d<<<2 * b.nBlocks, kThreads, 0, s>>>(b.blocks, keyS, b.xyrp, b.table, b.tableStride, b.evOff, prm.dc,
                                            prm.rhoc, prm.outlierDeltaFactor, prm.dm2, io.delta, io.nearestHigher,
                                            io.isSeed, parent);
  scan_exclusive(io.isSeed, rankAll, b.N, b.part, s);
  launch_labels(b.blocks, b.nBlocks, parent, io.isSeed, rankAll, b.evOff, io.clusterIndex, s);
}

size_t clue2d_workspace_bytes(int64_t N, int nEvents, int nBlocks) {
  const size_t n = (size_t)(N > 0 ? N : 1);
  size_t bytes = 0;
  auto add = [&](size_t sz) { bytes += (sz + 255) & ~(size_t)255; };
  for (int k = 0; k < 4; ++k) add(n * 4);                   // k1 i1 k2 i2
  for (int k = 0; k < 2; ++k) add(n * 16);                  // xyw xyrp
  add((size_t)kRadix * (size_t)(nBlocks > 0 ? nBlocks : 1) * 4);  // hist
  const int64_t maxScan = (int64_t)kRadix * nBlocks > N ? (int64_t)kRadix * nBlocks : N;
  add((size_t)scan_partials(maxScan > 0 ? maxScan : 1) * 4);  // partials
  add((size_t)clue2d_table_stride() * (size_t)(nEvents > 0 ? nEvents : 1) * 4);
  add((size_t)(table_tiles(k2dKeys) + 1) * (size_t)(nEvents > 0 ? nEvents : 1) * 4);  // tileStart
  add((size_t)(nEvents + 1) * 4);                                   // evOff
  add((size_t)(nBlocks > 0 ? nBlocks : 1) * sizeof(BlockDesc));     // blocks
  add(256);                                                         // error flags
  return bytes;
}

}  // namespace clue
