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

namespace clue {

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
  float* xs;
  float* ys;
  float* ws;
  __device__ __forceinline__ void operator()(int pos, int i, uint32_t k) const {
    const int idx = iIn[i];
    kOut[pos] = k;
    iOut[pos] = idx;
    xs[pos] = x[idx];
    ys[pos] = y[idx];
    ws[pos] = w[idx];
  }
};

using DigitY = Digit<1, k2dRows>;
using DigitX = Digit<k2dRows, k2dCols>;
using DigitL = Digit<k2dBinsPerLayer, kRadix>;

// rho (kernel_calculate_density, CLUEAlgoKernels.h:25-55), one thread per sorted point.
__global__ __launch_bounds__(kThreads) void k_density2d(const BlockDesc* __restrict__ blocks,
                                                        const uint32_t* __restrict__ keyS,
                                                        const float* __restrict__ xs, const float* __restrict__ ys,
                                                        const float* __restrict__ ws, const int* __restrict__ perm,
                                                        const int* __restrict__ table, int64_t stride, float dc,
                                                        float* __restrict__ rhoS, float* __restrict__ rhoOut) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int* T = table + (int64_t)bd.event * stride;
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int p = bd.begin + r * kThreads + (int)threadIdx.x;
    if (p >= bd.end) break;
    const uint32_t key = keyS[p];
    const int L = (int)(key / (uint32_t)k2dBinsPerLayer);
    const float xi = xs[p], yi = ys[p];
    const int xlo = xbin2d(xi - dc), xhi = xbin2d(xi + dc);
    const int ylo = ybin2d(yi - dc), yhi = ybin2d(yi + dc);
    const int* TL = T + L * k2dBinsPerLayer;
    float rho = 0.f;
    for (int xb = xlo; xb <= xhi; ++xb) {
      const int s = TL[xb * k2dRows + ylo];
      const int t = TL[xb * k2dRows + yhi + 1];
      for (int q = s; q < t; ++q) {
        const float dx = xi - xs[q];
        const float dy = yi - ys[q];
        const float d = sqrtf(dx * dx + dy * dy);
        if (d <= dc) rho += (q == p ? 1.f : 0.5f) * ws[q];
      }
    }
    rhoS[p] = rho;
    rhoOut[perm[p]] = rho;
  }
}

// delta / nearestHigher (kernel_calculate_distanceToHigher, CLUEAlgoKernels.h:57-106) plus the
// seed/outlier decision of kernel_findAndAssign_clusters (:108-137), one thread per sorted point.
__global__ __launch_bounds__(kThreads) void k_nh2d(const BlockDesc* __restrict__ blocks,
                                                   const uint32_t* __restrict__ keyS, const float* __restrict__ xs,
                                                   const float* __restrict__ ys, const float* __restrict__ rhoS,
                                                   const int* __restrict__ perm, const int* __restrict__ table,
                                                   int64_t stride, const int* __restrict__ evOff, float dc, float rhoc,
                                                   float odf, float* __restrict__ deltaOut, int32_t* __restrict__ nhOut,
                                                   int32_t* __restrict__ isSeedOut, int* __restrict__ parent) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int* T = table + (int64_t)bd.event * stride;
  const int evBase = evOff[bd.event];
  const float dm = odf * dc;
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int p = bd.begin + r * kThreads + (int)threadIdx.x;
    if (p >= bd.end) break;
    const uint32_t key = keyS[p];
    const int L = (int)(key / (uint32_t)k2dBinsPerLayer);
    const float xi = xs[p], yi = ys[p], rhoi = rhoS[p];
    const int i = perm[p];
    const int xlo = xbin2d(xi - dm), xhi = xbin2d(xi + dm);
    const int ylo = ybin2d(yi - dm), yhi = ybin2d(yi + dm);
    const int* TL = T + L * k2dBinsPerLayer;
    float delta = FLT_MAX;
    int nhq = -1;
    for (int xb = xlo; xb <= xhi; ++xb) {
      const int s = TL[xb * k2dRows + ylo];
      const int t = TL[xb * k2dRows + yhi + 1];
      for (int q = s; q < t; ++q) {
        const float rq = rhoS[q];
        const bool higher = (rq > rhoi) || ((rq == rhoi) && (perm[q] > i));
        const float dx = xi - xs[q];
        const float dy = yi - ys[q];
        const float d = sqrtf(dx * dx + dy * dy);
        if (higher && d <= dm && d < delta) {
          delta = d;
          nhq = q;
        }
      }
    }
    const int nhg = nhq >= 0 ? perm[nhq] : -1;
    deltaOut[i] = delta;
    nhOut[i] = nhg >= 0 ? nhg - evBase : -1;
    const bool seed = (delta > dc) && (rhoi >= rhoc);
    const bool outlier = (delta > odf * dc) && (rhoi < rhoc);
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
      EmitLast2D{b.i2, b.k1, b.i1, io.x, io.y, io.weight, b.xs, b.ys, b.ws});
  const uint32_t* keyS = b.k1;
  const int* perm = b.i1;
  build_table(keyS, b.evOff, b.E, k2dKeys, b.tableStride, b.table, s);
  k_density2d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, keyS, b.xs, b.ys, b.ws, perm, b.table, b.tableStride, prm.dc,
                                             b.rhoS, io.rho);
  int* parent = reinterpret_cast<int*>(b.k2);
  int* rankAll = b.i2;
  k_nh2d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, keyS, b.xs, b.ys, b.rhoS, perm, b.table, b.tableStride, b.evOff,
                                        prm.dc, prm.rhoc, prm.outlierDeltaFactor, io.delta, io.nearestHigher,
                                        io.isSeed, parent);
  scan_exclusive(io.isSeed, rankAll, b.N, b.part, s);
  launch_labels(b.blocks, b.nBlocks, parent, io.isSeed, rankAll, b.evOff, io.clusterIndex, s);
}

size_t clue2d_workspace_bytes(int64_t N, int nEvents, int nBlocks) {
  const size_t n = (size_t)(N > 0 ? N : 1);
  size_t bytes = 0;
  auto add = [&](size_t sz) { bytes += (sz + 255) & ~(size_t)255; };
  for (int k = 0; k < 8; ++k) add(n * 4);                   // k1 i1 k2 i2 xs ys ws rhoS
  add((size_t)kRadix * (size_t)(nBlocks > 0 ? nBlocks : 1) * 4);  // hist
  const int64_t maxScan = (int64_t)kRadix * nBlocks > N ? (int64_t)kRadix * nBlocks : N;
  add((size_t)scan_partials(maxScan > 0 ? maxScan : 1) * 4);  // partials
  add((size_t)clue2d_table_stride() * (size_t)(nEvents > 0 ? nEvents : 1) * 4);
  add((size_t)(nEvents + 1) * 4);                                   // evOff
  add((size_t)(nBlocks > 0 ? nBlocks : 1) * sizeof(BlockDesc));     // blocks
  add(256);                                                         // error flags
  return bytes;
}

}  // namespace clue
