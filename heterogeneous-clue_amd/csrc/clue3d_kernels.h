// Device code of the CLUE3D path: key and payload functors of the three radix passes and the
// density / nearest-higher kernels. Header-only so that the kernels also build for the CPU
// emulation of the HIP execution model used by the tests; the launch sequence is in clue3d.hip.
//
// Reference semantics: src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:19-24 (tiles),
// :26-150 (density), :286-392 (nearest higher), :509-580 (seeds and propagation).
// Serial scan order: sibling layer ascending, eta bin ascending, phi bin in window order
// (pb-w .. pb+w, wrapped modulo 126), SoA index ascending. With keys (layer, eta, phi) sorted and
// bins in index order, each (layer, eta row) window is at most two contiguous runs of the sorted
// array, visited in that order, so the fp32 sums and the "last minimum wins" updates see
// candidates in exactly the serial order.
#pragma once

#include <float.h>
#include <limits.h>
#include <math.h>

#include "clue3d.h"

namespace clue {
namespace g3d {

__device__ __forceinline__ int clamp_layer3d(int L) { return L < 0 ? 0 : (L >= k3dLayers ? k3dLayers - 1 : L); }
__device__ __forceinline__ int clamp_phibin(int p) { return p < 0 ? 0 : (p > k3dPhi ? k3dPhi : p); }

// Bin key of cluster i of event ev; range errors set bits in err[ev * errStride] (errStride 0: one
// flag word for the whole batch).
struct Keys3D {
  const float* eta;
  const float* phi;
  const int32_t* layer;
  uint32_t* keyOut;
  int* err;
  int errStride;
  __device__ __forceinline__ uint32_t operator()(int i, int ev) const {
    int L = layer[i];
    if (L < 0 || L >= k3dLayers) {
      atomicOr(err + ev * errStride, kErrLayerRange);
      L = clamp_layer3d(L);
    }
    int pb = phi_bin3d(phi[i]);
    if (pb < 0 || pb > k3dPhi) {
      atomicOr(err + ev * errStride, kErrBinRange);
      pb = clamp_phibin(pb);
    }
    // TICLLayerTileT::globalBin = phiBin + etaBin * 126 (TICLLayerTile.h:52); a phiBin of 126 is
    // the first bin of the next row there too.
    const uint32_t k = (uint32_t)(L * k3dBinsPerLayer + eta_bin3d(eta[i]) * k3dPhi + pb);
    keyOut[i] = k;
    return k;
  }
};

// Last pass: also gathers {r_over_absz, phi, energy, z} of every point into tile order.
struct EmitLast3D {
  const int* iIn;
  uint32_t* kOut;
  int* iOut;
  const float* roz;
  const float* phi;
  const float* en;
  const float* z;
  float4* cre;
  // the cluster index passes through LDS; the four fields are gathered once it is in sorted order
  using Payload = int;
  struct Out {
    int idx;
    float4 v;
  };
  __device__ __forceinline__ void block_starts(const BlockDesc&, const int*) const {}
  __device__ __forceinline__ Payload load(int i) const { return iIn[i]; }
  __device__ __forceinline__ Out fetch(int idx) const {
    return Out{idx, make_float4(roz[idx], phi[idx], en[idx], z[idx])};
  }
  __device__ __forceinline__ void store(int pos, uint32_t k, const Out& o) const {
    kOut[pos] = k;
    iOut[pos] = o.idx;
    cre[pos] = o.v;
  }
};

using DigitPhi = Digit<1, k3dPhi>;
using DigitEta = Digit<k3dPhi, k3dEta>;
using DigitLay = Digit<k3dBinsPerLayer, kRadix>;

// ---------------------------------------------------------------- search windows

// The window of one point: sibling layers [minL, maxL], eta rows [emin, emax] (emax may be 34: that
// row is the next layer's first row, as in the reference's contiguous tile array), and per row the
// phi bins a0..b0 then a1..b1 (a1 > b1: no second run) - the reference's iteration
// ((it % 126) + 126) % 126 over pb-W .. pb+W split into ascending runs.
struct Win3D {
  int L, minL, maxL, emin, emax;
  int a0, b0, a1, b1;
};

// sideLo: the lowest layer the second detector side lets the window reach (47 for the density,
// 48 for the nearest higher: CLUE3DAlgoKernels.h:56-59 and :312-318).
template <int W>
__device__ __forceinline__ Win3D make_win(int L, int eb, int pb, int S, int sideLo) {
  Win3D w;
  w.L = L;
  if (L < k3dLastPerSide) {
    w.minL = L - S > 0 ? L - S : 0;
    w.maxL = L + S < k3dLastPerSide - 1 ? L + S : k3dLastPerSide - 1;
  } else {
    w.minL = L - S > sideLo ? L - S : sideLo;
    w.maxL = L + S < 2 * k3dLastPerSide - 1 ? L + S : 2 * k3dLastPerSide - 1;
  }
  w.emin = eb - W > 0 ? eb - W : 0;
  w.emax = eb + W < k3dEta ? eb + W : k3dEta;
  const int lo = pb - W, hi = pb + W;
  if (lo < 0) {
    w.a0 = lo + k3dPhi;
    w.b0 = k3dPhi - 1;
    w.a1 = 0;
    w.b1 = hi;
  } else if (hi > k3dPhi - 1) {
    w.a0 = lo;
    w.b0 = k3dPhi - 1;
    w.a1 = 0;
    w.b1 = hi - k3dPhi;
  } else {
    w.a0 = lo;
    w.b0 = hi;
    w.a1 = 0;
    w.b1 = -1;
  }
  return w;
}

__device__ __forceinline__ bool win_wraps(const Win3D& w) { return w.a1 <= w.b1; }

// First and one-past-last sorted positions the window reaches (lo > hi: no sibling layer).
__device__ __forceinline__ void win_reach(const int* __restrict__ T, const Win3D& w, int& lo, int& hi) {
  if (w.minL > w.maxL) {
    lo = INT_MAX;
    hi = -1;
    return;
  }
  const bool wrap = win_wraps(w);
  lo = T[w.minL * k3dBinsPerLayer + w.emin * k3dPhi + (wrap ? 0 : w.a0)];
  hi = T[w.maxL * k3dBinsPerLayer + w.emax * k3dPhi + (wrap ? k3dPhi : w.b0 + 1)];
}

// reco::reduceRange (Math/deltaPhi.h:17-24) returns its argument unchanged when |x| <= pi (the
// float nearest M_PI). Candidates of one window differ in phi by less than that unless the window
// wraps or a phi lies outside [-pi, pi], so the candidate steps test both differences of a pair and
// take the reduction only when one of them needs it.
constexpr float kPiF = (float)3.14159265358979323846;
__device__ __forceinline__ bool phi_wide(float x) { return !(fabsf(x) <= kPiF); }
__device__ __forceinline__ float reduce_wide(float x) {  // reduce_range for |x| > pi or NaN
  const float o2pi = (float)(1. / (2. * 3.14159265358979323846));
  const float n = roundf(x * o2pi);
  return x - n * (float)(2. * 3.14159265358979323846);
}
__device__ __forceinline__ void reduce_pair(float& dp0, float& dp1) {
  const bool w0 = phi_wide(dp0), w1 = phi_wide(dp1);
  if (__builtin_expect(w0 | w1, 0)) {
    dp0 = w0 ? reduce_wide(dp0) : dp0;
    dp1 = w1 ? reduce_wide(dp1) : dp1;
  }
}

// ---------------------------------------------------------------- kernel shapes

// The density and nearest-higher kernels give each block ITEMS * kThreads consecutive sorted points
// of one event. Everything their windows can reach is one contiguous range of the sorted array;
// when it has at most STAGE points it is staged into LDS and each lane lists the non-empty runs of
// its window in LDS, then walks them flattened: one loop iteration per candidate pair, moving to the
// next run only when the current one ends, so a wave iterates its lanes' largest candidate count
// instead of the sum over runs of each run's largest length. Otherwise (or for more than 31
// sibling layers) the window is scanned row by row from global memory. Both visit the candidates in
// the serial order, so the results do not depend on ITEMS, STAGE or the path taken.
template <int ITEMS>
struct CompShape3D {
  static constexpr int kHits = kThreads * ITEMS;
  static constexpr int kPerSort = kSortBlock / kHits;
  static_assert(kSortBlock % kHits == 0, "compute blocks tile each sort block");
};

struct Variant3D {
  int items;  // points per thread
  int stage;  // staged points
};
constexpr int kNumVariants3D = 4;
constexpr Variant3D kVariants3D[kNumVariants3D] = {{2, 2048}, {1, 2048}, {4, 2048}, {1, 1024}};

constexpr int kRunCap = 16;      // run-list entries per lane and round
constexpr int kOffBits = 13;     // staged offsets (<= STAGE <= 4096) per entry field
constexpr int kMaxListS = 31;    // sibling-layer reach the 6-bit layer field of an entry covers
constexpr uint32_t kOffMask = (1u << kOffBits) - 1u;

__device__ __forceinline__ uint32_t run_entry(int s, int t, int layerOff) {
  return (uint32_t)s | ((uint32_t)t << kOffBits) | ((uint32_t)layerOff << (2 * kOffBits));
}

// Appends the non-empty runs of layer cl of the window, in serial order, to a lane's list (entries
// kThreads apart, positions relative to the staged range start rlo). All table reads of the layer
// are issued before any is used.
template <int W>
__device__ __forceinline__ int list_layer(const int* __restrict__ T, const Win3D& w, int cl, int rlo, uint32_t* list,
                                          int cnt) {
  constexpr int kRows = 2 * W + 1;
  const int* TL = T + cl * k3dBinsPerLayer;
  int s0[kRows], t0[kRows], s1[kRows], t1[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int ie = w.emin + r <= w.emax ? w.emin + r : w.emax;
    const int* Tr = TL + ie * k3dPhi;
    s0[r] = Tr[w.a0];
    t0[r] = Tr[w.b0 + 1];
    s1[r] = 0;
    t1[r] = 0;
  }
  if (win_wraps(w)) {  // the second run of each row (phi bins 0..b1) only exists for wrapped windows
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int ie = w.emin + r <= w.emax ? w.emin + r : w.emax;
      const int* Tr = TL + ie * k3dPhi;
      s1[r] = Tr[w.a1];
      t1[r] = Tr[w.b1 + 1];
    }
  }
  const int off = cl - w.minL;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (w.emin + r <= w.emax) {
      if (s0[r] < t0[r]) {
        list[cnt * kThreads] = run_entry(s0[r] - rlo, t0[r] - rlo, off);
        ++cnt;
      }
      if (s1[r] < t1[r]) {
        list[cnt * kThreads] = run_entry(s1[r] - rlo, t1[r] - rlo, off);
        ++cnt;
      }
    }
  }
  return cnt;
}

// Walks the cnt (> 0 or none) listed runs, calling f.run(layerOff) at the start of each run and
// f.pair(xo, two) for the staged candidates at {r_over_absz, phi} byte offset xo (8 bytes per
// candidate) and, if two, the next one: a step costs one compare and two adds of bookkeeping. The
// second candidate of a run's last step is read but masked off; the staged range is kept below
// the staging capacity, so that read stays inside the arrays. The next entry is read ahead.
template <class F>
__device__ __forceinline__ void walk_list(const uint32_t* list, int cnt, F& f) {
  if (cnt <= 0) return;
  uint32_t e = list[0];
  int xo = (int)(e & kOffMask) << 3, tx = (int)((e >> kOffBits) & kOffMask) << 3;
  f.run((int)(e >> (2 * kOffBits)));
  int j = 1;
  e = list[(cnt > 1 ? 1 : 0) * kThreads];
  while (xo < tx) {
    f.pair(xo, xo + 8 < tx);
    xo += 16;
    if (xo >= tx && j < cnt) {
      xo = (int)(e & kOffMask) << 3;
      tx = (int)((e >> kOffBits) & kOffMask) << 3;
      f.run((int)(e >> (2 * kOffBits)));
      ++j;
      e = list[(j < cnt ? j : cnt - 1) * kThreads];
    }
  }
}

// Lists and walks the window layer by layer, as many layers per round as the list is sure to hold.
// skipL: a layer left out (the nearest higher's own layer), or -1.
template <int W, class F>
__device__ __forceinline__ void walk_window(const int* __restrict__ T, const Win3D& w, int skipL, int rlo,
                                            uint32_t* list, F& f) {
  const int perLayer = (2 * W + 1) * (win_wraps(w) ? 2 : 1);
  int cl = w.minL;
  while (cl <= w.maxL) {
    int cnt = 0;
    while (cl <= w.maxL && cnt + perLayer <= kRunCap) {
      if (cl != skipL) cnt = list_layer<W>(T, w, cl, rlo, list, cnt);
      ++cl;
    }
    walk_list(list, cnt, f);
  }
}

// The window row by row from global memory (unstaged blocks): f.run(cl - minL) per layer and
// f.one(q) per candidate, positions absolute.
template <int W, class F>
__device__ __forceinline__ void scan_window(const int* __restrict__ T, const Win3D& w, int skipL, F& f) {
  const bool wrap = win_wraps(w);
  for (int cl = w.minL; cl <= w.maxL; ++cl) {
    if (cl == skipL) continue;
    f.run(cl - w.minL);
    for (int ie = w.emin; ie <= w.emax; ++ie) {
      const int* Tr = T + cl * k3dBinsPerLayer + ie * k3dPhi;
      for (int q = Tr[w.a0], t = Tr[w.b0 + 1]; q < t; ++q) f.one(q);
      if (wrap)
        for (int q = Tr[w.a1], t = Tr[w.b1 + 1]; q < t; ++q) f.one(q);
    }
  }
}

// ---------------------------------------------------------------- staged candidates

// The three fields a candidate step reads: {r_over_absz, phi, energy} for the density, {r_over_absz,
// phi, rho} for the nearest higher (the fourth field of the tile-order records, z, is only read for
// the query point, from registers). Staged as 12 bytes per candidate instead of 16, in two arrays so
// that the two consecutive candidates of a walk step load with one 64-bit pair read and one 32-bit
// pair read: xy[k] = {r_over_absz, phi}, z[k] = energy or rho.
struct Cand3 {
  float x, y, z;
};
struct Staged3 {
  const float2* xy;
  const float* z;
  __device__ __forceinline__ Cand3 operator[](int q) const {
    const float2 a = xy[q];
    return Cand3{a.x, a.y, z[q]};
  }
  // the candidates at xy byte offset xo (z at xo / 2) and the one after it
  __device__ __forceinline__ void pair(int xo, Cand3& c0, Cand3& c1) const {
    const char* bx = reinterpret_cast<const char*>(xy) + xo;
    const char* bz = reinterpret_cast<const char*>(z) + (xo >> 1);
    const float2 a = *reinterpret_cast<const float2*>(bx);
    const float2 b = *reinterpret_cast<const float2*>(bx + 8);
    c0 = Cand3{a.x, a.y, *reinterpret_cast<const float*>(bz)};
    c1 = Cand3{b.x, b.y, *reinterpret_cast<const float*>(bz + 4)};
  }
};
// Candidates read from the tile-order records in global memory (unstaged blocks).
struct Global4 {
  const float4* p;
  __device__ __forceinline__ float4 operator[](int q) const { return p[q]; }
};

// Copy [lo, hi) (1 <= hi - lo <= STAGE) of the tile-order records into LDS, all loads of a thread in
// flight together.
template <int STAGE>
__device__ __forceinline__ void stage_range3(const float4* __restrict__ g, float2* sxy, float* sz, int lo, int hi) {
  static_assert(STAGE % kThreads == 0, "staging splits evenly over the block");
  constexpr int kPer = STAGE / kThreads;
  const int n = hi - lo;
  float4 v[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = u * kThreads + (int)threadIdx.x;
    v[u] = g[lo + (k < n ? k : n - 1)];
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = u * kThreads + (int)threadIdx.x;
    if (k < n) {
      sxy[k] = make_float2(v[u].x, v[u].y);
      sz[k] = v[u].z;
    }
  }
}

// ---------------------------------------------------------------- density

// isReachable and the energy to add (CLUE3DAlgoKernels.h:39-43, :118-134) for candidates read
// from `src` (the staged range or the whole sorted array). A contribution that is not reachable
// adds +0.f, which leaves rho unchanged: rho starts at +0 and a round-to-nearest sum started from
// +0 is never -0.
template <class Src>
struct DensityStep3D {
  int p;  // the point's own position (same frame as the candidates')
  int L, minL;
  float r0, zi, phii, dsq, kdf, kdfSame, kdfRun;
  Src src;
  float rho;
  __device__ __forceinline__ void run(int layerOff) { kdfRun = (minL + layerOff == L) ? kdfSame : kdf; }
  template <class C>
  __device__ __forceinline__ float term(const C& c, float dp, bool self) const {
    const float r1 = c.x * zi;
    const float d = (r0 - r1) * (r0 - r1) + r1 * r1 * dp * dp;
    const float e = (self ? 1.f : kdfRun) * c.z;
    return d < dsq ? e : 0.f;
  }
  // staged candidates at byte offset xo and (if two) xo + 8 (see walk_list)
  __device__ __forceinline__ void pair(int xo, bool two) {
    Cand3 c0, c1;
    src.pair(xo, c0, c1);
    float dp0 = phii - c0.y, dp1 = phii - c1.y;
    reduce_pair(dp0, dp1);
    const int self = p << 3;
    const float e0 = term(c0, dp0, xo == self);
    const float e1 = term(c1, dp1, xo + 8 == self);
    rho += e0;
    rho += two ? e1 : 0.f;
  }
  __device__ __forceinline__ void one(int q) {
    const auto c = src[q];
    rho += term(c, reduce_range(phii - c.y), q == p);
  }
};

// rho of every point (KernelCalculateDensitySoA). cre: sorted {r_over_absz, phi, energy, z};
// writes crp = sorted {r_over_absz, phi, rho, z} and rho by SoA index.
template <int ITEMS, int STAGE>
__global__ __launch_bounds__(kThreads) void k_density3d(
    const BlockDesc* __restrict__ blocks, const int* __restrict__ perm, const float4* __restrict__ cre,
    const float* __restrict__ eta, const float* __restrict__ radius, const int32_t* __restrict__ layer,
    const int32_t* __restrict__ isSi, const int* __restrict__ table, int64_t stride, Clue3DParams prm,
    float4* __restrict__ crp, float* __restrict__ rhoOut) {
  using Shape = CompShape3D<ITEMS>;
  static_assert(STAGE < (1 << kOffBits), "staged offsets fit an entry field");
  __shared__ float2 s_xy[STAGE];
  __shared__ float s_z[STAGE];
  // the run lists; before them (block_range, then a barrier) its first words are the reduction scratch
  __shared__ uint32_t s_runs[kRunCap * kThreads];
  int* s_red = reinterpret_cast<int*>(s_runs);
  // consecutive compute blocks (overlapping staged ranges) run on one XCD and share its L2
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const BlockDesc bd = blocks[blk / Shape::kPerSort];
  const int b0 = bd.begin + (blk % Shape::kPerSort) * Shape::kHits;
  const int b1 = (b0 + Shape::kHits) < bd.end ? (b0 + Shape::kHits) : bd.end;
  if (b0 >= b1) return;  // uniform over the block
  const int* T = table + (int64_t)bd.event * stride;
  const int S = prm.densitySiblingLayers;
  Win3D win[ITEMS];
  float4 me[ITEMS];
  int lo = INT_MAX, hi = -1;
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p < b1) {
      const int i = perm[p];
      me[r] = cre[p];
      win[r] = make_win<2>(clamp_layer3d(layer[i]), eta_bin3d(eta[i]), clamp_phibin(phi_bin3d(me[r].y)), S,
                           k3dLastPerSide);
      int l, h;
      win_reach(T, win[r], l, h);
      lo = l < lo ? l : lo;
      hi = h > hi ? h : hi;
    }
  }
  int rlo, rhi;
  block_range(lo, hi, s_red, rlo, rhi);
  const bool staged = S <= kMaxListS && rhi - rlo < STAGE;  // a spare slot for the walk's masked read
  if (staged && rhi > rlo) stage_range3<STAGE>(cre, s_xy, s_z, rlo, rhi);
  __syncthreads();
  const float kdf = prm.kernelDensityFactor;
  const float kdfSame = prm.kernelDensityFactor * (prm.densityOnSameLayer ? 1.f : 0.f);
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p < b1) {
      const int i = perm[p];
      const float zi = me[r].w;
      const float dsq = isSi[i] ? prm.densityXYDistanceSqr : radius[i] * radius[i];
      float rho;
      if (staged) {
        DensityStep3D<Staged3> f{p - rlo, win[r].L, win[r].minL, me[r].x * zi, zi, me[r].y, dsq, kdf, kdfSame, kdf,
                                 Staged3{s_xy, s_z}, 0.f};
        walk_window<2>(T, win[r], -1, rlo, s_runs + threadIdx.x, f);
        rho = f.rho;
      } else {
        DensityStep3D<Global4> f{p, win[r].L, win[r].minL, me[r].x * zi, zi, me[r].y, dsq, kdf, kdfSame, kdf,
                                 Global4{cre}, 0.f};
        scan_window<2>(T, win[r], -1, f);
        rho = f.rho;
      }
      crp[p] = make_float4(me[r].x, me[r].y, rho, zi);
      rhoOut[i] = rho;
    }
  }
}

// ---------------------------------------------------------------- nearest higher

// distanceSqr and the "last minimum wins" update (CLUE3DAlgoKernels.h:296-299, :357-366) for
// candidates read from `src`; bo keeps the sibling-layer offset of the current best.
template <class Src>
struct NhStep3D {
  float r0, zi, phii, rhoi;
  Src src;
  float best;
  int bq, bo, cur;
  __device__ __forceinline__ void run(int layerOff) { cur = layerOff; }
  template <class C>
  __device__ __forceinline__ void take(const C& c, float dp, int q, bool valid) {
    const float r1 = c.x * zi;
    const float d = (r0 - r1) * (r0 - r1) + r1 * r1 * dp * dp;
    // bitwise forms: lane masks combined without branches
    const bool upd = valid & (c.z > rhoi) & (d <= best);
    best = upd ? d : best;
    bq = upd ? q : bq;
    bo = upd ? cur : bo;
  }
  // staged candidates at byte offset xo and (if two) xo + 8 (see walk_list); bq keeps the byte offset
  __device__ __forceinline__ void pair(int xo, bool two) {
    Cand3 c0, c1;
    src.pair(xo, c0, c1);
    float dp0 = phii - c0.y, dp1 = phii - c1.y;
    reduce_pair(dp0, dp1);
    take(c0, dp0, xo, true);
    take(c1, dp1, xo + 8, two);
  }
  __device__ __forceinline__ void one(int q) {
    const auto c = src[q];
    take(c, reduce_range(phii - c.y), q, true);
  }
};

// delta / nearestHigher (KernelComputeDistanceToHigherSoA) plus the seed / outlier decision of
// KernelFindAndAssignClustersSoA (CLUE3DAlgoKernels.h:522-543). crp: sorted {r_over_absz, phi,
// rho, z}; cre: sorted {.., energy, ..}.
template <int ITEMS, int STAGE>
__global__ __launch_bounds__(kThreads) void k_nh3d(
    const BlockDesc* __restrict__ blocks, const int* __restrict__ perm, const float4* __restrict__ crp,
    const float4* __restrict__ cre, const float* __restrict__ eta, const float* __restrict__ radius,
    const int32_t* __restrict__ layer, const int32_t* __restrict__ isSi, const int* __restrict__ table,
    int64_t stride, const int* __restrict__ evOff, Clue3DParams prm, float* __restrict__ deltaF,
    int32_t* __restrict__ deltaL, int32_t* __restrict__ nhLayer, int32_t* __restrict__ nhIdx,
    int32_t* __restrict__ isSeedOut, int* __restrict__ parent) {
  using Shape = CompShape3D<ITEMS>;
  static_assert(STAGE < (1 << kOffBits), "staged offsets fit an entry field");
  __shared__ float2 s_xy[STAGE];
  __shared__ float s_z[STAGE];
  // the run lists; before them (block_range, then a barrier) its first words are the reduction scratch
  __shared__ uint32_t s_runs[kRunCap * kThreads];
  int* s_red = reinterpret_cast<int*>(s_runs);
  // consecutive compute blocks (overlapping staged ranges) run on one XCD and share its L2
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const BlockDesc bd = blocks[blk / Shape::kPerSort];
  const int b0 = bd.begin + (blk % Shape::kPerSort) * Shape::kHits;
  const int b1 = (b0 + Shape::kHits) < bd.end ? (b0 + Shape::kHits) : bd.end;
  if (b0 >= b1) return;
  const int* T = table + (int64_t)bd.event * stride;
  const int evBase = evOff[bd.event];
  const int S = prm.densitySiblingLayers;
  Win3D win[ITEMS];
  float4 me[ITEMS];
  int lo = INT_MAX, hi = -1;
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p < b1) {
      const int i = perm[p];
      me[r] = crp[p];
      win[r] = make_win<1>(clamp_layer3d(layer[i]), eta_bin3d(eta[i]), clamp_phibin(phi_bin3d(me[r].y)), S,
                           k3dLastPerSide + 1);
      int l, h;
      win_reach(T, win[r], l, h);
      lo = l < lo ? l : lo;
      hi = h > hi ? h : hi;
    }
  }
  int rlo, rhi;
  block_range(lo, hi, s_red, rlo, rhi);
  const bool staged = S <= kMaxListS && rhi - rlo < STAGE;  // a spare slot for the walk's masked read
  if (staged && rhi > rlo) stage_range3<STAGE>(crp, s_xy, s_z, rlo, rhi);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = b0 + r * kThreads + (int)threadIdx.x;
    if (p < b1) {
      const int i = perm[p];
      const int L = win[r].L;
      const float zi = me[r].w;
      const float rhoi = me[r].z;
      const int skipL = prm.nearestHigherOnSameLayer ? -1 : L;
      float best;
      int bq, bo;
      if (staged) {
        NhStep3D<Staged3> f{me[r].x * zi, zi, me[r].y, rhoi, Staged3{s_xy, s_z}, FLT_MAX, -1, 0, 0};
        walk_window<1>(T, win[r], skipL, rlo, s_runs + threadIdx.x, f);
        best = f.best;
        bq = (f.bq >> 3) + rlo;
        bo = f.bo;
      } else {
        NhStep3D<Global4> f{me[r].x * zi, zi, me[r].y, rhoi, Global4{crp}, FLT_MAX, -1, 0, 0};
        scan_window<1>(T, win[r], skipL, f);
        best = f.best;
        bq = f.bq;
        bo = f.bo;
      }
      float df;
      int dl, hl, hiG;
      if (best != FLT_MAX) {
        df = sqrtf(best);
        hl = win[r].minL + bo;
        dl = hl > L ? hl - L : L - hl;
        hiG = perm[bq];
      } else {
        df = FLT_MAX;
        dl = INT_MAX;
        hl = -1;
        hiG = -1;
      }
      deltaF[i] = df;
      deltaL[i] = dl;
      nhLayer[i] = hl;
      nhIdx[i] = hiG >= 0 ? hiG - evBase : -1;
      const float energy = cre[p].z;
      const float cut = isSi[i] ? prm.criticalXYDistance : radius[i];
      const bool seed = (df > cut || (float)dl > prm.criticalZDistanceLyr) && (rhoi >= prm.criticalDensity) &&
                        (energy / rhoi > prm.criticalSelfDensity);
      const bool outlier = (df > prm.outlierMultiplier * prm.criticalXYDistance) && (rhoi < prm.criticalDensity);
      isSeedOut[i] = seed ? 1 : 0;
      // Followers point at their nearest higher; seeds, outliers (and a point without a nearest
      // higher) are roots.
      parent[i] = (seed || outlier || hl < 0) ? i : hiG;
    }
  }
}

}  // namespace g3d
}  // namespace clue
