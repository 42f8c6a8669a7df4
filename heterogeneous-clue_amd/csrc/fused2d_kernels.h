// Device code of the CLUE2D fused segment path (see fused2d.h): the layer bucket pass functors, the
// segment lists, the fused segment kernel and the seed numbering. Header-only so that the same
// kernels also build for the CPU emulation of the HIP execution model used by the tests.
//
// Bit-exactness inside a segment rests on the same facts as the global-sort path: the segment's
// points are ordered by (xBin, yBin, point index), so a search-box column (xBin, yLo..yHi) is one
// contiguous run and walking the columns in ascending xBin visits candidates in the serial order
// of kernel_calculate_density / kernel_calculate_distanceToHigher
// (src/serial/plugin-CLUEClusterizer/CLUEAlgoKernels.h:25-106); the arithmetic is the reference's
// fp32 expressions without contraction, with a correctly rounded sqrtf.
#pragma once

#include <float.h>
#include <limits.h>

#include "fused2d.h"

namespace clue {
namespace fused {

constexpr int kTabEntries = k2dBinsPerLayer + 1;  // bin table of one layer: 10^4 bins + end
constexpr uint16_t kNone = 0xFFFF;

constexpr int align16(int b) { return (b + 15) / 16 * 16; }

// ---------------------------------------------------------------- (1) bucket pass by layer

struct LayerKey {
  const int32_t* layer;
  int* err;
  __device__ __forceinline__ uint32_t operator()(int i, int) const {
    int L = layer[i];
    if (L < 0 || L >= k2dLayers) {
      atomicOr(err, kErrLayerRange);
      L = L < 0 ? 0 : k2dLayers - 1;
    }
    return (uint32_t)L;
  }
};

struct EmitPoint2D {
  const float* x;
  const float* y;
  const float* w;
  float4* lg;
  using Payload = float4;
  using Out = float4;
  __device__ __forceinline__ void block_starts(const BlockDesc&, const int*) const {}
  __device__ __forceinline__ Payload load(int i) const { return make_float4(x[i], y[i], w[i], __int_as_float(i)); }
  __device__ __forceinline__ Out fetch(const Payload& v) const { return v; }
  __device__ __forceinline__ void store(int pos, uint32_t, const Out& v) const { lg[pos] = v; }
};

using DigitLayer = Digit<1, kRadix>;

// ---------------------------------------------------------------- (2) segment lists

// One block per event, one thread per layer. After the scan, hist[kRadix*b0 + d*nb] is the batch
// position of the event's first point of layer d (block 0 of the event, digit d), so a layer's
// segment is [that, the same for d + 1); digits past the last layer are empty.
__global__ __launch_bounds__(kRadix) void k_segments2d(const int* __restrict__ evOff, const int* __restrict__ evBlk,
                                                       const int* __restrict__ histScanned, int4* __restrict__ segs,
                                                       int segCap, int* __restrict__ ctl, int strict,
                                                       int* __restrict__ err) {
  const int e = blockIdx.x, l = (int)threadIdx.x;
  if (l >= k2dLayers) return;
  const int b0 = evBlk[e], nb = evBlk[e + 1] - b0;
  if (nb == 0) return;
  const int* H = histScanned + (int64_t)kRadix * b0;
  const int start = H[l * nb];
  const int n = H[(l + 1) * nb] - start;
  if (n == 0) return;
  atomicMax(&ctl[8], n);
  const int c = n <= 1024 ? 0 : n <= 4096 ? 1 : n <= kFusedCap ? 2 : -1;
  if (c < 0) {
    if (strict) atomicOr(err, kErrCapacity);
    return;
  }
  const int slot = atomicAdd(&ctl[c], 1);
  segs[(int64_t)c * segCap + slot] = make_int4(start, n, evOff[e], 0);
}

// ---------------------------------------------------------------- (3) fused segment kernel

template <int CAP, int NT>
struct FusedLayout {
  static constexpr int W = NT / kWave;
  static constexpr int R = (CAP + NT - 1) / NT;  // points per thread
  static constexpr int offPos = 0;                // u16 [CAP]: sorted position -> position in segment
  static constexpr int offTab = align16(2 * CAP);  // u16 [kTabEntries]: bin -> first sorted position
  static constexpr int offX = offTab + align16(2 * kTabEntries);
  static constexpr int bytes = offX + 12 * CAP;
  // region X while sorting: two key buffers and the per-wave digit counters
  static constexpr int offA = offX;
  static constexpr int offB = offX + 4 * CAP;
  static constexpr int offCnt = offX + 8 * CAP;
  static constexpr int offOff = offCnt + W * kRadix * 4;
  // region X while searching: {x, y} and weight (then rho) in sorted order
  static constexpr int offXY = offX;
  static constexpr int offW = offX + 8 * CAP;
  // region X while walking: parent (sorted position) and seed flag
  static constexpr int offPar = offX;
  static constexpr int offFlag = offX + 2 * CAP;
  static_assert(offOff + W * kRadix * 4 <= bytes, "sort scratch fits region X");
  static_assert(CAP <= 16384, "segment positions take 14 bits");
  static_assert(NT % kWave == 0 && NT >= 2 * kWave, "whole waves, at least 128 threads");
};

template <int NT>
__device__ __forceinline__ int block_exclusive_min_nt(int v, int* s_tmp) {
  constexpr int W = NT / kWave;
  const int lane = lane_id(), wave = (int)threadIdx.x / kWave;
  const int inc = wave_inclusive_min(v);
  if (lane == kWave - 1) s_tmp[wave] = inc;
  __syncthreads();
  int base = INT_MAX;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int s = s_tmp[w];
    if (w < wave) base = s < base ? s : base;
  }
  __syncthreads();
  int excl = __shfl_up(inc, 1, kWave);
  if (lane == 0) excl = INT_MAX;
  return excl < base ? excl : base;
}

// Stable counting-sort pass over the segment's n keys in[0, n) into out: item k = r * NT + tid is
// ranked in round r (rounds in order, waves in order inside a round), so equal digits keep their
// input order. The rank of an item inside its digit comes from a wave ballot match, per-wave
// counters for the waves before it and a running count per digit for the rounds before; it waits
// in s_rank (u16) until the digit starts are known. DIV selects the digit (key >> 14) / DIV % 100.
template <int NT, int R, int DIV>
__device__ __forceinline__ void lds_sort_pass(const uint32_t* in, int n, uint32_t* out, uint16_t* s_rank, int* s_cnt,
                                              int* s_off, int* s_run, int* s_start) {
  constexpr int W = NT / kWave;
  const int tid = (int)threadIdx.x, lane = lane_id(), wave = tid / kWave;
  if (tid < kRadix) s_run[tid] = 0;
  for (int k = lane; k < kRadix; k += kWave) s_cnt[wave * kRadix + k] = 0;
  __syncthreads();
#pragma unroll 1
  for (int r = 0; r < R; ++r) {
    const int k = r * NT + tid;
    const bool valid = k < n;
    const int d = valid ? (int)((in[k] >> 14) / DIV % k2dRows) : kRadix - 1;
    const uint64_t m = match_digit7(d);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (valid && before == 0) s_cnt[wave * kRadix + d] = (int)__popcll(m);
    __syncthreads();
    if (tid < kRadix) {
      int run = s_run[tid];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        s_off[w * kRadix + tid] = run;
        run += s_cnt[w * kRadix + tid];
      }
      s_run[tid] = run;
    }
    __syncthreads();
    if (valid) s_rank[k] = (uint16_t)(s_off[wave * kRadix + d] + before);
    for (int j = lane; j < kRadix; j += kWave) s_cnt[wave * kRadix + j] = 0;
  }
  if (tid < kWave) {
    const int v0 = s_run[2 * tid], v1 = s_run[2 * tid + 1];
    const int inc = wave_inclusive_sum(v0 + v1);
    s_start[2 * tid] = inc - v0 - v1;
    s_start[2 * tid + 1] = inc - v1;
  }
  __syncthreads();
#pragma unroll 1
  for (int k = tid; k < n; k += NT) {
    const uint32_t c = in[k];
    out[s_start[(c >> 14) / DIV % k2dRows] + s_rank[k]] = c;
  }
  __syncthreads();
}

// rho contributions of the sorted positions [s, t) to point k (CLUEAlgoKernels.h:36-52). A failed
// distance test adds +0.f, which leaves rho unchanged (rho starts at +0 and never becomes -0).
__device__ __forceinline__ float dens_run(int k, float xi, float yi, int s, int t, const float2* sxy, const float* sw,
                                          float dc2, float rho) {
  int q = s;
  for (; q + 1 < t; q += 2) {
    const float2 c0 = sxy[q], c1 = sxy[q + 1];
    const float w0 = sw[q], w1 = sw[q + 1];
    const float dx0 = xi - c0.x, dy0 = yi - c0.y;
    const float dx1 = xi - c1.x, dy1 = yi - c1.y;
    const float d20 = dx0 * dx0 + dy0 * dy0;
    const float d21 = dx1 * dx1 + dy1 * dy1;
    const float v0 = (q == k ? 1.f : 0.5f) * w0;
    const float v1 = (q + 1 == k ? 1.f : 0.5f) * w1;
    rho += d20 <= dc2 ? v0 : 0.f;
    rho += d21 <= dc2 ? v1 : 0.f;
  }
  if (q < t) {
    const float2 c0 = sxy[q];
    const float dx0 = xi - c0.x, dy0 = yi - c0.y;
    const float d20 = dx0 * dx0 + dy0 * dy0;
    const float v0 = (q == k ? 1.f : 0.5f) * sw[q];
    rho += d20 <= dc2 ? v0 : 0.f;
  }
  return rho;
}

// Nearest-higher scan of [s, t) with the deferred square root: the running minimum of d2 (first of
// equal values) and the minimum before it. "Higher" is (rho_j > rho_i) or (rho_j == rho_i and
// j > i); positions in the segment are in point-index order, so j > i compares them.
__device__ __forceinline__ void nh_run(int lpi, float xi, float yi, float rhoi, int s, int t, const float2* sxy,
                                       const float* sr, const uint16_t* spos, float dm2, float& best2, int& best,
                                       float& prev2, int& prev) {
  for (int q = s; q < t; ++q) {
    const float2 c = sxy[q];
    const float rq = sr[q];
    const float dx = xi - c.x;
    const float dy = yi - c.y;
    const float d2 = dx * dx + dy * dy;
    bool higher = rq > rhoi;
    if (rq == rhoi) higher = (int)spos[q] > lpi;
    const bool upd = higher && d2 <= dm2 && d2 < best2;
    prev2 = upd ? best2 : prev2;
    prev = upd ? best : prev;
    best2 = upd ? d2 : best2;
    best = upd ? q : best;
  }
}

// The reference's own update rule (CLUEAlgoKernels.h:84-100), for the rare case where two
// different d2 values share the minimal sqrtf.
__device__ __forceinline__ void nh_run_exact(int lpi, float xi, float yi, float rhoi, int s, int t,
                                             const float2* sxy, const float* sr, const uint16_t* spos, float dm2,
                                             float& delta, int& nh) {
  for (int q = s; q < t; ++q) {
    const float2 c = sxy[q];
    const float rq = sr[q];
    const float dx = xi - c.x;
    const float dy = yi - c.y;
    const float d2 = dx * dx + dy * dy;
    const bool higher = (rq > rhoi) || ((rq == rhoi) && ((int)spos[q] > lpi));
    if (higher && d2 <= dm2) {
      const float d = sqrtf(d2);
      if (d < delta) {
        delta = d;
        nh = q;
      }
    }
  }
}

// Batch index of the point at grouped position p (the index rides in the w lane of lg).
__device__ __forceinline__ int lg_index(const float4* lg, int p) {
  return reinterpret_cast<const int*>(lg)[4 * (int64_t)p + 3];
}

template <int CAP, int NT>
__global__ __launch_bounds__(NT) void k_fused2d(const int4* __restrict__ segs, const int* __restrict__ segCount,
                                                int* __restrict__ queue, const float4* __restrict__ lg, float dc,
                                                float dc2, float rhoc, float odf, float dm2,
                                                float* __restrict__ rhoOut, float* __restrict__ deltaOut,
                                                int32_t* __restrict__ nhOut, int32_t* __restrict__ isSeedOut,
                                                int* __restrict__ rootOut) {
  using Ly = FusedLayout<CAP, NT>;
  constexpr int R = Ly::R;
  __shared__ __align__(16) unsigned char s_mem[Ly::bytes];
  __shared__ int s_run[kRadix], s_start[kRadix], s_tmp[Ly::W];
  __shared__ int s_item;
  uint16_t* s_pos = reinterpret_cast<uint16_t*>(s_mem + Ly::offPos);
  uint16_t* s_tab = reinterpret_cast<uint16_t*>(s_mem + Ly::offTab);
  uint32_t* s_a = reinterpret_cast<uint32_t*>(s_mem + Ly::offA);
  uint32_t* s_b = reinterpret_cast<uint32_t*>(s_mem + Ly::offB);
  int* s_cnt = reinterpret_cast<int*>(s_mem + Ly::offCnt);
  int* s_off = reinterpret_cast<int*>(s_mem + Ly::offOff);
  float2* s_xy = reinterpret_cast<float2*>(s_mem + Ly::offXY);
  float* s_w = reinterpret_cast<float*>(s_mem + Ly::offW);
  uint16_t* s_par = reinterpret_cast<uint16_t*>(s_mem + Ly::offPar);
  uint8_t* s_flag = reinterpret_cast<uint8_t*>(s_mem + Ly::offFlag);

  const int tid = (int)threadIdx.x;
  const int count = *segCount;
  const float dm = odf * dc;
  for (;;) {
    if (tid == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int item = s_item;
    if (item >= count) break;  // uniform: every thread read the same item
    const int4 sg = segs[item];
    const int start = sg.x, n = sg.y, evBase = sg.z;
    for (int b = tid; b < kTabEntries; b += NT) s_tab[b] = kNone;

    // ---- sort the segment by (xBin, yBin, position): LSD, yBin then xBin (s_pos holds the ranks)
#pragma unroll 1
    for (int lp = tid; lp < n; lp += NT) {
      const float4 v = lg[start + lp];
      s_b[lp] = ((uint32_t)(xbin2d(v.x) * k2dRows + ybin2d(v.y)) << 14) | (uint32_t)lp;
    }
    __syncthreads();
    lds_sort_pass<NT, R, 1>(s_b, n, s_a, s_pos, s_cnt, s_off, s_run, s_start);
    lds_sort_pass<NT, R, k2dRows>(s_a, n, s_b, s_pos, s_cnt, s_off, s_run, s_start);

    // ---- bin table: s_tab[b] = first sorted position with bin >= b (n if none)
#pragma unroll 2
    for (int k = tid; k < n; k += NT) {
      const uint32_t c = s_b[k];
      s_pos[k] = (uint16_t)(c & 0x3FFFu);
      const uint32_t key = c >> 14;
      if (k == 0 || (s_b[k - 1] >> 14) != key) s_tab[key] = (uint16_t)k;
    }
    __syncthreads();
    {
      constexpr int chunk = (kTabEntries + NT - 1) / NT;
      const int j0 = (NT - 1 - tid) * chunk;  // thread order runs from the high bins down
      int m = INT_MAX;
      for (int j = 0; j < chunk; ++j) {
        const int b = j0 + j;
        if (b < kTabEntries) {
          const int v = s_tab[b];
          if (v != kNone) m = v < m ? v : m;
        }
      }
      int run = block_exclusive_min_nt<NT>(m, s_tmp);
      run = run < n ? run : n;
      for (int j = chunk - 1; j >= 0; --j) {
        const int b = j0 + j;
        if (b < kTabEntries) {
          const int v = s_tab[b];
          if (v != kNone) run = v < run ? v : run;
          s_tab[b] = (uint16_t)run;
        }
      }
    }
    __syncthreads();

    // ---- points in sorted order
#pragma unroll 2
    for (int k = tid; k < n; k += NT) {
      const float4 v = lg[start + s_pos[k]];
      s_xy[k] = make_float2(v.x, v.y);
      s_w[k] = v.z;
    }
    __syncthreads();

    // ---- density
    float mrho[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = r * NT + tid;
      __builtin_amdgcn_sched_barrier(0);  // keep the points' searches apart (register pressure)
      float rho = 0.f;
      if (k < n) {
        const float2 me = s_xy[k];
        const int xlo = xbin2d(me.x - dc), xhi = xbin2d(me.x + dc);
        const int ylo = ybin2d(me.y - dc), yhi = ybin2d(me.y + dc);
        for (int c = xlo; c <= xhi; ++c)
          rho = dens_run(k, me.x, me.y, s_tab[c * k2dRows + ylo], s_tab[c * k2dRows + yhi + 1], s_xy, s_w, dc2, rho);
      }
      mrho[r] = rho;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = r * NT + tid;
      if (k < n) s_w[k] = mrho[r];
    }
    __syncthreads();

    // ---- nearest higher, seed / outlier decision (CLUEAlgoKernels.h:57-137)
    int mpar[R];  // parent (sorted position) | seed flag << 16
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = r * NT + tid;
      __builtin_amdgcn_sched_barrier(0);
      mpar[r] = k;
      if (k < n) {
        const float2 me = s_xy[k];
        const float xi = me.x, yi = me.y, rhoi = s_w[k];
        const int lpi = s_pos[k];
        const int xlo = xbin2d(xi - dm), xhi = xbin2d(xi + dm);
        const int ylo = ybin2d(yi - dm), yhi = ybin2d(yi + dm);
        float best2 = INFINITY, prev2 = INFINITY;
        int best = -1, prev = -1;
        for (int c = xlo; c <= xhi; ++c)
          nh_run(lpi, xi, yi, rhoi, s_tab[c * k2dRows + ylo], s_tab[c * k2dRows + yhi + 1], s_xy, s_w, s_pos, dm2,
                 best2, best, prev2, prev);
        float delta = FLT_MAX;
        if (best >= 0) {
          delta = sqrtf(best2);
          if (prev >= 0 && sqrtf(prev2) == delta) {
            delta = FLT_MAX;
            best = -1;
            for (int c = xlo; c <= xhi; ++c)
              nh_run_exact(lpi, xi, yi, rhoi, s_tab[c * k2dRows + ylo], s_tab[c * k2dRows + yhi + 1], s_xy, s_w,
                           s_pos, dm2, delta, best);
          }
        }
        const int i = lg_index(lg, start + lpi);
        rhoOut[i] = rhoi;
        deltaOut[i] = delta;
        nhOut[i] = best >= 0 ? lg_index(lg, start + s_pos[best]) - evBase : -1;
        const bool seed = (delta > dc) && (rhoi >= rhoc);
        const bool outlier = (delta > dm) && (rhoi < rhoc);
        isSeedOut[i] = seed ? 1 : 0;
        // followers point at their nearest higher; seeds, outliers and a point without a nearest
        // higher are roots
        mpar[r] = ((seed || outlier || best < 0) ? k : best) | (seed ? 1 << 16 : 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = r * NT + tid;
      if (k < n) {
        s_par[k] = (uint16_t)(mpar[r] & 0xFFFF);
        s_flag[k] = (uint8_t)(mpar[r] >> 16);
      }
    }
    __syncthreads();

    // ---- follower trees: root of every point (the forest is acyclic: a parent has a strictly
    // larger (rho, index))
#pragma unroll 2
    for (int k = tid; k < n; k += NT) {
      int cur = k;
      for (;;) {
        const int p = s_par[cur];
        if (p == cur) break;
        cur = p;
      }
      rootOut[lg_index(lg, start + s_pos[k])] = s_flag[cur] ? lg_index(lg, start + s_pos[cur]) : -1;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- (4) seed numbering

__global__ __launch_bounds__(kThreads) void k_cluster2d(const BlockDesc* __restrict__ blocks,
                                                        const int* __restrict__ evOff, const int* __restrict__ root,
                                                        const int* __restrict__ rank, int32_t* __restrict__ label) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int evBegin = evOff[bd.event], evEnd = evOff[bd.event + 1];
  const int base = rank[evBegin];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = bd.begin + r * kThreads + (int)threadIdx.x;
    if (i < bd.end) {
      // root is -1 for outlier trees and for points of a segment no kernel took (over capacity)
      const int rt = root[i];
      label[i] = (rt >= evBegin && rt < evEnd) ? rank[rt] - base : -1;
    }
  }
}

}  // namespace fused
}  // namespace clue
