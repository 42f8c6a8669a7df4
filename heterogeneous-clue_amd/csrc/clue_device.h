// Device-side building blocks shared by the CLUE2D and CLUE3D HIP kernels (gfx950, wave64).
//
// Arithmetic contract: every expression that reaches a result is evaluated in IEEE fp32 (fp64
// where the reference uses double), with the reference's operand order, no FMA contraction
// (the library is compiled with -ffp-contract=off and this header also pins it with a pragma),
// correctly rounded sqrt/division, and denormals preserved.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

// Minimum waves per SIMD a kernel's register allocation must admit (the CPU emulation of the HIP
// model defines it away).
#ifndef CLUE_WAVES_PER_EU
#define CLUE_WAVES_PER_EU(...) __attribute__((amdgpu_waves_per_eu(__VA_ARGS__)))
#endif

namespace clue {

// Device-side error bits, OR-ed into the context's flag word and reported after the run.
constexpr int kErrLayerRange = 1;  // a layer outside the tile array (reference: out-of-bounds access)
constexpr int kErrBinRange = 2;    // a phi bin outside [0, 126] (phi far outside [-pi, pi])

constexpr int kWave = 64;
constexpr int kThreads = 256;              // every kernel of this library runs 256-thread blocks
constexpr int kWaves = kThreads / kWave;

// ---------------------------------------------------------------- float -> int as on x86
// The reference converts with `int v = float_expr;` on x86-64 (cvttss2si / cvttsd2si): truncation
// toward zero, and INT_MIN for NaN or values outside the int32 range. The GPU conversion saturates
// instead, so the out-of-range case is spelled out.
__device__ __forceinline__ int f2i_x86(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000;
}
__device__ __forceinline__ int d2i_x86(double v) {
  return (v >= -2147483648.0 && v < 2147483648.0) ? (int)v : (int)0x80000000;
}

// ---------------------------------------------------------------- CLUE2D binning
// src/serial/DataFormats/LayerTilesConstants.h:7-27 and LayerTilesSerial.h:24-52:
// 100 layers of 100 x 100 tiles of 5 cm over [-250, 250)^2; rX = 100 / 500.f in float.
constexpr int k2dLayers = 100;
constexpr int k2dCols = 100;
constexpr int k2dRows = 100;
constexpr int k2dBinsPerLayer = k2dCols * k2dRows;
constexpr int k2dKeys = k2dLayers * k2dBinsPerLayer;  // key = layer*10^4 + xBin*100 + yBin (x-major)

// clamp(f2i_x86(v), 0, hi) in fewer instructions: NaN and v >= 2^31 convert to INT_MIN on x86 and
// clamp to 0; every other v is clamped to [0, hi] before the (then exact) truncation.
__device__ __forceinline__ int clamp_bin(float v, int hi) {
  return v < 2147483648.0f ? (int)fminf(fmaxf(v, 0.f), (float)hi) : 0;
}
__device__ __forceinline__ int xbin2d(float x) {
  const float rX = (float)k2dCols / (250.f - (-250.f));
  return clamp_bin((x - (-250.f)) * rX, k2dCols - 1);
}
__device__ __forceinline__ int ybin2d(float y) {
  const float rY = (float)k2dRows / (250.f - (-250.f));
  return clamp_bin((y - (-250.f)) * rY, k2dRows - 1);
}

// ---------------------------------------------------------------- CLUE3D binning
// src/serial/DataFormats/Common.h:9-16, TICLLayerTile.h:15-54, Math/deltaPhi.h:17-24.
constexpr int k3dLayers = 94;
constexpr int k3dEta = 34;
constexpr int k3dPhi = 126;
constexpr int k3dBinsPerLayer = k3dEta * k3dPhi;
// Flat key = layer*4284 + etaBin*126 + phiBin, plus one spare row of 126 always-empty bins after
// the last layer: a window row etaBin == 34 addresses (layer+1, eta 0, phi), as the reference's
// contiguous tile array does, and reads empty bins past layer 93.
constexpr int k3dKeys = k3dLayers * k3dBinsPerLayer + k3dPhi;
constexpr int k3dLastPerSide = k3dLayers / 2;  // 47

__device__ __forceinline__ float reduce_range(float x) {
  const float o2pi = (float)(1. / (2. * 3.14159265358979323846));
  if (fabsf(x) <= (float)3.14159265358979323846) return x;
  const float n = roundf(x * o2pi);
  return x - n * (float)(2. * 3.14159265358979323846);
}
__device__ __forceinline__ int eta_bin3d(float eta) {
  const float r = (float)k3dEta / (3.2f - 1.5f);
  return clamp_bin((fabsf(eta) - 1.5f) * r, k3dEta - 1);
}
__device__ __forceinline__ int phi_bin3d(float phi) {
  const float normPhi = reduce_range(phi);
  const float r = (float)(k3dPhi * 0.318309886183790671538 * 0.5f);
  return d2i_x86(((double)normPhi + 3.14159265358979323846) * (double)r);
}

// ---------------------------------------------------------------- workgroup placement
// Workgroups are handed to the 8 XCDs of the chip round-robin (workgroup b runs on XCD b % 8), and
// each XCD has its own L2. xcd_block() renumbers the grid so that XCD x works through one contiguous
// range of logical blocks - neighbouring sort blocks, hence the same events' data, share an L2.
// A bijection of [0, G).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int x = b % kXcds, k = b / kXcds;
  const int q = G / kXcds, r = G % kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// ---------------------------------------------------------------- wave / block primitives
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Mask of the lanes of this wave whose `digit` (< 128) equals this lane's digit (every lane of the
// wave active). When all lanes hold lane 0's digit, as for keys that arrive grouped, one ballot
// answers; otherwise each of the digit's seven bits is matched with a ballot.
__device__ __forceinline__ uint64_t match_digit7(int digit) {
  const uint64_t same = __ballot(digit == __shfl(digit, 0, kWave));
  if (same == ~0ull) return same;
  uint64_t m = ~0ull;
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    const uint64_t bal = __ballot((digit >> b) & 1);
    m &= ((digit >> b) & 1) ? bal : ~bal;
  }
  return m;
}

__device__ __forceinline__ int wave_inclusive_sum(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  return v;
}

// Exclusive prefix sum over the 256 threads of a block; `total` receives the block sum.
// `s_tmp` needs kWaves ints. Contains two barriers: call from uniform control flow.
__device__ __forceinline__ int block_exclusive_sum(int v, int* s_tmp, int& total) {
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int inc = wave_inclusive_sum(v);
  if (lane == kWave - 1) s_tmp[wave] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const int s = s_tmp[w];
    base += (w < wave) ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

__device__ __forceinline__ int wave_inclusive_min(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int o = __shfl_up(v, off, kWave);
    if (lane >= off) v = o < v ? o : v;
  }
  return v;
}

// Exclusive prefix min over the block (identity INT_MAX).
__device__ __forceinline__ int block_exclusive_min(int v, int* s_tmp) {
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int inc = wave_inclusive_min(v);
  if (lane == kWave - 1) s_tmp[wave] = inc;
  __syncthreads();
  int base = 0x7fffffff;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const int s = s_tmp[w];
    if (w < wave) base = s < base ? s : base;
  }
  __syncthreads();
  int excl = __shfl_up(inc, 1, kWave);
  if (lane == 0) excl = 0x7fffffff;
  return excl < base ? excl : base;
}

// Block-wide [min lo, max hi).
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

// Copy [lo, hi) (1 <= hi - lo <= STAGE points) into LDS, all loads of a thread in flight together
// (indices past the end re-read the last point and are not stored).
template <int STAGE>
__device__ __forceinline__ void stage_range(const float4* __restrict__ g, float4* st, int lo, int hi) {
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
    if (k < n) st[k] = v[u];
  }
}

}  // namespace clue
