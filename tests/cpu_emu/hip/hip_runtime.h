// CPU emulation of the parts of the HIP device model the library's kernels use, so that the kernel
// sources build with g++ and run on the host (one OS thread per lane, real barriers, wave-wide
// ballots and shuffles through a per-wave mailbox) under AddressSanitizer. Test infrastructure
// only: it checks indexing, synchronisation structure and results of the device code on the CPU.
//
// Model: workgroups run one after another (so `__shared__` maps to function-local statics, one
// copy per kernel instantiation); a workgroup's threads run concurrently; __syncthreads is a
// barrier over the workgroup's live threads; a wave is 64 consecutive threads.
#pragma once

#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define CLUE_WAVES_PER_EU(...)
#ifdef EMU_LDS_REDZONES
// Every `__shared__` array a plain static: AddressSanitizer puts a redzone after each one, so an
// index past the end of an LDS array stops the run here (on the device such an access is dropped
// or reads another array, silently). LDS is not poisoned in this build.
#define __shared__ static
#else
// Every `__shared__` array lives in one linker section, which launch() fills with a poison byte
// before each workgroup: LDS starts with arbitrary contents on the device, so code that reads LDS
// before writing it must not pass here by reading zeros or a previous workgroup's values.
// (AddressSanitizer leaves globals of a named section without redzones.)
#define __shared__ static __attribute__((section("emu_lds")))
extern "C" char __start_emu_lds[];
extern "C" char __stop_emu_lds[];
#endif
#define __align__(n) __attribute__((aligned(n)))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
};
typedef void* hipStream_t;
typedef int hipError_t;

struct float2 {
  float x, y;
};
struct float4 {
  float x, y, z, w;
};
struct int2 {
  int x, y;
};
struct int4 {
  int x, y, z, w;
};
struct uint2 {
  unsigned x, y;
};
struct uint4 {
  unsigned x, y, z, w;
};
inline float2 make_float2(float x, float y) { return {x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline int2 make_int2(int x, int y) { return {x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return {x, y, z, w}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return {x, y, z, w}; }

inline float __int_as_float(int v) {
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}
inline int __float_as_int(float v) {
  int i;
  std::memcpy(&i, &v, 4);
  return i;
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

namespace emu {

constexpr int kWave = 64;

struct Wave {
  std::unique_ptr<std::barrier<>> bar;
  int64_t mail[kWave];
};

struct Group {
  std::unique_ptr<std::barrier<>> bar;
  std::vector<Wave> waves;
};

inline thread_local dim3 tIdx, bIdx;
inline dim3 bDim, gDim;
inline Group* group = nullptr;

inline Wave& my_wave() { return group->waves[tIdx.x / kWave]; }

// LDS poison byte (environment EMU_LDS_POISON, default 0x5A).
inline int lds_poison() {
  static const int v = [] {
    const char* e = std::getenv("EMU_LDS_POISON");
    return e ? (int)std::strtol(e, nullptr, 0) & 0xff : 0x5A;
  }();
  return v;
}

// Run body() over grid x block threads, one workgroup at a time; the block's threads are created
// once per launch and step through the workgroups together.
template <class F>
void launch(unsigned grid, unsigned block, F&& body) {
  gDim = dim3{grid, 1, 1};
  bDim = dim3{block, 1, 1};
  if (grid == 0) return;
  std::barrier<> start(block + 1), done(block + 1);
  std::vector<std::thread> ts;
  ts.reserve(block);
  for (unsigned t = 0; t < block; ++t)
    ts.emplace_back([&, t] {
      tIdx = dim3{t, 0, 0};
      for (unsigned b = 0; b < grid; ++b) {
        start.arrive_and_wait();
        bIdx = dim3{b, 0, 0};
        body();
        // a thread that has left the kernel no longer takes part in barriers
        my_wave().bar->arrive_and_drop();
        group->bar->arrive_and_drop();
        done.arrive_and_wait();
      }
    });
  for (unsigned b = 0; b < grid; ++b) {
    Group g;
    g.bar = std::make_unique<std::barrier<>>(block);
    const unsigned nw = (block + kWave - 1) / kWave;
    g.waves.resize(nw);
    for (unsigned w = 0; w < nw; ++w) {
      const unsigned lanes = (w + 1) * kWave <= block ? kWave : block - w * kWave;
      g.waves[w].bar = std::make_unique<std::barrier<>>(lanes);
    }
    group = &g;
#ifndef EMU_LDS_REDZONES
    std::memset(__start_emu_lds, lds_poison(), (size_t)(__stop_emu_lds - __start_emu_lds));
#endif
    start.arrive_and_wait();
    done.arrive_and_wait();
    group = nullptr;
  }
  for (auto& t : ts) t.join();
}

template <class T>
T exchange(T v, int srcLane) {
  Wave& w = my_wave();
  const int lane = (int)(tIdx.x % kWave);
  int64_t bits = 0;
  std::memcpy(&bits, &v, sizeof(T));
  w.mail[lane] = bits;
  w.bar->arrive_and_wait();
  T out;
  const int64_t got = w.mail[srcLane];
  std::memcpy(&out, &got, sizeof(T));
  w.bar->arrive_and_wait();
  return out;
}

}  // namespace emu

#define threadIdx emu::tIdx
#define blockIdx emu::bIdx
#define blockDim emu::bDim
#define gridDim emu::gDim

inline void __syncthreads() { emu::group->bar->arrive_and_wait(); }
inline unsigned __lane_id() { return emu::tIdx.x % emu::kWave; }

inline unsigned long long __ballot(int pred) {
  emu::Wave& w = emu::my_wave();
  const int lane = (int)__lane_id();
  w.mail[lane] = pred ? 1 : 0;
  w.bar->arrive_and_wait();
  unsigned long long m = 0;
  for (int l = 0; l < emu::kWave; ++l)
    if (w.mail[l]) m |= 1ull << l;
  w.bar->arrive_and_wait();
  return m;
}

template <class T>
T __shfl_up(T v, unsigned off, int width = 64) {
  (void)width;
  const int lane = (int)__lane_id();
  const T o = emu::exchange(v, lane >= (int)off ? lane - (int)off : lane);
  return lane >= (int)off ? o : v;
}
template <class T>
T __shfl_down(T v, unsigned off, int width = 64) {
  (void)width;
  const int lane = (int)__lane_id();
  const T o = emu::exchange(v, lane + (int)off < emu::kWave ? lane + (int)off : lane);
  return lane + (int)off < emu::kWave ? o : v;
}
template <class T>
T __shfl_xor(T v, int mask, int width = 64) {
  (void)width;
  return emu::exchange(v, (int)(__lane_id() ^ (unsigned)mask));
}
template <class T>
T __shfl(T v, int src, int width = 64) {
  (void)width;
  return emu::exchange(v, src);
}

inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicMax(int* p, int v) {
  int cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
inline int atomicMin(int* p, int v) {
  int cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur > v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
