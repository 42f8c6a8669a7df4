// Stream-ordered caching allocator of a context: the role of the reference's CachingDeviceAllocator
// (src/cuda/CUDACore/CachingDeviceAllocator.h, configured in getCachingDeviceAllocator.h:14-94),
// sized for a 288 GB device.
//
// Sizes round up to power-of-two bins from 2^8 B to 2^34 B (16 GiB); a larger request gets an
// exact block that is freed, not cached, when given back. A block given back keeps the stream it was
// allocated on and an event recorded there; a later request for the same bin takes it at once on
// that stream (stream order protects it) and on another stream once the event has completed,
// otherwise a new block is allocated. Every block counts against the context's HBM budget; when a
// new block would exceed the budget, or hipMalloc fails, the idle blocks are released first and the
// allocation retried. Idle bytes above the cache limit (the budget if one is set, else 80 % of the
// device memory free when the cache was first used) are released at once. Like the rest of a
// context, the cache is thread-compatible, not thread-safe (one context per EDM stream).
#include <hip/hip_runtime.h>

#include <map>
#include <unordered_map>

#include "context.h"

namespace clue {
namespace detail {

struct CacheBlock {
  void* ptr = nullptr;
  size_t bytes = 0;
  int bin = -1;  // -1: exact block above the largest bin
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;  // recorded on `stream` when the block was given back
};

struct DeviceCache {
  std::multimap<int, CacheBlock> idle;
  std::unordered_map<void*, CacheBlock> live;
  uint64_t liveBytes = 0, idleBytes = 0, limit = 0;
  int64_t hits = 0, misses = 0;
};

namespace {

constexpr int kMinBin = 8, kMaxBin = 34;

int bin_of(uint64_t bytes) {
  int b = kMinBin;
  while (b <= kMaxBin && (uint64_t(1) << b) < bytes) ++b;
  return b <= kMaxBin ? b : -1;
}

DeviceCache& cache_of(clue_ctx* c) {
  if (!c->cache) {
    c->cache = new DeviceCache();
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) freeB = 0;
    c->cache->limit = c->budget ? c->budget : static_cast<uint64_t>(0.8 * static_cast<double>(freeB));
  }
  return *c->cache;
}

// Frees one idle block after its last use on the device has finished.
void drop(clue_ctx* c, CacheBlock& b) {
  if (b.ready) {
    (void)hipEventSynchronize(b.ready);
    (void)hipEventDestroy(b.ready);
  }
  (void)dev_free(c, b.ptr);
}

void release_idle(clue_ctx* c, DeviceCache& dc) {
  for (auto& kv : dc.idle) drop(c, kv.second);
  dc.idle.clear();
  dc.idleBytes = 0;
}

}  // namespace

void cache_destroy(clue_ctx* c) {
  if (!c->cache) return;
  for (auto& kv : c->cache->idle)
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  for (auto& kv : c->cache->live)
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  delete c->cache;  // the blocks themselves are freed with the context's other allocations
  c->cache = nullptr;
}

}  // namespace detail
}  // namespace clue

using namespace clue::detail;

void* clue_cache_alloc(clue_ctx* c, uint64_t bytes, void* stream) {
  if (!c) return nullptr;
  (void)hipSetDevice(c->device);
  DeviceCache& dc = cache_of(c);
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
  const int bin = bin_of(bytes ? bytes : 1);
  const uint64_t size = bin >= 0 ? (uint64_t(1) << bin) : bytes;
  if (bin >= 0) {
    auto range = dc.idle.equal_range(bin);
    auto take = range.second;
    for (auto it = range.first; it != range.second; ++it) {
      if (it->second.stream == s || hipEventQuery(it->second.ready) == hipSuccess) {
        take = it;
        if (it->second.stream == s) break;  // prefer the caller's own stream
      }
    }
    if (take != range.second) {
      CacheBlock b = take->second;
      dc.idle.erase(take);
      dc.idleBytes -= b.bytes;
      b.stream = s;
      dc.live[b.ptr] = b;
      dc.liveBytes += b.bytes;
      ++dc.hits;
      return b.ptr;
    }
  }
  CacheBlock b;
  b.bytes = size;
  b.bin = bin;
  b.stream = s;
  int rc = dev_alloc(c, &b.ptr, size);
  if (rc != CLUE_OK && !dc.idle.empty()) {
    release_idle(c, dc);
    rc = dev_alloc(c, &b.ptr, size);
  }
  if (rc != CLUE_OK) return nullptr;
  dc.live[b.ptr] = b;
  dc.liveBytes += b.bytes;
  ++dc.misses;
  return b.ptr;
}

int32_t clue_cache_free(clue_ctx* c, void* ptr) {
  if (!c) return CLUE_ERR_ARG;
  if (!ptr) return CLUE_OK;
  if (!c->cache) return set_err(c, CLUE_ERR_ARG, "clue_cache_free: pointer not from clue_cache_alloc");
  DeviceCache& dc = *c->cache;
  auto it = dc.live.find(ptr);
  if (it == dc.live.end()) return set_err(c, CLUE_ERR_ARG, "clue_cache_free: pointer not from clue_cache_alloc");
  CacheBlock b = it->second;
  dc.live.erase(it);
  dc.liveBytes -= b.bytes;
  if (b.bin < 0 || dc.idleBytes + b.bytes > dc.limit) {
    HIP_TRY(c, hipStreamSynchronize(b.stream));
    if (b.ready) (void)hipEventDestroy(b.ready);
    return dev_free(c, b.ptr);
  }
  if (!b.ready) HIP_TRY(c, hipEventCreateWithFlags(&b.ready, hipEventDisableTiming));
  HIP_TRY(c, hipEventRecord(b.ready, b.stream));
  dc.idle.emplace(b.bin, b);
  dc.idleBytes += b.bytes;
  return CLUE_OK;
}

int32_t clue_cache_release(clue_ctx* c) {
  if (!c) return CLUE_ERR_ARG;
  if (!c->cache) return CLUE_OK;
  release_idle(c, *c->cache);
  return CLUE_OK;
}

int32_t clue_cache_stats(clue_ctx* c, uint64_t* live_bytes, uint64_t* idle_bytes, int64_t* hits, int64_t* misses) {
  if (!c) return CLUE_ERR_ARG;
  DeviceCache& dc = cache_of(c);
  if (live_bytes) *live_bytes = dc.liveBytes;
  if (idle_bytes) *idle_bytes = dc.idleBytes;
  if (hits) *hits = dc.hits;
  if (misses) *misses = dc.misses;
  return CLUE_OK;
}
