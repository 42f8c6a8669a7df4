// Internal: the context object behind clue_ctx and the allocation helpers shared by the C ABI and
// the serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>

#include "../../include/clue_hip.h"

namespace clue {
namespace detail {

constexpr int kMarks = 16;
struct DeviceCache;

struct Workspace {
  void* base = nullptr;
  size_t bytes = 0;
};

// Pinned host buffers that the asynchronous batch calls fill with a batch's block table and
// offsets before copying them to the device on the caller's stream. A buffer is reused once the
// copy that last read it has run (its event), so kTableRing calls can be in flight before a call
// waits for the device.
constexpr int kTableRing = 4;
struct TableRing {
  void* buf[kTableRing] = {};
  size_t bytes[kTableRing] = {};
  hipEvent_t read[kTableRing] = {};
  bool pending[kTableRing] = {};
  int next = 0;
};

}  // namespace detail
}  // namespace clue

struct clue_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t inUse = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int* dErr = nullptr;  // device error flag word
  hipEvent_t marks[clue::detail::kMarks] = {};
  clue::detail::Workspace scratch2d, scratch3d, hostio;  // run_* workspaces and host-path I/O buffers
  clue::detail::Workspace scratchF2d;                     // run_* workspace of the fused CLUE2D path
  clue::detail::TableRing tables;                         // pinned staging of batch block tables
  int path2d = CLUE_PATH_GLOBAL;  // CLUE_PATH_AUTO resolves to the global-sort path
  int cus = 256;                                          // compute units of the device
  hipStream_t scratchStream = nullptr;                    // stream of the last run_* that used the scratch
  std::unordered_map<void*, size_t> allocs;
  clue::detail::DeviceCache* cache = nullptr;  // clue_cache_* blocks (created on first use)
};

namespace clue {
namespace detail {

int set_err(clue_ctx* c, int code, const std::string& msg);
int hip_fail(clue_ctx* c, hipError_t e, const char* what);
// Device allocation counted against the context's HBM budget (CLUE_ERR_BUDGET when over it).
int dev_alloc(clue_ctx* c, void** p, size_t bytes);
int dev_free(clue_ctx* c, void* p);
// Grow w to at least `bytes` (synchronises the device before freeing the old buffer).
int ensure(clue_ctx* c, Workspace& w, size_t bytes);
// Frees the caching allocator's bookkeeping (its blocks are among the context's allocations).
void cache_destroy(clue_ctx* c);

}  // namespace detail
}  // namespace clue

#define HIP_TRY(ctx, expr)                                                 \
  do {                                                                     \
    hipError_t _e = (expr);                                                \
    if (_e != hipSuccess) return clue::detail::hip_fail((ctx), _e, #expr); \
  } while (0)
