// Per-module-instance handle to a clue_ctx, created on the first event on the device the event's
// stream maps to (streamID % number of devices, as the reference's chooseDevice does,
// src/cuda/CUDACore/chooseDevice.cc:5-13). C ABI error codes become std::runtime_error, the
// reference's error convention (src/cuda/CUDACore/cudaCheck.h:20-36).
#ifndef CLUE_PLUGINS_ClueContext_h
#define CLUE_PLUGINS_ClueContext_h

#include <cstddef>
#include <stdexcept>
#include <string>

#include "clue_hip.h"

class ClueContext {
public:
  ClueContext() = default;
  ClueContext(ClueContext const&) = delete;
  ClueContext& operator=(ClueContext const&) = delete;
  ~ClueContext() {
    if (ctx_) clue_destroy(ctx_);
  }

  clue_ctx* get(int streamId) {
    if (!ctx_) {
      const int n = clue_device_count();
      if (n <= 0) throw std::runtime_error("CLUE HIP backend: no HIP device available");
      ctx_ = clue_create(streamId % n, 0);
      if (!ctx_) throw std::runtime_error("CLUE HIP backend: clue_create failed on device " + std::to_string(streamId % n));
    }
    return ctx_;
  }

  void check(int rc, char const* what) const {
    if (rc != CLUE_OK) throw std::runtime_error(std::string(what) + ": " + clue_last_error(ctx_));
  }

private:
  clue_ctx* ctx_ = nullptr;
};

// Pinned host and device I/O buffers of one module instance, grown to the largest event seen and
// reused. An event's input columns are packed into the pinned buffer and moved to the device with
// one copy, its output columns come back with one copy, all on the context's stream, so the host
// never waits for the device inside acquire(). Columns are padded to 256 bytes.
class ClueStaging {
public:
  ClueStaging() = default;
  ClueStaging(ClueStaging const&) = delete;
  ClueStaging& operator=(ClueStaging const&) = delete;
  ~ClueStaging() { release(); }

  static size_t column_bytes(size_t n) { return (n * 4 + 255) & ~static_cast<size_t>(255); }

  // Both buffers at least `bytes` long. The previous event's copies must have finished (an EDM
  // stream's next acquire() follows its previous produce(), which waited for them).
  void reserve(ClueContext const& cc, clue_ctx* ctx, size_t bytes) {
    if (bytes <= cap_) return;
    release();
    ctx_ = ctx;
    host_ = static_cast<char*>(clue_malloc_pinned(ctx, bytes));
    if (!host_) cc.check(CLUE_ERR_HIP, "pinned staging buffer");
    dev_ = static_cast<char*>(clue_malloc_device(ctx, bytes));
    if (!dev_) cc.check(CLUE_ERR_HIP, "device staging buffer");
    cap_ = bytes;
  }

  char* host(size_t column, size_t n) const { return host_ + column * column_bytes(n); }
  char* device(size_t column, size_t n) const { return dev_ + column * column_bytes(n); }

private:
  void release() {
    if (host_) clue_free_pinned(ctx_, host_);
    if (dev_) clue_free_device(ctx_, dev_);
    host_ = dev_ = nullptr;
    cap_ = 0;
  }

  clue_ctx* ctx_ = nullptr;
  char* host_ = nullptr;
  char* dev_ = nullptr;
  size_t cap_ = 0;
};

#endif
