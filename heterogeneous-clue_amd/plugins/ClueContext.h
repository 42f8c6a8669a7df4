// Per-module-instance handle to a clue_ctx, created on the first event on the device the event's
// stream maps to (streamID % number of devices, as the reference's chooseDevice does,
// src/cuda/CUDACore/chooseDevice.cc:5-13). C ABI error codes become std::runtime_error, the
// reference's error convention (src/cuda/CUDACore/cudaCheck.h:20-36).
#ifndef CLUE_PLUGINS_ClueContext_h
#define CLUE_PLUGINS_ClueContext_h

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

#endif
