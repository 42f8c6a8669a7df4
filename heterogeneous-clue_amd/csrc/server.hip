// Serving loop: continuous batching of CLUE2D and CLUE3D events on one GPU.
//
// Requests (one event each, caller-owned host buffers) are queued by clue_server_submit*. A worker
// thread owns a fixed set of slots; a slot is one batch in flight with its own HIP stream, pinned
// staging memory and one device arena (SoA inputs and outputs, workspaces and per-event error words
// of its batch), kept resident. The arenas share the context's HBM budget: each slot may hold up to
// budget / slots bytes (without a budget, what its largest batch of max_events events and
// max_points points per dimension needs), and an arena grows to that share only when a batch needs
// it. Whenever a slot is free and requests are queued, the worker admits the queue head in arrival
// order while the batch's device bytes fit the slot's share (and its event and point counts fit the
// staging), packs the batch into the slot's pinned staging, and enqueues copy-in, the batched
// kernels and copy-out on the slot's stream. Requests that arrive meanwhile wait for the next free
// slot, so with two or more slots a new batch starts while earlier ones still run and their work
// overlaps on the device. When a slot's completion event fires the worker copies each result to its
// caller's buffers, reads the event's device error flags, and timestamps it; latency = completion
// time - submission time.
//
// Every event's arithmetic depends only on its own points (blocks never straddle events), so a
// result is bit-identical whether the event ran alone or in any mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "clue2d.h"
#include "clue3d.h"
#include "context.h"

using namespace clue;
using namespace clue::detail;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  int64_t id = 0;
  int dim = 2;
  int32_t n = 0;
  const void* in[8] = {};
  void* out[7] = {};
  Clock::time_point tSubmit, tDone;
  int state = 0;          // 0 queued, 1 running, 2 done, < 0 error code
  bool pinnedIn = false;  // every input buffer is pinned host memory: copied by DMA directly
  bool pinnedOut = false;  // every output buffer is pinned host memory
};

// Host memcpys of a batch (pageable caller buffers <-> pinned staging), split into pieces of at
// most kPiece bytes and run by a few threads plus the caller. add() queues copies; run_all() arms
// the pool, helps, and returns once every queued copy is done. Only one thread (the server's
// worker) uses a pool.
class CopyPool {
 public:
  explicit CopyPool(int threads) {
    for (int t = 0; t < threads; ++t) workers_.emplace_back([this] { run(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  void add(void* dst, const void* src, size_t bytes) {
    std::lock_guard<std::mutex> g(m_);
    for (size_t o = 0; o < bytes; o += kPiece)
      tasks_.push_back({static_cast<char*>(dst) + o, static_cast<const char*>(src) + o,
                        bytes - o < kPiece ? bytes - o : kPiece});
  }
  void run_all() {
    {
      std::lock_guard<std::mutex> g(m_);
      if (tasks_.empty()) return;
      next_ = 0;
      pending_ = tasks_.size();
      armed_ = true;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return pending_ == 0; });
    armed_ = false;
    tasks_.clear();
  }

 private:
  struct Task {
    char* dst;
    const char* src;
    size_t bytes;
  };
  static constexpr size_t kPiece = 1u << 22;
  // Takes and runs tasks until none is left to take.
  void work() {
    for (;;) {
      Task t;
      {
        std::lock_guard<std::mutex> g(m_);
        if (!armed_ || next_ >= tasks_.size()) return;
        t = tasks_[next_++];
      }
      std::memcpy(t.dst, t.src, t.bytes);
      std::lock_guard<std::mutex> g(m_);
      if (--pending_ == 0) done_.notify_all();
    }
  }
  void run() {
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || (armed_ && next_ < tasks_.size()); });
      if (stop_) return;
      lk.unlock();
      work();
      lk.lock();
    }
  }
  std::vector<std::thread> workers_;
  std::vector<Task> tasks_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  size_t next_ = 0, pending_ = 0;
  bool armed_ = false, stop_ = false;
};

bool is_pinned(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory is reported as an error by some runtimes
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

constexpr int kIn2 = 4, kOut2 = 5, kIn3 = 8, kOut3 = 7;

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  Workspace arena;        // device: the batch's SoA, workspaces and error words (batch_layout)
  char* stage = nullptr;  // pinned host staging
  size_t stageBytes = 0;
  std::vector<Req*> reqs2, reqs3;
  std::vector<int64_t> offs2, offs3;
  bool busy = false;
};

}  // namespace

struct clue_server {
  clue_ctx* ctx = nullptr;
  Clue2DParams p2{};
  Clue3DParams p3{};
  int64_t cap2 = 0, cap3 = 0;  // points per batch and dimension (pinned staging capacity)
  int capEvents = 1;
  size_t share = 0;            // device bytes one slot's batch may use
  std::vector<Slot> slots;
  std::mutex m;
  std::condition_variable cvWork, cvDone;
  std::deque<Req*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
  std::thread worker;
  bool stop = false;
  int64_t nextId = 0;
  int64_t nBatches = 0, nEvents = 0, nPoints = 0;
  std::string err;
  std::unique_ptr<CopyPool> copies;
};

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Pinned staging layout of a slot: 2D inputs and outputs, 3D inputs and outputs (field-major,
// capacity-sized), then block tables and offsets for both dimensions and the per-event error words.
struct StageLayout {
  size_t in2, out2, in3, out3, blk2, off2, blk3, off3, err, total;
};

StageLayout stage_layout(int64_t cap2, int64_t cap3, int capEvents) {
  StageLayout L;
  size_t p = 0;
  L.in2 = p;
  p += align256((size_t)cap2 * 4) * kIn2;
  L.out2 = p;
  p += align256((size_t)cap2 * 4) * kOut2;
  L.in3 = p;
  p += align256((size_t)cap3 * 4) * kIn3;
  L.out3 = p;
  p += align256((size_t)cap3 * 4) * kOut3;
  const size_t maxBlk2 = (size_t)(cap2 / kSortBlock + capEvents + 1);
  const size_t maxBlk3 = (size_t)(cap3 / kSortBlock + capEvents + 1);
  L.blk2 = p;
  p += align256(maxBlk2 * sizeof(BlockDesc));
  L.off2 = p;
  p += align256((size_t)(capEvents + 1) * 4);
  L.blk3 = p;
  p += align256(maxBlk3 * sizeof(BlockDesc));
  L.off3 = p;
  p += align256((size_t)(capEvents + 1) * 4);
  L.err = p;
  p += align256((size_t)capEvents * 2 * 4);
  L.total = p;
  return L;
}

int64_t sort_blocks(int64_t n) { return (n + kSortBlock - 1) / kSortBlock; }

// Device layout of one batch in a slot's arena: the error words of its 2D and 3D events, then per
// dimension the SoA inputs and outputs (field-major, f bytes per field) and the workspace.
struct BatchLayout {
  size_t err2, err3, in2, ws2, in3, ws3, f2, f3, total;
};

BatchLayout batch_layout(int64_t N2, int E2, int64_t nb2, int64_t N3, int E3, int64_t nb3) {
  BatchLayout L;
  size_t p = 0;
  L.err2 = p;
  p += align256((size_t)(E2 > 0 ? E2 : 1) * 4);
  L.err3 = p;
  p += align256((size_t)(E3 > 0 ? E3 : 1) * 4);
  L.f2 = align256((size_t)N2 * 4);
  L.in2 = p;
  p += L.f2 * (kIn2 + kOut2);
  L.ws2 = p;
  if (E2 > 0) p += align256(clue2d_workspace_bytes(N2, E2, (int)nb2));
  L.f3 = align256((size_t)N3 * 4);
  L.in3 = p;
  p += L.f3 * (kIn3 + kOut3);
  L.ws3 = p;
  if (E3 > 0) p += align256(clue3d_workspace_bytes(N3, E3, (int)nb3));
  L.total = p;
  return L;
}

int server_fail(clue_server* s, int code, const std::string& msg) {
  std::lock_guard<std::mutex> g(s->m);
  s->err = msg;
  return code;
}

// Inputs of a batch into the slot's device SoA (field k at dev + k*stride, event e at offs[e]):
// pinned caller buffers by DMA directly; pageable ones through the pinned staging (parallel host
// copies) and one DMA per contiguous staged run.
hipError_t copy_in(clue_server* s, Slot& sl, const std::vector<Req*>& reqs, const std::vector<int64_t>& offs,
                   int nf, char* stage, size_t sstride, char* dev, size_t stride) {
  bool staged = false;
  for (size_t e = 0; e < reqs.size(); ++e)
    if (!reqs[e]->pinnedIn) {
      staged = true;
      for (int k = 0; k < nf; ++k)
        s->copies->add(stage + k * sstride + (size_t)offs[e] * 4, reqs[e]->in[k], (size_t)reqs[e]->n * 4);
    }
  if (staged) s->copies->run_all();
  for (int k = 0; k < nf; ++k) {
    size_t e = 0;
    while (e < reqs.size()) {
      if (reqs[e]->pinnedIn) {
        hipError_t r = hipMemcpyAsync(dev + k * stride + (size_t)offs[e] * 4, reqs[e]->in[k], (size_t)reqs[e]->n * 4,
                                      hipMemcpyHostToDevice, sl.stream);
        if (r != hipSuccess) return r;
        ++e;
        continue;
      }
      size_t f = e;
      while (f < reqs.size() && !reqs[f]->pinnedIn) ++f;  // a run of staged events
      const size_t a = (size_t)offs[e] * 4, b = (size_t)offs[f] * 4;
      hipError_t r = hipMemcpyAsync(dev + k * stride + a, stage + k * sstride + a, b - a, hipMemcpyHostToDevice,
                                    sl.stream);
      if (r != hipSuccess) return r;
      e = f;
    }
  }
  return hipSuccess;
}

// Outputs of a batch: into pinned caller buffers directly, otherwise into the staging (copied to
// the caller by finish_slot).
hipError_t copy_out(Slot& sl, const std::vector<Req*>& reqs, const std::vector<int64_t>& offs, int nf, char* stage,
                    size_t sstride, char* dev, size_t stride) {
  for (int k = 0; k < nf; ++k) {
    size_t e = 0;
    while (e < reqs.size()) {
      if (reqs[e]->pinnedOut) {
        hipError_t r = hipMemcpyAsync(reqs[e]->out[k], dev + k * stride + (size_t)offs[e] * 4, (size_t)reqs[e]->n * 4,
                                      hipMemcpyDeviceToHost, sl.stream);
        if (r != hipSuccess) return r;
        ++e;
        continue;
      }
      size_t f = e;
      while (f < reqs.size() && !reqs[f]->pinnedOut) ++f;
      const size_t a = (size_t)offs[e] * 4, b = (size_t)offs[f] * 4;
      hipError_t r = hipMemcpyAsync(stage + k * sstride + a, dev + k * stride + a, b - a, hipMemcpyDeviceToHost,
                                    sl.stream);
      if (r != hipSuccess) return r;
      e = f;
    }
  }
  return hipSuccess;
}

// Pack and launch the requests assigned to a slot. Returns CLUE_OK or an error code.
int launch_slot(clue_server* s, Slot& sl) {
  clue_ctx* c = s->ctx;
  const StageLayout L = stage_layout(s->cap2, s->cap3, s->capEvents);
  const size_t sf2 = align256((size_t)s->cap2 * 4), sf3 = align256((size_t)s->cap3 * 4);
  char* st = sl.stage;
  const int E2 = (int)sl.reqs2.size(), E3 = (int)sl.reqs3.size();
  sl.offs2.assign((size_t)E2 + 1, 0);
  for (int e = 0; e < E2; ++e) sl.offs2[e + 1] = sl.offs2[e] + sl.reqs2[e]->n;
  sl.offs3.assign((size_t)E3 + 1, 0);
  for (int e = 0; e < E3; ++e) sl.offs3[e + 1] = sl.offs3[e] + sl.reqs3[e]->n;
  int64_t nb2 = 0, nb3 = 0;
  for (int e = 0; e < E2; ++e) nb2 += sort_blocks(sl.reqs2[e]->n);
  for (int e = 0; e < E3; ++e) nb3 += sort_blocks(sl.reqs3[e]->n);
  const BatchLayout B = batch_layout(sl.offs2[E2], E2, nb2, sl.offs3[E3], E3, nb3);
  if (B.total > sl.arena.bytes) {
    // the slot is idle (its previous batch finished), so its arena can be replaced at once; it grows
    // by at least half so that a few early batches settle its size within the slot's share
    size_t want = sl.arena.bytes + sl.arena.bytes / 2;
    want = want > B.total ? want : B.total;
    want = want < s->share ? want : s->share;
    if (sl.arena.base) {
      const int rc = dev_free(c, sl.arena.base);
      sl.arena.base = nullptr;
      sl.arena.bytes = 0;
      if (rc) return rc;
    }
    const int rc = dev_alloc(c, &sl.arena.base, want);
    if (rc) return rc;
    sl.arena.bytes = want;
  }
  char* dev = reinterpret_cast<char*>(sl.arena.base);
  int* err2 = reinterpret_cast<int*>(dev + B.err2);
  int* err3 = reinterpret_cast<int*>(dev + B.err3);
  fill_i32(err2, 0, (int64_t)(B.err3 - B.err2) / 4, sl.stream);
  fill_i32(err3, 0, (int64_t)(B.in2 - B.err3) / 4, sl.stream);
  // ---- 2D
  if (E2 > 0) {
    const int64_t N = sl.offs2[E2];
    char* din = dev + B.in2;
    HIP_TRY(c, copy_in(s, sl, sl.reqs2, sl.offs2, kIn2, st + L.in2, sf2, din, B.f2));
    Clue2DBuffers b;
    clue2d_carve(dev + B.ws2, N, E2, (int)nb2, b);
    clue2d_segment_bounds(b, sl.offs2.data(), E2);
    b.err = err2;
    b.errStride = 1;
    const std::vector<BlockDesc> blocks = make_blocks(sl.offs2.data(), E2);
    std::memcpy(st + L.blk2, blocks.data(), blocks.size() * sizeof(BlockDesc));
    int* off32 = reinterpret_cast<int*>(st + L.off2);
    for (int e = 0; e <= E2; ++e) off32[e] = (int)sl.offs2[e];
    HIP_TRY(c, hipMemcpyAsync(const_cast<BlockDesc*>(b.blocks), st + L.blk2, blocks.size() * sizeof(BlockDesc),
                              hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(c, hipMemcpyAsync(const_cast<int*>(b.evOff), off32, (size_t)(E2 + 1) * 4, hipMemcpyHostToDevice,
                              sl.stream));
    char* dout = din + kIn2 * B.f2;
    const size_t f = B.f2;
    Clue2DIO io{reinterpret_cast<float*>(din), reinterpret_cast<float*>(din + f),
                reinterpret_cast<int32_t*>(din + 2 * f), reinterpret_cast<float*>(din + 3 * f),
                reinterpret_cast<float*>(dout), reinterpret_cast<float*>(dout + f),
                reinterpret_cast<int32_t*>(dout + 2 * f), reinterpret_cast<int32_t*>(dout + 3 * f),
                reinterpret_cast<int32_t*>(dout + 4 * f)};
    clue2d_enqueue(b, io, s->p2, sl.stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, copy_out(sl, sl.reqs2, sl.offs2, kOut2, st + L.out2, sf2, dout, B.f2));
  }
  // ---- 3D
  if (E3 > 0) {
    const int64_t N = sl.offs3[E3];
    char* din = dev + B.in3;
    HIP_TRY(c, copy_in(s, sl, sl.reqs3, sl.offs3, kIn3, st + L.in3, sf3, din, B.f3));
    Clue3DBuffers b;
    clue3d_carve(dev + B.ws3, N, E3, (int)nb3, b);
    b.err = err3;
    b.errStride = 1;
    const std::vector<BlockDesc> blocks = make_blocks(sl.offs3.data(), E3);
    std::memcpy(st + L.blk3, blocks.data(), blocks.size() * sizeof(BlockDesc));
    int* off32 = reinterpret_cast<int*>(st + L.off3);
    for (int e = 0; e <= E3; ++e) off32[e] = (int)sl.offs3[e];
    HIP_TRY(c, hipMemcpyAsync(const_cast<BlockDesc*>(b.blocks), st + L.blk3, blocks.size() * sizeof(BlockDesc),
                              hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(c, hipMemcpyAsync(const_cast<int*>(b.evOff), off32, (size_t)(E3 + 1) * 4, hipMemcpyHostToDevice,
                              sl.stream));
    char* dout = din + kIn3 * B.f3;
    auto F = [&](char* base, int k) { return reinterpret_cast<float*>(base + k * B.f3); };
    auto I = [&](char* base, int k) { return reinterpret_cast<int32_t*>(base + k * B.f3); };
    Clue3DIO io{F(din, 0), F(din, 1), F(din, 2), F(din, 3), F(din, 4), F(din, 5), I(din, 6), I(din, 7),
                F(dout, 0), F(dout, 1), I(dout, 2), I(dout, 3), I(dout, 4), I(dout, 5), I(dout, 6)};
    clue3d_enqueue(b, io, s->p3, sl.stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, copy_out(sl, sl.reqs3, sl.offs3, kOut3, st + L.out3, sf3, dout, B.f3));
  }
  // the events' error words come back with the results
  int* herr = reinterpret_cast<int*>(st + L.err);
  if (E2 > 0) HIP_TRY(c, hipMemcpyAsync(herr, err2, (size_t)E2 * 4, hipMemcpyDeviceToHost, sl.stream));
  if (E3 > 0)
    HIP_TRY(c, hipMemcpyAsync(herr + s->capEvents, err3, (size_t)E3 * 4, hipMemcpyDeviceToHost, sl.stream));
  HIP_TRY(c, hipEventRecord(sl.done, sl.stream));
  return CLUE_OK;
}

// Status of an event from its device error word (0: none).
int event_status(int flags) {
  if (flags & kErrLayerRange) return CLUE_ERR_LAYER;
  if (flags & kErrBinRange) return CLUE_ERR_BIN;
  return flags ? CLUE_ERR_HIP : CLUE_OK;
}

// Copy the results of a finished slot to the callers' buffers and mark the requests done.
void finish_slot(clue_server* s, Slot& sl, int status) {
  const StageLayout L = stage_layout(s->cap2, s->cap3, s->capEvents);
  const size_t f2 = align256((size_t)s->cap2 * 4), f3 = align256((size_t)s->cap3 * 4);
  if (status == CLUE_OK) {
    for (size_t e = 0; e < sl.reqs2.size(); ++e)
      if (!sl.reqs2[e]->pinnedOut)
        for (int k = 0; k < kOut2; ++k)
          s->copies->add(sl.reqs2[e]->out[k], sl.stage + L.out2 + k * f2 + (size_t)sl.offs2[e] * 4,
                         (size_t)sl.reqs2[e]->n * 4);
    for (size_t e = 0; e < sl.reqs3.size(); ++e)
      if (!sl.reqs3[e]->pinnedOut)
        for (int k = 0; k < kOut3; ++k)
          s->copies->add(sl.reqs3[e]->out[k], sl.stage + L.out3 + k * f3 + (size_t)sl.offs3[e] * 4,
                         (size_t)sl.reqs3[e]->n * 4);
    s->copies->run_all();
  }
  const auto now = Clock::now();
  const int* herr = reinterpret_cast<const int*>(sl.stage + L.err);
  {
    std::lock_guard<std::mutex> g(s->m);
    for (size_t e = 0; e < sl.reqs2.size(); ++e) {
      Req* r = sl.reqs2[e];
      const int st = status == CLUE_OK ? event_status(herr[e]) : status;
      r->tDone = now;
      r->state = st == CLUE_OK ? 2 : st;
      if (st != CLUE_OK && status == CLUE_OK) s->err = "event flagged by the device: input outside the tile array";
    }
    for (size_t e = 0; e < sl.reqs3.size(); ++e) {
      Req* r = sl.reqs3[e];
      const int st = status == CLUE_OK ? event_status(herr[s->capEvents + e]) : status;
      r->tDone = now;
      r->state = st == CLUE_OK ? 2 : st;
      if (st != CLUE_OK && status == CLUE_OK) s->err = "event flagged by the device: input outside the tile array";
    }
  }
  sl.reqs2.clear();
  sl.reqs3.clear();
  sl.busy = false;
  s->cvDone.notify_all();
}

void worker_loop(clue_server* s) {
  (void)hipSetDevice(s->ctx->device);
  for (;;) {
    bool active = false;
    // completions
    for (auto& sl : s->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipSuccess) {
        finish_slot(s, sl, CLUE_OK);
      } else if (q != hipErrorNotReady) {
        server_fail(s, CLUE_ERR_HIP, std::string("batch failed: ") + hipGetErrorString(q));
        finish_slot(s, sl, CLUE_ERR_HIP);
      } else {
        active = true;
      }
    }
    // admissions: the queue head, in arrival order, into a free slot while the batch's device bytes
    // fit the slot's share of the HBM budget and its points and events fit the staging
    Slot* free = nullptr;
    bool anyBusy = false;
    for (auto& sl : s->slots) {
      anyBusy = anyBusy || sl.busy;
      if (!sl.busy && !free) free = &sl;
    }
    if (free) {
      std::unique_lock<std::mutex> lk(s->m);
      if (s->stop && s->queue.empty() && !anyBusy) break;  // drained: every batch has been finished
      int64_t used2 = 0, used3 = 0, nb2 = 0, nb3 = 0;
      while (!s->queue.empty()) {
        Req* r = s->queue.front();
        const bool d2 = r->dim == 2;
        std::vector<Req*>& dst = d2 ? free->reqs2 : free->reqs3;
        const int64_t cap = d2 ? s->cap2 : s->cap3;
        if ((int)(free->reqs2.size() + free->reqs3.size()) >= s->capEvents || (d2 ? used2 : used3) + r->n > cap) break;
        const int64_t nb = sort_blocks(r->n);
        const BatchLayout B =
            d2 ? batch_layout(used2 + r->n, (int)free->reqs2.size() + 1, nb2 + nb, used3, (int)free->reqs3.size(), nb3)
               : batch_layout(used2, (int)free->reqs2.size(), nb2, used3 + r->n, (int)free->reqs3.size() + 1, nb3 + nb);
        if (B.total > s->share) break;
        (d2 ? used2 : used3) += r->n;
        (d2 ? nb2 : nb3) += nb;
        dst.push_back(r);
        r->state = 1;
        s->queue.pop_front();
      }
      const bool any = !free->reqs2.empty() || !free->reqs3.empty();
      if (any) {
        s->nBatches += 1;
        s->nEvents += (int64_t)(free->reqs2.size() + free->reqs3.size());
        s->nPoints += used2 + used3;
      }
      lk.unlock();
      if (any) {
        free->busy = true;
        const int rc = launch_slot(s, *free);
        if (rc != CLUE_OK) {
          server_fail(s, rc, s->ctx->err);
          (void)hipStreamSynchronize(free->stream);
          finish_slot(s, *free, rc);
        }
        continue;
      }
    }
    if (active) {
      // a batch is on the device: poll its completion without sleeping (a timed sleep would add the
      // kernel timer's slack to every event's latency), letting other threads run in between
      std::this_thread::yield();
    } else {
      std::unique_lock<std::mutex> lk(s->m);
      if (s->stop && s->queue.empty()) break;  // no slot is busy here
      s->cvWork.wait_for(lk, std::chrono::milliseconds(1), [&] { return s->stop || !s->queue.empty(); });
    }
  }
}

int64_t submit(clue_server* s, int dim, int32_t n, const void* const* in, void* const* out) {
  if (!s || n < 0) return CLUE_ERR_ARG;
  const int nin = dim == 2 ? kIn2 : kIn3, nout = dim == 2 ? kOut2 : kOut3;
  for (int k = 0; k < nin; ++k)
    if (n > 0 && !in[k]) return server_fail(s, CLUE_ERR_ARG, "null input buffer");
  for (int k = 0; k < nout; ++k)
    if (n > 0 && !out[k]) return server_fail(s, CLUE_ERR_ARG, "null output buffer");
  if (n > (dim == 2 ? s->cap2 : s->cap3))
    return server_fail(s, CLUE_ERR_BUDGET, "event larger than the server's per-batch capacity");
  const BatchLayout B = dim == 2 ? batch_layout(n, 1, sort_blocks(n), 0, 0, 0) : batch_layout(0, 0, 0, n, 1, sort_blocks(n));
  if (B.total > s->share)
    return server_fail(s, CLUE_ERR_BUDGET, "event needs more device memory than a slot's share of the HBM budget");
  // Layers are checked here so that a bad event is rejected alone instead of failing its batch.
  const int32_t* layer = static_cast<const int32_t*>(in[dim == 2 ? 2 : 6]);
  const int nLayers = dim == 2 ? k2dLayers : k3dLayers;
  for (int32_t i = 0; i < n; ++i)
    if (layer[i] < 0 || layer[i] >= nLayers) return server_fail(s, CLUE_ERR_LAYER, "input layer outside the tile array");
  auto r = std::make_unique<Req>();
  r->dim = dim;
  r->n = n;
  for (int k = 0; k < nin; ++k) r->in[k] = in[k];
  for (int k = 0; k < nout; ++k) r->out[k] = out[k];
  r->pinnedIn = true;
  for (int k = 0; k < nin && n > 0; ++k) r->pinnedIn = r->pinnedIn && is_pinned(in[k]);
  r->pinnedOut = true;
  for (int k = 0; k < nout && n > 0; ++k) r->pinnedOut = r->pinnedOut && is_pinned(out[k]);
  r->tSubmit = Clock::now();
  int64_t id;
  {
    std::lock_guard<std::mutex> g(s->m);
    id = s->nextId++;
    r->id = id;
    if (n == 0) {
      r->state = 2;
      r->tDone = r->tSubmit;
    } else {
      s->queue.push_back(r.get());
    }
    s->reqs.emplace(id, std::move(r));
  }
  s->cvWork.notify_one();
  return id;
}

}  // namespace

extern "C" {

clue_server* clue_server_create(clue_ctx* c, const clue2d_params* p2, const clue3d_params* p3,
                                int64_t max_points_2d, int64_t max_points_3d, int32_t max_events, int32_t slots) {
  if (!c || !p2 || !p3 || max_points_2d < 1 || max_points_3d < 1 || max_events < 1 || slots < 1 ||
      max_points_2d >= (int64_t)0x7fffffff || max_points_3d >= (int64_t)0x7fffffff) {
    if (c) set_err(c, CLUE_ERR_ARG, "bad server configuration");
    return nullptr;
  }
  auto s = std::make_unique<clue_server>();
  s->ctx = c;
  s->p2 = make_clue2d_params(p2->dc, p2->rhoc, p2->outlierDeltaFactor);
  s->p3 = Clue3DParams{p3->densitySiblingLayers, p3->densityXYDistanceSqr, p3->kernelDensityFactor,
                       p3->densityOnSameLayer, p3->nearestHigherOnSameLayer, p3->criticalXYDistance,
                       p3->criticalZDistanceLyr, p3->criticalDensity, p3->criticalSelfDensity,
                       p3->outlierMultiplier};
  s->cap2 = max_points_2d;
  s->cap3 = max_points_3d;
  s->capEvents = max_events;
  s->slots.resize((size_t)slots);
  const StageLayout L = stage_layout(s->cap2, s->cap3, s->capEvents);
  // the bytes of the largest batch the capacities allow: max_events events holding max_points points
  // per dimension (each event adds at most one sort block beyond its points' share)
  const BatchLayout full = batch_layout(s->cap2, s->capEvents, s->cap2 / kSortBlock + s->capEvents + 1, s->cap3,
                                        s->capEvents, s->cap3 / kSortBlock + s->capEvents + 1);
  if (c->budget) {
    const uint64_t left = c->budget > c->inUse ? c->budget - c->inUse : 0;
    s->share = (size_t)(left / (uint64_t)slots) & ~(size_t)255;
  } else {
    s->share = full.total;
  }
  const size_t first = full.total < s->share ? full.total : s->share;
  bool ok = s->share > 0;
  for (auto& sl : s->slots) {
    ok = ok && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess;
    ok = ok && ensure(c, sl.arena, first) == CLUE_OK;
    if (ok) {
      void* p = nullptr;
      ok = hipHostMalloc(&p, L.total, hipHostMallocDefault) == hipSuccess;
      sl.stage = static_cast<char*>(p);
      sl.stageBytes = L.total;
    }
    if (!ok) break;
  }
  if (!ok) {
    for (auto& sl : s->slots) {
      if (sl.stage) (void)hipHostFree(sl.stage);
      if (sl.arena.base) dev_free(c, sl.arena.base);
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (c->err.empty()) set_err(c, CLUE_ERR_HIP, "server resources could not be created");
    return nullptr;
  }
  s->copies = std::make_unique<CopyPool>(7);
  clue_server* raw = s.release();
  raw->worker = std::thread(worker_loop, raw);
  return raw;
}

void clue_server_destroy(clue_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->m);
    s->stop = true;
  }
  s->cvWork.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.stage) (void)hipHostFree(sl.stage);
    if (sl.arena.base) dev_free(s->ctx, sl.arena.base);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete s;
}

int64_t clue_server_submit2d(clue_server* s, int32_t n, const float* x, const float* y, const int32_t* layer,
                             const float* weight, float* rho, float* delta, int32_t* nearestHigher, int32_t* isSeed,
                             int32_t* clusterIndex) {
  const void* in[kIn2] = {x, y, layer, weight};
  void* out[kOut2] = {rho, delta, nearestHigher, isSeed, clusterIndex};
  return submit(s, 2, n, in, out);
}

int64_t clue_server_submit3d(clue_server* s, int32_t n, const float* z, const float* eta, const float* phi,
                             const float* r_over_absz, const float* radius, const float* energy, const int32_t* layer,
                             const int32_t* isSilicon, float* rho, float* delta_first, int32_t* delta_second,
                             int32_t* nh_layer, int32_t* nh_index, int32_t* isSeed, int32_t* tracksterIndex) {
  const void* in[kIn3] = {z, eta, phi, r_over_absz, radius, energy, layer, isSilicon};
  void* out[kOut3] = {rho, delta_first, delta_second, nh_layer, nh_index, isSeed, tracksterIndex};
  return submit(s, 3, n, in, out);
}

int32_t clue_server_wait(clue_server* s, int64_t id, int32_t timeout_ms, double* latency_ms) {
  if (!s) return CLUE_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) {
    s->err = "unknown request id";
    return CLUE_ERR_ARG;
  }
  Req* r = it->second.get();
  const bool done = s->cvDone.wait_for(lk, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                                       [&] { return r->state == 2 || r->state < 0; });
  if (!done) return CLUE_ERR_TIMEOUT;
  const int state = r->state;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->tDone - r->tSubmit).count();
  s->reqs.erase(it);
  return state == 2 ? CLUE_OK : state;
}

int32_t clue_server_stats(clue_server* s, int64_t* batches, int64_t* events, int64_t* points) {
  if (!s) return CLUE_ERR_ARG;
  std::lock_guard<std::mutex> g(s->m);
  if (batches) *batches = s->nBatches;
  if (events) *events = s->nEvents;
  if (points) *points = s->nPoints;
  return CLUE_OK;
}

const char* clue_server_last_error(clue_server* s) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> g(s->m);
  thread_local std::string copy;
  copy = s->err;
  return copy.c_str();
}

}  // extern "C"
