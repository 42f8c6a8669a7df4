// C ABI of the clusterizer: contexts, memory helpers, per-event / batch / plan entry points.
// See include/clue_hip.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/clue_hip.h"
#include "clue2d.h"
#include "clue3d.h"
#include "fused2d.h"

#include "context.h"

using namespace clue;

struct clue_plan {
  clue_ctx* ctx = nullptr;
  int dim = 2;
  clue::detail::Workspace ws;   // global-sort path workspace
  clue::detail::Workspace wsF;  // fused CLUE2D path workspace
  int route2d = -1;             // CLUE_PATH_FUSED or CLUE_PATH_GLOBAL once decided
  std::vector<int64_t> offsets;
  Fused2DBuffers f2;
  Clue2DBuffers b2;
  Clue2DIO io2{};
  Clue2DParams p2{};
  Clue3DBuffers b3;
  Clue3DIO io3{};
  Clue3DParams p3{};
  bool useGraph = false;
  bool tune = false;   // CLUE_PLAN_AUTOTUNE requested and not done yet
  bool indexPass = false;  // CLUE_PLAN_INDEX_PASS: the nearest-higher step's index-pass output form
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace clue {
namespace detail {

int set_err(clue_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(clue_ctx* c, hipError_t e, const char* what) {
  return set_err(c, CLUE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

hipStream_t pick(clue_ctx* c, void* s) { return s ? reinterpret_cast<hipStream_t>(s) : c->stream; }

int dev_alloc(clue_ctx* c, void** p, size_t bytes) {
  if (c->budget && c->inUse + bytes > c->budget)
    return set_err(c, CLUE_ERR_BUDGET, "allocation of " + std::to_string(bytes) + " bytes exceeds the HBM budget (" +
                                           std::to_string(c->inUse) + " of " + std::to_string(c->budget) +
                                           " bytes in use)");
  hipError_t e = hipMalloc(p, bytes ? bytes : 256);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
  c->inUse += bytes;
  c->allocs[*p] = bytes;
  return CLUE_OK;
}

int dev_free(clue_ctx* c, void* p) {
  if (!p) return CLUE_OK;
  auto it = c->allocs.find(p);
  if (it == c->allocs.end()) return set_err(c, CLUE_ERR_ARG, "pointer not allocated by this context");
  c->inUse -= it->second;
  c->allocs.erase(it);
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return hip_fail(c, e, "hipFree");
  return CLUE_OK;
}

int ensure(clue_ctx* c, Workspace& w, size_t bytes) {
  if (w.bytes >= bytes) return CLUE_OK;
  if (w.base) {
    HIP_TRY(c, hipDeviceSynchronize());  // the old workspace may still be read by queued kernels
    int rc = dev_free(c, w.base);
    if (rc) return rc;
    w.base = nullptr;
    w.bytes = 0;
  }
  int rc = dev_alloc(c, &w.base, bytes);
  if (rc) return rc;
  w.bytes = bytes;
  return CLUE_OK;
}

// The scratch workspace is shared by all run_* calls of a context: calls on another stream wait for
// the previous user of the scratch first.
int claim_scratch(clue_ctx* c, hipStream_t s) {
  if (c->scratchStream && c->scratchStream != s) HIP_TRY(c, hipStreamSynchronize(c->scratchStream));
  c->scratchStream = s;
  return CLUE_OK;
}

}  // namespace detail
}  // namespace clue

using namespace clue::detail;

namespace {

bool check_offsets(const int64_t* offs, int n) {
  if (!offs || n < 0 || offs[0] != 0) return false;
  for (int e = 0; e < n; ++e)
    if (offs[e + 1] < offs[e]) return false;
  return offs[n] < (int64_t)0x7fffffff;
}

// ---------------------------------------------------------------- allocation-free CLUE2D scratch
// clue2d_run_device_noalloc carves its workspace out of this fixed array instead of allocating: one
// copy per GPU, held until the process ends. In host code `scratch` names a host-side stand-in, not
// device memory, so it is never passed to a kernel or a copy; the device address of each GPU's copy
// is looked up once per device (noalloc_setup, called from clue_create with the device current)
// and only that address is used.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratchDev[kMaxDevices] = {};

int noalloc_setup(clue_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, CLUE_ERR_ARG, "device index above 63");
  std::lock_guard<std::mutex> g(g_scratchMu);
  if (g_scratchDev[c->device]) return CLUE_OK;
  void* p = nullptr;
  HIP_TRY(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return set_err(c, CLUE_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratchDev[c->device] = p;
  return CLUE_OK;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratchMu);
  return g_scratchDev[device];
}

__global__ void k_single_event(int n, int nb, BlockDesc* blocks, int* evOff, int* evBlk) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    BlockDesc d;
    d.event = 0;
    d.begin = b * kSortBlock;
    d.end = (b + 1) * kSortBlock < n ? (b + 1) * kSortBlock : n;
    d.histBase = 0;
    d.nbEvent = nb;
    d.bLocal = b;
    blocks[b] = d;
  }
  if (b == 0) {
    evOff[0] = 0;
    evOff[1] = n;
    if (evBlk) {
      evBlk[0] = 0;
      evBlk[1] = nb;
    }
  }
}

// Upload a host block table and int32 offsets into the workspace (synchronous copies).
int upload_tables(clue_ctx* c, const int64_t* offsets, int E, const BlockDesc* blocksDev, const int* evOffDev,
                  std::vector<BlockDesc>& blocks, const int* evBlkDev = nullptr) {
  blocks = make_blocks(offsets, E);
  if (evBlkDev) {
    const std::vector<int> eb = make_event_blocks(offsets, E);
    HIP_TRY(c, hipMemcpy(const_cast<int*>(evBlkDev), eb.data(), eb.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  std::vector<int> off32((size_t)E + 1);
  for (int e = 0; e <= E; ++e) off32[e] = (int)offsets[e];
  if (!blocks.empty())
    HIP_TRY(c, hipMemcpy(const_cast<BlockDesc*>(blocksDev), blocks.data(), blocks.size() * sizeof(BlockDesc),
                         hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(const_cast<int*>(evOffDev), off32.data(), off32.size() * sizeof(int), hipMemcpyHostToDevice));
  return CLUE_OK;
}

// The same tables for a call that must not wait for the device: built on the host into a pinned
// buffer of the context's ring and copied with hipMemcpyAsync on `s`, so they overwrite the shared
// workspace only after the work enqueued on `s` before them (other streams are ordered by
// claim_scratch). Waits only when the ring buffer's previous copy has not run yet.
int upload_tables_async(clue_ctx* c, const int64_t* offsets, int E, const BlockDesc* blocksDev, const int* evOffDev,
                        hipStream_t s, const int* evBlkDev = nullptr) {
  const std::vector<BlockDesc> blocks = make_blocks(offsets, E);
  const size_t bBlk = blocks.size() * sizeof(BlockDesc), bOff = ((size_t)E + 1) * sizeof(int);
  const size_t oOff = (bBlk + 255) & ~(size_t)255, oEvb = oOff + ((bOff + 255) & ~(size_t)255);
  const size_t need = oEvb + (evBlkDev ? bOff : 0);
  TableRing& R = c->tables;
  const int k = R.next;
  R.next = (k + 1) % kTableRing;
  if (R.pending[k]) {
    HIP_TRY(c, hipEventSynchronize(R.read[k]));
    R.pending[k] = false;
  }
  if (!R.read[k]) HIP_TRY(c, hipEventCreateWithFlags(&R.read[k], hipEventDisableTiming));
  if (R.bytes[k] < need) {
    if (R.buf[k]) HIP_TRY(c, hipHostFree(R.buf[k]));
    R.buf[k] = nullptr;
    R.bytes[k] = 0;
    HIP_TRY(c, hipHostMalloc(&R.buf[k], need, hipHostMallocDefault));
    R.bytes[k] = need;
  }
  char* h = static_cast<char*>(R.buf[k]);
  if (bBlk) std::memcpy(h, blocks.data(), bBlk);
  int* off32 = reinterpret_cast<int*>(h + oOff);
  for (int e = 0; e <= E; ++e) off32[e] = (int)offsets[e];
  if (bBlk)
    HIP_TRY(c, hipMemcpyAsync(const_cast<BlockDesc*>(blocksDev), h, bBlk, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(const_cast<int*>(evOffDev), off32, bOff, hipMemcpyHostToDevice, s));
  if (evBlkDev) {
    const std::vector<int> eb = make_event_blocks(offsets, E);
    std::memcpy(h + oEvb, eb.data(), bOff);
    HIP_TRY(c, hipMemcpyAsync(const_cast<int*>(evBlkDev), h + oEvb, bOff, hipMemcpyHostToDevice, s));
  }
  HIP_TRY(c, hipEventRecord(R.read[k], s));
  R.pending[k] = true;
  return CLUE_OK;
}

int nblocks_of(const int64_t* offsets, int E) {
  int64_t nb = 0;
  for (int e = 0; e < E; ++e) nb += (offsets[e + 1] - offsets[e] + kSortBlock - 1) / kSortBlock;
  return (int)nb;
}

Clue2DParams to2d(const clue2d_params* p) { return make_clue2d_params(p->dc, p->rhoc, p->outlierDeltaFactor); }

Clue3DParams to3d(const clue3d_params* p) {
  Clue3DParams q;
  q.densitySiblingLayers = p->densitySiblingLayers;
  q.densityXYDistanceSqr = p->densityXYDistanceSqr;
  q.kernelDensityFactor = p->kernelDensityFactor;
  q.densityOnSameLayer = p->densityOnSameLayer;
  q.nearestHigherOnSameLayer = p->nearestHigherOnSameLayer;
  q.criticalXYDistance = p->criticalXYDistance;
  q.criticalZDistanceLyr = p->criticalZDistanceLyr;
  q.criticalDensity = p->criticalDensity;
  q.criticalSelfDensity = p->criticalSelfDensity;
  q.outlierMultiplier = p->outlierMultiplier;
  return q;
}

int64_t max_event(const int64_t* offsets, int E) {
  int64_t m = 0;
  for (int e = 0; e < E; ++e) m = offsets[e + 1] - offsets[e] > m ? offsets[e + 1] - offsets[e] : m;
  return m;
}

int single_setup(clue_ctx* c, int n, int nb, const BlockDesc* blocks, const int* evOff, hipStream_t s,
                 const int* evBlk = nullptr) {
  k_single_event<<<(nb + kThreads - 1) / kThreads > 0 ? (nb + kThreads - 1) / kThreads : 1, kThreads, 0, s>>>(
      n, nb, const_cast<BlockDesc*>(blocks), const_cast<int*>(evOff), const_cast<int*>(evBlk));
  HIP_TRY(c, hipGetLastError());
  return CLUE_OK;
}

// Global-sort workspace and block tables of a CLUE2D plan.
int plan_setup_global(clue_plan* pl) {
  clue_ctx* c = pl->ctx;
  const int E = (int)pl->offsets.size() - 1;
  const int64_t N = pl->offsets[E];
  const int nb = nblocks_of(pl->offsets.data(), E);
  int rc = ensure(c, pl->ws, clue2d_workspace_bytes(N, E, nb));
  if (rc) return rc;
  clue2d_carve(pl->ws.base, N, E, nb, pl->b2);
  clue2d_segment_bounds(pl->b2, pl->offsets.data(), E);
  pl->b2.err = c->dErr;
  pl->b2.nhIndexPass = pl->indexPass ? 1 : 0;
  std::vector<BlockDesc> blocks;
  return upload_tables(c, pl->offsets.data(), E, pl->b2.blocks, pl->b2.evOff, blocks);
}

// First launch of a CLUE2D plan whose route is open: group by layer, read the largest layer, then
// finish this launch on the route that fits and keep it for the later launches.
int plan_route_first(clue_plan* pl, hipStream_t s) {
  clue_ctx* c = pl->ctx;
  fused2d_enqueue_segments(pl->f2, pl->io2, false, s);
  int big = 0;
  HIP_TRY(c, hipMemcpyAsync(&big, pl->f2.ctl + 8, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  if (big <= kFusedCap) {
    pl->route2d = CLUE_PATH_FUSED;
    fused2d_enqueue_cluster(pl->f2, pl->io2, pl->p2, s);
  } else {
    pl->route2d = CLUE_PATH_GLOBAL;
    int rc = plan_setup_global(pl);
    if (rc) return rc;
    clue2d_enqueue(pl->b2, pl->io2, pl->p2, s);
  }
  HIP_TRY(c, hipGetLastError());
  return CLUE_OK;
}

// CLUE2D / CLUE3D plan autotuning: one full run (so this launch produces its results), then each
// variant of the density and of the nearest-higher kernel timed in place on the same data (best of
// kTuneRounds interleaved rounds); the fastest of each is kept. Re-running either kernel rewrites its
// outputs with the same values.
constexpr int kTuneRounds = 5;
static_assert(kClue2DVariants <= 8 && kClue3DVariants <= 8, "plan_autotune keeps at most 8 timings");
int plan_autotune(clue_plan* pl, hipStream_t s) {
  clue_ctx* c = pl->ctx;
  const bool d3 = pl->dim == 3;
  if (d3)
    clue3d_enqueue(pl->b3, pl->io3, pl->p3, s);
  else
    clue2d_enqueue(pl->b2, pl->io2, pl->p2, s);
  HIP_TRY(c, hipGetLastError());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  const int nVariants = d3 ? kClue3DVariants : kClue2DVariants;
  int best[2] = {0, 0};
  for (int which = 0; which < 2; ++which) {
    float ms[8] = {};
    // a warm-up round, then kTuneRounds rounds over all variants in turn (best of each), so that a
    // slow stretch of the device weighs on every variant alike
    for (int round = 0; round <= kTuneRounds; ++round) {
      for (int v = 0; v < nVariants; ++v) {
        HIP_TRY(c, hipEventRecord(e0, s));
        if (d3)
          clue3d_enqueue_step(pl->b3, pl->io3, pl->p3, which, v, s);
        else
          clue2d_enqueue_step(pl->b2, pl->io2, pl->p2, which, v, s);
        HIP_TRY(c, hipEventRecord(e1, s));
        HIP_TRY(c, hipEventSynchronize(e1));
        float t = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&t, e0, e1));
        if (round == 1 || (round > 1 && t < ms[v])) ms[v] = t;
      }
    }
    for (int v = 1; v < nVariants; ++v)
      if (ms[v] < ms[best[which]]) best[which] = v;
  }
  if (d3) {
    pl->b3.densityVariant = best[0];
    pl->b3.nhVariant = best[1];
  } else {
    pl->b2.densityVariant = best[0];
    pl->b2.nhVariant = best[1];
    // the output form of the nearest-higher step: both forms of that kernel plus its labels step,
    // timed the same way (hits that arrive grouped by position favour the direct writes, hits in an
    // order unrelated to position the index pass)
    float ms[2] = {};
    for (int round = 0; round <= kTuneRounds; ++round) {
      for (int v = 0; v < 2; ++v) {
        HIP_TRY(c, hipEventRecord(e0, s));
        clue2d_enqueue_step(pl->b2, pl->io2, pl->p2, 2, v, s);
        HIP_TRY(c, hipEventRecord(e1, s));
        HIP_TRY(c, hipEventSynchronize(e1));
        float t = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&t, e0, e1));
        if (round == 1 || (round > 1 && t < ms[v])) ms[v] = t;
      }
    }
    pl->b2.nhIndexPass = ms[1] < ms[0] ? 1 : 0;
    pl->indexPass = pl->b2.nhIndexPass != 0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pl->tune = false;
  return CLUE_OK;
}

// CLUE2D over a batch on the context's scratch: the fused path when the context allows it and
// every (event, layer) fits, the global-sort path otherwise. `single`: one event whose block table
// is written on the device (no host copy, no synchronisation).
int run2d(clue_ctx* c, const Clue2DParams& p, int E, const int64_t* offsets, const Clue2DIO& io, hipStream_t s,
          bool single) {
  const int64_t N = offsets[E];
  const int nb = nblocks_of(offsets, E);
  const int64_t maxEv = max_event(offsets, E);
  if (c->path2d != CLUE_PATH_GLOBAL) {
    int rc = ensure(c, c->scratchF2d, fused2d_workspace_bytes(N, E, nb));
    if (rc) return rc;
    Fused2DBuffers fb;
    fused2d_carve(c->scratchF2d.base, N, E, nb, fb);
    fb.err = c->dErr;
    fb.cus = c->cus;
    if (single) {
      rc = single_setup(c, (int)N, nb, fb.blocks, fb.evOff, s, fb.evBlk);
    } else {
      rc = upload_tables_async(c, offsets, E, fb.blocks, fb.evOff, s, fb.evBlk);
    }
    if (rc) return rc;
    const bool forced = c->path2d == CLUE_PATH_FUSED;
    fused2d_enqueue_segments(fb, io, forced, s);
    bool fits = forced || maxEv <= kFusedCap;
    if (!fits) {
      int big = 0;
      HIP_TRY(c, hipMemcpyAsync(&big, fb.ctl + 8, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipStreamSynchronize(s));
      fits = big <= kFusedCap;
    }
    if (fits) {
      fused2d_enqueue_cluster(fb, io, p, s);
      HIP_TRY(c, hipGetLastError());
      return CLUE_OK;
    }
  }
  int rc = ensure(c, c->scratch2d, clue2d_workspace_bytes(N, E, nb));
  if (rc) return rc;
  Clue2DBuffers b;
  clue2d_carve(c->scratch2d.base, N, E, nb, b);
  clue2d_segment_bounds(b, offsets, E);
  b.err = c->dErr;
  if (single) {
    rc = single_setup(c, (int)N, nb, b.blocks, b.evOff, s);
  } else {
    rc = upload_tables_async(c, offsets, E, b.blocks, b.evOff, s);
  }
  if (rc) return rc;
  clue2d_enqueue(b, io, p, s);
  HIP_TRY(c, hipGetLastError());
  return CLUE_OK;
}

int launch_plan_body(clue_plan* pl, hipStream_t s) {
  if (pl->dim == 2 && pl->route2d == CLUE_PATH_FUSED) {
    fused2d_enqueue_segments(pl->f2, pl->io2, true, s);
    fused2d_enqueue_cluster(pl->f2, pl->io2, pl->p2, s);
  } else if (pl->dim == 2)
    clue2d_enqueue(pl->b2, pl->io2, pl->p2, s);
  else
    clue3d_enqueue(pl->b3, pl->io3, pl->p3, s);
  HIP_TRY(pl->ctx, hipGetLastError());
  return CLUE_OK;
}

// ---------------------------------------------------------------- arithmetic self-check kernel
__global__ void k_selftest(int n, const float* v, const float* w, int* xb, int* yb, int* eb, int* pb, float* red,
                           float* dist, float* quot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a = v[i], b = w[i];
  xb[i] = xbin2d(a);
  yb[i] = ybin2d(a);
  eb[i] = eta_bin3d(a);
  pb[i] = phi_bin3d(a);
  red[i] = reduce_range(a);
  dist[i] = sqrtf(a * a + b * b);
  quot[i] = a / b;
}

}  // namespace

extern "C" {

clue3d_params clue3d_default_params(void) {
  clue3d_params p;
  p.densitySiblingLayers = 3;
  p.densityXYDistanceSqr = 3.24f;
  p.kernelDensityFactor = 0.2f;
  p.densityOnSameLayer = 0;
  p.nearestHigherOnSameLayer = 0;
  p.criticalXYDistance = 1.8f;
  p.criticalZDistanceLyr = 5.f;
  p.criticalDensity = 0.6f;
  p.criticalSelfDensity = 0.15f;
  p.outlierMultiplier = 2.f;
  return p;
}

int32_t clue_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

clue_ctx* clue_create(int32_t device, uint64_t hbm_budget_bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  clue_ctx* c = new (std::nothrow) clue_ctx();
  if (!c) return nullptr;
  c->device = device;
  c->budget = hbm_budget_bytes;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return nullptr;
  }
  void* p = nullptr;
  if (dev_alloc(c, &p, 256) != CLUE_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return nullptr;
  }
  c->dErr = reinterpret_cast<int*>(p);
  if (noalloc_setup(c) != CLUE_OK) {
    (void)hipFree(p);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return nullptr;
  }
  fill_i32(c->dErr, 0, 64, c->stream);
  for (int k = 0; k < kMarks; ++k) (void)hipEventCreate(&c->marks[k]);
  (void)hipStreamSynchronize(c->stream);
  return c;
}

void clue_destroy(clue_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  cache_destroy(c);
  for (auto& kv : c->allocs) (void)hipFree(kv.first);
  c->allocs.clear();
  for (int k = 0; k < kMarks; ++k)
    if (c->marks[k]) (void)hipEventDestroy(c->marks[k]);
  for (int k = 0; k < kTableRing; ++k) {
    if (c->tables.buf[k]) (void)hipHostFree(c->tables.buf[k]);
    if (c->tables.read[k]) (void)hipEventDestroy(c->tables.read[k]);
  }
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int32_t clue_set_path(clue_ctx* c, int32_t dim, int32_t path) {
  if (!c) return CLUE_ERR_ARG;
  if (dim != 2 || path < CLUE_PATH_AUTO || path > CLUE_PATH_GLOBAL)
    return set_err(c, CLUE_ERR_ARG, "clue_set_path: dim must be 2 and path one of CLUE_PATH_*");
  if (path == CLUE_PATH_FUSED)
    return set_err(c, CLUE_ERR_ARG,
                   "clue_set_path: the fused CLUE2D path is disabled in this build (its kernels are verified "
                   "on the CPU emulation only); CLUE_PATH_AUTO selects the global-sort path");
  c->path2d = CLUE_PATH_GLOBAL;
  return CLUE_OK;
}

const char* clue_last_error(const clue_ctx* c) { return c ? c->err.c_str() : "null context"; }

uint64_t clue_bytes_in_use(const clue_ctx* c) { return c ? c->inUse : 0; }

void* clue_stream_create(clue_ctx* c) {
  if (!c) return nullptr;
  hipStream_t s = nullptr;
  (void)hipSetDevice(c->device);
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return reinterpret_cast<void*>(s);
}

int32_t clue_stream_destroy(clue_ctx* c, void* s) {
  if (!c || !s) return CLUE_ERR_ARG;
  hipStream_t hs = reinterpret_cast<hipStream_t>(s);
  if (c->scratchStream == hs) c->scratchStream = nullptr;
  HIP_TRY(c, hipStreamSynchronize(hs));
  HIP_TRY(c, hipStreamDestroy(hs));
  return CLUE_OK;
}

int32_t clue_stream_sync(clue_ctx* c, void* s) {
  if (!c) return CLUE_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(pick(c, s)));
  return CLUE_OK;
}

int32_t clue_stream_callback(clue_ctx* c, void* s, void (*fn)(void* user), void* user) {
  if (!c) return CLUE_ERR_ARG;
  if (!fn) return set_err(c, CLUE_ERR_ARG, "null callback");
  HIP_TRY(c, hipLaunchHostFunc(pick(c, s), fn, user));
  return CLUE_OK;
}

void* clue_malloc_device(clue_ctx* c, uint64_t bytes) {
  if (!c) return nullptr;
  void* p = nullptr;
  if (dev_alloc(c, &p, bytes) != CLUE_OK) return nullptr;
  return p;
}

int32_t clue_free_device(clue_ctx* c, void* p) {
  if (!c) return CLUE_ERR_ARG;
  HIP_TRY(c, hipDeviceSynchronize());
  return dev_free(c, p);
}

void* clue_malloc_pinned(clue_ctx* c, uint64_t bytes) {
  if (!c) return nullptr;
  void* p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes ? bytes : 64, hipHostMallocDefault);
  if (e != hipSuccess) {
    hip_fail(c, e, "hipHostMalloc");
    return nullptr;
  }
  return p;
}

int32_t clue_free_pinned(clue_ctx* c, void* p) {
  if (!c) return CLUE_ERR_ARG;
  HIP_TRY(c, hipHostFree(p));
  return CLUE_OK;
}

int32_t clue_memcpy_htod(clue_ctx* c, void* dst, const void* src, uint64_t bytes, void* s) {
  if (!c || (!dst && bytes) || (!src && bytes)) return CLUE_ERR_ARG;
  if (!bytes) return CLUE_OK;
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pick(c, s)));
  return CLUE_OK;
}

int32_t clue_memcpy_dtoh(clue_ctx* c, void* dst, const void* src, uint64_t bytes, void* s) {
  if (!c || (!dst && bytes) || (!src && bytes)) return CLUE_ERR_ARG;
  if (!bytes) return CLUE_OK;
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, pick(c, s)));
  return CLUE_OK;
}

int32_t clue_memcpy_dtod(clue_ctx* c, void* dst, const void* src, uint64_t bytes, void* s) {
  if (!c || (!dst && bytes) || (!src && bytes)) return CLUE_ERR_ARG;
  if (!bytes) return CLUE_OK;
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, pick(c, s)));
  return CLUE_OK;
}

int32_t clue_check(clue_ctx* c, void* s) {
  if (!c) return CLUE_ERR_ARG;
  hipStream_t hs = pick(c, s);
  HIP_TRY(c, hipStreamSynchronize(hs));
  int flags = 0;
  HIP_TRY(c, hipMemcpy(&flags, c->dErr, sizeof(int), hipMemcpyDeviceToHost));
  if (flags) {
    fill_i32(c->dErr, 0, 1, hs);
    HIP_TRY(c, hipStreamSynchronize(hs));
    if (flags & kErrLayerRange) return set_err(c, CLUE_ERR_LAYER, "input layer outside the tile array");
    if (flags & kErrCapacity)
      return set_err(c, CLUE_ERR_CAPACITY,
                     "an (event, layer) has more points than the fused path holds (" + std::to_string(kFusedCap) +
                         "); select CLUE_PATH_AUTO or CLUE_PATH_GLOBAL");
    return set_err(c, CLUE_ERR_BIN, "phi outside the range the tile array can bin");
  }
  return CLUE_OK;
}

int32_t clue_mark(clue_ctx* c, int32_t k, void* s) {
  if (!c || k < 0 || k >= kMarks) return CLUE_ERR_ARG;
  HIP_TRY(c, hipEventRecord(c->marks[k], pick(c, s)));
  return CLUE_OK;
}

float clue_elapsed_ms(clue_ctx* c, int32_t a, int32_t b) {
  if (!c || a < 0 || b < 0 || a >= kMarks || b >= kMarks) return -1.f;
  float ms = -1.f;
  if (hipEventElapsedTime(&ms, c->marks[a], c->marks[b]) != hipSuccess) return -1.f;
  return ms;
}

// ---------------------------------------------------------------- CLUE2D

int32_t clue2d_run_device(clue_ctx* c, const clue2d_params* p, int32_t n, const float* x, const float* y,
                          const int32_t* layer, const float* weight, float* rho, float* delta, int32_t* nh,
                          int32_t* isSeed, int32_t* clusterIndex, void* stream) {
  if (!c || !p || n < 0) return set_err(c, CLUE_ERR_ARG, "bad argument");
  if (n == 0) return CLUE_OK;
  if (!x || !y || !layer || !weight || !rho || !delta || !nh || !isSeed || !clusterIndex)
    return set_err(c, CLUE_ERR_ARG, "null buffer");
  hipStream_t s = pick(c, stream);
  int rc = claim_scratch(c, s);
  if (rc) return rc;
  const int64_t offs[2] = {0, n};
  Clue2DIO io{x, y, layer, weight, rho, delta, nh, isSeed, clusterIndex};
  return run2d(c, to2d(p), 1, offs, io, s, true);
}

int32_t clue2d_run_device_noalloc(clue_ctx* c, const clue2d_params* p, int32_t n, const float* x, const float* y,
                                  const int32_t* layer, const float* weight, float* rho, float* delta, int32_t* nh,
                                  int32_t* isSeed, int32_t* clusterIndex, void* stream) {
  if (!c || !p || n < 0) return set_err(c, CLUE_ERR_ARG, "bad argument");
  if (n == 0) return CLUE_OK;
  if (!x || !y || !layer || !weight || !rho || !delta || !nh || !isSeed || !clusterIndex)
    return set_err(c, CLUE_ERR_ARG, "null buffer");
  if (n > CLUE_NOALLOC_MAX_POINTS)
    return set_err(c, CLUE_ERR_CAPACITY, "clue2d_run_device_noalloc: " + std::to_string(n) +
                                             " points exceed its fixed scratch (at most " +
                                             std::to_string(CLUE_NOALLOC_MAX_POINTS) + "); use clue2d_run_device");
  void* base = noalloc_scratch(c->device);
  if (!base) return set_err(c, CLUE_ERR_HIP, "the scratch array of this device has no device address");
  const int nb = (n + kSortBlock - 1) / kSortBlock;
  const size_t need = clue2d_workspace_bytes(n, 1, nb);
  if (need > kNoAllocScratchBytes)
    return set_err(c, CLUE_ERR_CAPACITY, "clue2d_run_device_noalloc: workspace of " + std::to_string(need) +
                                             " bytes exceeds the fixed scratch");
  hipStream_t s = pick(c, stream);
  // Every workspace word the launch sequence reads is written earlier in the same sequence (block
  // table, histograms, scans, bin table, reach records, seed words), so the previous call's contents
  // need no clearing.
  Clue2DBuffers b;
  clue2d_carve(base, n, 1, nb, b);
  const int64_t offs[2] = {0, n};
  clue2d_segment_bounds(b, offs, 1);
  b.err = c->dErr;
  int rc = single_setup(c, n, nb, b.blocks, b.evOff, s);
  if (rc) return rc;
  Clue2DIO io{x, y, layer, weight, rho, delta, nh, isSeed, clusterIndex};
  clue2d_enqueue(b, io, to2d(p), s);
  HIP_TRY(c, hipGetLastError());
  return CLUE_OK;
}

int32_t clue2d_run_batch_device(clue_ctx* c, const clue2d_params* p, int32_t E, const int64_t* offsets,
                                const float* x, const float* y, const int32_t* layer, const float* weight, float* rho,
                                float* delta, int32_t* nh, int32_t* isSeed, int32_t* clusterIndex, void* stream) {
  if (!c || !p || !check_offsets(offsets, E)) return set_err(c, CLUE_ERR_ARG, "bad argument or offsets");
  const int64_t N = offsets[E];
  if (N == 0) return CLUE_OK;
  hipStream_t s = pick(c, stream);
  int rc = claim_scratch(c, s);
  if (rc) return rc;
  Clue2DIO io{x, y, layer, weight, rho, delta, nh, isSeed, clusterIndex};
  return run2d(c, to2d(p), E, offsets, io, s, false);
}

int32_t clue2d_run_host(clue_ctx* c, const clue2d_params* p, int32_t n, const float* x, const float* y,
                        const int32_t* layer, const float* weight, float* rho, float* delta, int32_t* nh,
                        int32_t* isSeed, int32_t* clusterIndex) {
  if (!c || !p || n < 0) return set_err(c, CLUE_ERR_ARG, "bad argument");
  if (n == 0) return CLUE_OK;
  const size_t nf = (size_t)n * 4;
  int rc = claim_scratch(c, c->stream);
  if (rc) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  rc = ensure(c, c->hostio, 9 * ((nf + 255) & ~(size_t)255));
  if (rc) return rc;
  Carver cv{reinterpret_cast<char*>(c->hostio.base)};
  float* dx = cv.take<float>(n);
  float* dy = cv.take<float>(n);
  int32_t* dl = cv.take<int32_t>(n);
  float* dw = cv.take<float>(n);
  float* drho = cv.take<float>(n);
  float* ddel = cv.take<float>(n);
  int32_t* dnh = cv.take<int32_t>(n);
  int32_t* dseed = cv.take<int32_t>(n);
  int32_t* dci = cv.take<int32_t>(n);
  hipStream_t s = c->stream;
  HIP_TRY(c, hipMemcpyAsync(dx, x, nf, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(dy, y, nf, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(dl, layer, nf, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(dw, weight, nf, hipMemcpyHostToDevice, s));
  rc = clue2d_run_device(c, p, n, dx, dy, dl, dw, drho, ddel, dnh, dseed, dci, s);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(rho, drho, nf, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(delta, ddel, nf, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(nh, dnh, nf, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(isSeed, dseed, nf, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(clusterIndex, dci, nf, hipMemcpyDeviceToHost, s));
  return clue_check(c, s);
}

clue_plan* clue2d_plan_create(clue_ctx* c, const clue2d_params* p, int32_t E, const int64_t* offsets, const float* x,
                              const float* y, const int32_t* layer, const float* weight, float* rho, float* delta,
                              int32_t* nh, int32_t* isSeed, int32_t* clusterIndex, int32_t use_graph) {
  if (!c || !p || !check_offsets(offsets, E)) {
    set_err(c, CLUE_ERR_ARG, "bad argument or offsets");
    return nullptr;
  }
  clue_plan* pl = new (std::nothrow) clue_plan();
  if (!pl) return nullptr;
  pl->ctx = c;
  pl->dim = 2;
  pl->offsets.assign(offsets, offsets + E + 1);
  pl->io2 = Clue2DIO{x, y, layer, weight, rho, delta, nh, isSeed, clusterIndex};
  pl->p2 = to2d(p);
  pl->useGraph = (use_graph & CLUE_PLAN_GRAPH) != 0;
  pl->tune = (use_graph & CLUE_PLAN_AUTOTUNE) != 0;
  pl->indexPass = (use_graph & CLUE_PLAN_INDEX_PASS) != 0;
  if (c->path2d == CLUE_PATH_GLOBAL) {
    pl->route2d = CLUE_PATH_GLOBAL;
    if (plan_setup_global(pl) != CLUE_OK) {
      clue_plan_destroy(pl);
      return nullptr;
    }
    return pl;
  }
  const int64_t N = offsets[E];
  const int nb = nblocks_of(offsets, E);
  if (ensure(c, pl->wsF, fused2d_workspace_bytes(N, E, nb)) != CLUE_OK) {
    delete pl;
    return nullptr;
  }
  fused2d_carve(pl->wsF.base, N, E, nb, pl->f2);
  pl->f2.err = c->dErr;
  pl->f2.cus = c->cus;
  std::vector<BlockDesc> blocks;
  if (upload_tables(c, offsets, E, pl->f2.blocks, pl->f2.evOff, blocks, pl->f2.evBlk) != CLUE_OK) {
    clue_plan_destroy(pl);
    return nullptr;
  }
  // The route is fixed now when no layer can exceed the fused capacity, else at the first launch.
  if (c->path2d == CLUE_PATH_FUSED || max_event(offsets, E) <= kFusedCap) pl->route2d = CLUE_PATH_FUSED;
  return pl;
}

// ---------------------------------------------------------------- CLUE3D

int32_t clue3d_run_device(clue_ctx* c, const clue3d_params* p, int32_t n, const float* z, const float* eta,
                          const float* phi, const float* roz, const float* radius, const float* energy,
                          const int32_t* layer, const int32_t* isSi, float* rho, float* df, int32_t* dl,
                          int32_t* nhl, int32_t* nhi, int32_t* isSeed, int32_t* ti, void* stream) {
  if (!c || !p || n < 0) return set_err(c, CLUE_ERR_ARG, "bad argument");
  if (n == 0) return CLUE_OK;
  if (!z || !eta || !phi || !roz || !radius || !energy || !layer || !isSi || !rho || !df || !dl || !nhl || !nhi ||
      !isSeed || !ti)
    return set_err(c, CLUE_ERR_ARG, "null buffer");
  hipStream_t s = pick(c, stream);
  int rc = claim_scratch(c, s);
  if (rc) return rc;
  const int nb = (n + kSortBlock - 1) / kSortBlock;
  rc = ensure(c, c->scratch3d, clue3d_workspace_bytes(n, 1, nb));
  if (rc) return rc;
  Clue3DBuffers b;
  clue3d_carve(c->scratch3d.base, n, 1, nb, b);
  b.err = c->dErr;
  rc = single_setup(c, n, nb, b.blocks, b.evOff, s);
  if (rc) return rc;
  Clue3DIO io{z, eta, phi, roz, radius, energy, layer, isSi, rho, df, dl, nhl, nhi, isSeed, ti};
  clue3d_enqueue(b, io, to3d(p), s);
  HIP_TRY(c, hipGetLastError());
  return CLUE_OK;
}

int32_t clue3d_run_batch_device(clue_ctx* c, const clue3d_params* p, int32_t E, const int64_t* offsets,
                                const float* z, const float* eta, const float* phi, const float* roz,
                                const float* radius, const float* energy, const int32_t* layer, const int32_t* isSi,
                                float* rho, float* df, int32_t* dl, int32_t* nhl, int32_t* nhi, int32_t* isSeed,
                                int32_t* ti, void* stream) {
  if (!c || !p || !check_offsets(offsets, E)) return set_err(c, CLUE_ERR_ARG, "bad argument or offsets");
  const int64_t N = offsets[E];
  if (N == 0) return CLUE_OK;
  hipStream_t s = pick(c, stream);
  int rc = claim_scratch(c, s);
  if (rc) return rc;
  const int nb = nblocks_of(offsets, E);
  rc = ensure(c, c->scratch3d, clue3d_workspace_bytes(N, E, nb));
  if (rc) return rc;
  Clue3DBuffers b;
  clue3d_carve(c->scratch3d.base, N, E, nb, b);
  b.err = c->dErr;
  rc = upload_tables_async(c, offsets, E, b.blocks, b.evOff, s);
  if (rc) return rc;
  Clue3DIO io{z, eta, phi, roz, radius, energy, layer, isSi, rho, df, dl, nhl, nhi, isSeed, ti};
  clue3d_enqueue(b, io, to3d(p), s);
  HIP_TRY(c, hipGetLastError());
  return CLUE_OK;
}

int32_t clue3d_run_host(clue_ctx* c, const clue3d_params* p, int32_t n, const float* z, const float* eta,
                        const float* phi, const float* roz, const float* radius, const float* energy,
                        const int32_t* layer, const int32_t* isSi, float* rho, float* df, int32_t* dl, int32_t* nhl,
                        int32_t* nhi, int32_t* isSeed, int32_t* ti) {
  if (!c || !p || n < 0) return set_err(c, CLUE_ERR_ARG, "bad argument");
  if (n == 0) return CLUE_OK;
  const size_t nf = (size_t)n * 4;
  int rc = claim_scratch(c, c->stream);
  if (rc) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  rc = ensure(c, c->hostio, 15 * ((nf + 255) & ~(size_t)255));
  if (rc) return rc;
  Carver cv{reinterpret_cast<char*>(c->hostio.base)};
  const void* in[8] = {z, eta, phi, roz, radius, energy, layer, isSi};
  void* din[8];
  for (int k = 0; k < 8; ++k) din[k] = cv.take<float>(n);
  void* hout[7] = {rho, df, dl, nhl, nhi, isSeed, ti};
  void* dout[7];
  for (int k = 0; k < 7; ++k) dout[k] = cv.take<float>(n);
  hipStream_t s = c->stream;
  for (int k = 0; k < 8; ++k) HIP_TRY(c, hipMemcpyAsync(din[k], in[k], nf, hipMemcpyHostToDevice, s));
  rc = clue3d_run_device(c, p, n, (const float*)din[0], (const float*)din[1], (const float*)din[2],
                         (const float*)din[3], (const float*)din[4], (const float*)din[5], (const int32_t*)din[6],
                         (const int32_t*)din[7], (float*)dout[0], (float*)dout[1], (int32_t*)dout[2],
                         (int32_t*)dout[3], (int32_t*)dout[4], (int32_t*)dout[5], (int32_t*)dout[6], s);
  if (rc) return rc;
  for (int k = 0; k < 7; ++k) HIP_TRY(c, hipMemcpyAsync(hout[k], dout[k], nf, hipMemcpyDeviceToHost, s));
  return clue_check(c, s);
}

clue_plan* clue3d_plan_create(clue_ctx* c, const clue3d_params* p, int32_t E, const int64_t* offsets, const float* z,
                              const float* eta, const float* phi, const float* roz, const float* radius,
                              const float* energy, const int32_t* layer, const int32_t* isSi, float* rho, float* df,
                              int32_t* dl, int32_t* nhl, int32_t* nhi, int32_t* isSeed, int32_t* ti,
                              int32_t use_graph) {
  if (!c || !p || !check_offsets(offsets, E)) {
    set_err(c, CLUE_ERR_ARG, "bad argument or offsets");
    return nullptr;
  }
  clue_plan* pl = new (std::nothrow) clue_plan();
  if (!pl) return nullptr;
  pl->ctx = c;
  pl->dim = 3;
  const int64_t N = offsets[E];
  const int nb = nblocks_of(offsets, E);
  if (ensure(c, pl->ws, clue3d_workspace_bytes(N, E, nb)) != CLUE_OK) {
    delete pl;
    return nullptr;
  }
  clue3d_carve(pl->ws.base, N, E, nb, pl->b3);
  pl->b3.err = c->dErr;
  std::vector<BlockDesc> blocks;
  if (upload_tables(c, offsets, E, pl->b3.blocks, pl->b3.evOff, blocks) != CLUE_OK) {
    clue_plan_destroy(pl);
    return nullptr;
  }
  pl->io3 = Clue3DIO{z, eta, phi, roz, radius, energy, layer, isSi, rho, df, dl, nhl, nhi, isSeed, ti};
  pl->p3 = to3d(p);
  pl->useGraph = (use_graph & CLUE_PLAN_GRAPH) != 0;
  pl->tune = (use_graph & CLUE_PLAN_AUTOTUNE) != 0;
  return pl;
}

// ---------------------------------------------------------------- plans

int32_t clue_plan_launch(clue_plan* pl, void* stream) {
  if (!pl) return CLUE_ERR_ARG;
  clue_ctx* c = pl->ctx;
  hipStream_t s = pick(c, stream);
  if (pl->dim == 2 && pl->route2d < 0) return plan_route_first(pl, s);
  if (pl->tune && (pl->dim == 3 || pl->route2d == CLUE_PATH_GLOBAL)) return plan_autotune(pl, s);
  if (!pl->useGraph) return launch_plan_body(pl, s);
  if (!pl->exec) {
    // Record the launch sequence once on a private stream, then replay it.
    hipStream_t cap = nullptr;
    HIP_TRY(c, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    HIP_TRY(c, hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    int rc = launch_plan_body(pl, cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    pl->graph = g;
    HIP_TRY(c, hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  }
  HIP_TRY(c, hipGraphLaunch(pl->exec, s));
  return CLUE_OK;
}

void clue_plan_destroy(clue_plan* pl) {
  if (!pl) return;
  clue_ctx* c = pl->ctx;
  (void)hipDeviceSynchronize();
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  if (pl->ws.base) dev_free(c, pl->ws.base);
  if (pl->wsF.base) dev_free(c, pl->wsF.base);
  delete pl;
}

int32_t clue_plan_path(const clue_plan* pl) {
  if (!pl) return CLUE_ERR_ARG;
  return pl->dim == 2 ? pl->route2d : CLUE_PATH_GLOBAL;
}

int32_t clue_plan_index_pass(const clue_plan* pl) {
  if (!pl) return CLUE_ERR_ARG;
  return pl->dim == 2 && pl->indexPass ? 1 : 0;
}

int32_t clue_plan_tuning(const clue_plan* pl, int32_t* out4) {
  if (!pl || !out4) return CLUE_ERR_ARG;
  int dp = 0, ds = 0, np = 0, ns = 0;  // points per block and staged points of each kernel
  if (pl->dim == 3) {
    clue3d_variant_shape(pl->b3.densityVariant, dp, ds);
    clue3d_variant_shape(pl->b3.nhVariant, np, ns);
    dp *= kThreads;  // the CLUE3D kernels run 256-thread blocks of that many points per thread
    np *= kThreads;
  } else {
    clue2d_variant_shape(pl->b2.densityVariant, dp, ds);
    clue2d_variant_shape(pl->b2.nhVariant, np, ns);
  }
  out4[0] = dp;
  out4[1] = ds;
  out4[2] = np;
  out4[3] = ns;
  return CLUE_OK;
}

uint64_t clue_plan_workspace_bytes(const clue_plan* pl) { return pl ? pl->ws.bytes + pl->wsF.bytes : 0; }

float clue_plan_kernel_ms(clue_plan* pl, int32_t which, int32_t reps, void* stream) {
  if (!pl || which < 0 || which > 1 || reps < 1) return -1.f;
  clue_ctx* c = pl->ctx;
  const bool d3 = pl->dim == 3;
  if (!d3 && pl->route2d != CLUE_PATH_GLOBAL) return -1.f;  // not launched yet (or not the global-sort path)
  hipStream_t s = pick(c, stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.f;
  float ms = -1.f;
  if (hipEventRecord(e0, s) == hipSuccess) {
    for (int r = 0; r < reps; ++r) {
      if (d3)
        clue3d_enqueue_step(pl->b3, pl->io3, pl->p3, which, which ? pl->b3.nhVariant : pl->b3.densityVariant, s);
      else
        clue2d_enqueue_step(pl->b2, pl->io2, pl->p2, which, which ? pl->b2.nhVariant : pl->b2.densityVariant, s);
    }
    if (hipGetLastError() == hipSuccess && hipEventRecord(e1, s) == hipSuccess &&
        hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
      ms /= (float)reps;
    else
      ms = -1.f;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

// ---------------------------------------------------------------- self-check

int32_t clue_selftest_arith(clue_ctx* c, int32_t n, const float* v, const float* w, int32_t* xb, int32_t* yb,
                            int32_t* eb, int32_t* pb, float* red, float* dist, float* quot) {
  if (!c || n < 0 || !v || !w) return set_err(c, CLUE_ERR_ARG, "bad argument");
  if (n == 0) return CLUE_OK;
  const size_t nf = (size_t)n * 4;
  Workspace tmp;
  int rc = ensure(c, tmp, 9 * ((nf + 255) & ~(size_t)255));
  if (rc) return rc;
  Carver cv{reinterpret_cast<char*>(tmp.base)};
  float* dv = cv.take<float>(n);
  float* dw = cv.take<float>(n);
  int* o[4];
  for (int k = 0; k < 4; ++k) o[k] = cv.take<int>(n);
  float* f[3];
  for (int k = 0; k < 3; ++k) f[k] = cv.take<float>(n);
  hipStream_t s = c->stream;
  HIP_TRY(c, hipMemcpyAsync(dv, v, nf, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(dw, w, nf, hipMemcpyHostToDevice, s));
  k_selftest<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(n, dv, dw, o[0], o[1], o[2], o[3], f[0], f[1], f[2]);
  HIP_TRY(c, hipGetLastError());
  int32_t* ho[4] = {xb, yb, eb, pb};
  float* hf[3] = {red, dist, quot};
  for (int k = 0; k < 4; ++k)
    if (ho[k]) HIP_TRY(c, hipMemcpyAsync(ho[k], o[k], nf, hipMemcpyDeviceToHost, s));
  for (int k = 0; k < 3; ++k)
    if (hf[k]) HIP_TRY(c, hipMemcpyAsync(hf[k], f[k], nf, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return dev_free(c, tmp.base);
}

}  // extern "C"
