/*
 * clue_hip.h — C ABI of the MI355X (gfx950) CLUE / CLUE3D clusterizer.
 *
 * Plain C: pointers, sizes and opaque handles only. HIP streams travel as `void*` (a hipStream_t;
 * NULL selects the context's own stream). Entry points return CLUE_OK (0) or a negative code and
 * never throw; clue_last_error() gives the message. A C++ caller (the EDProducer plugins in
 * heterogeneous-clue_amd/plugins) turns a nonzero code into std::runtime_error, the reference's
 * error convention (src/cuda/CUDACore/cudaCheck.h:20-36, bin/main.cc:199-210).
 *
 * Threading: one clue_ctx per EDM stream, as the reference builds one module instance per stream
 * (src/serial/bin/StreamSchedule.cc:24-38). A context is thread-compatible, not thread-safe.
 *
 * Outputs are those of the reference's serial backend, bit for bit:
 *   CLUE2D: PointsCloudSerial {rho, delta, nearestHigher, isSeed, clusterIndex}
 *           (src/serial/DataFormats/PointsCloud.h:15-43)
 *   CLUE3D: ClusterCollectionSerial {rho, delta.first, delta.second, nearestHigher.first,
 *           nearestHigher.second, isSeed, tracksterIndex} (src/serial/DataFormats/ClusterCollection.h:21-51)
 * nearestHigher / nearestHigher.second / clusterIndex / tracksterIndex are indices within the event.
 */
#ifndef CLUE_HIP_H
#define CLUE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLUE_OK 0
#define CLUE_ERR_ARG (-1)      /* bad argument (null pointer, negative size, bad offsets) */
#define CLUE_ERR_HIP (-2)      /* HIP runtime error */
#define CLUE_ERR_BUDGET (-3)   /* the allocation would exceed the context's HBM budget */
#define CLUE_ERR_LAYER (-4)    /* an input layer outside the tile array (2D: [0,100), 3D: [0,94)) */
#define CLUE_ERR_BIN (-5)      /* CLUE3D: a phi so far outside [-pi, pi] that its bin leaves the tile */
#define CLUE_ERR_TIMEOUT (-6)  /* clue_server_wait: the request did not finish in time */
#define CLUE_ERR_CAPACITY (-7) /* a fixed capacity exceeded: CLUE2D fused path layer, no-alloc scratch */

/* CLUE2D execution paths (clue_set_path). Both give the serial backend's results bit for bit.
 *   CLUE_PATH_GLOBAL the batch is sorted into bins in device memory (no size limit);
 *   CLUE_PATH_FUSED  the batch is grouped by (event, layer) and each layer is clustered by one
 *                    workgroup in LDS (layers of up to CLUE_FUSED_CAPACITY points). Disabled in this
 *                    build: clue_set_path refuses it with CLUE_ERR_ARG;
 *   CLUE_PATH_AUTO   (default) the global-sort path. */
#define CLUE_PATH_AUTO 0
#define CLUE_PATH_FUSED 1
#define CLUE_PATH_GLOBAL 2
#define CLUE_FUSED_CAPACITY 10112

typedef struct clue_ctx clue_ctx;
typedef struct clue_plan clue_plan;
typedef struct clue_server clue_server;

/* CLUE2D parameters: Parameters {dc, rhoc, outlierDeltaFactor} (src/serial/DataFormats/CLUE_config.h:8-13),
 * produced by CLUESerialClusterizerESProducer from the CSV `dc,rhoc,odf,produceOutput`
 * (src/serial/plugin-CLUEClusterizer/CLUESerialClusterizerESProducer.cc:17-34). */
typedef struct {
  float dc;
  float rhoc;
  float outlierDeltaFactor;
} clue2d_params;

/* CLUE3D parameters: the kernel arguments of KernelCalculateDensitySoA, KernelComputeDistanceToHigherSoA
 * and KernelFindAndAssignClustersSoA (src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:26-32,
 * 286-290, 509-515). clue3d_default_params() returns the reference's compiled-in defaults. */
typedef struct {
  int32_t densitySiblingLayers;     /* 3 */
  float densityXYDistanceSqr;       /* 3.24 cm^2 */
  float kernelDensityFactor;        /* 0.2 */
  int32_t densityOnSameLayer;       /* 0 */
  int32_t nearestHigherOnSameLayer; /* 0 */
  float criticalXYDistance;         /* 1.8 cm */
  float criticalZDistanceLyr;       /* 5 */
  float criticalDensity;            /* 0.6 GeV */
  float criticalSelfDensity;        /* 0.15 */
  float outlierMultiplier;          /* 2 */
} clue3d_params;

clue3d_params clue3d_default_params(void);

/* ---------------------------------------------------------------- context, streams, memory */

/* Replaces the per-stream device binding of ScopedContextProduce (src/cuda/CUDACore/ScopedContext.h:106-160,
 * chooseDevice.cc:5-13) and the caching allocator's budget (getCachingDeviceAllocator.h:14-94).
 * hbm_budget_bytes = 0: no limit. */
clue_ctx* clue_create(int32_t device, uint64_t hbm_budget_bytes);
void clue_destroy(clue_ctx* ctx);
const char* clue_last_error(const clue_ctx* ctx);
int32_t clue_device_count(void);
uint64_t clue_bytes_in_use(const clue_ctx* ctx);

void* clue_stream_create(clue_ctx* ctx);
int32_t clue_stream_destroy(clue_ctx* ctx, void* stream);
int32_t clue_stream_sync(clue_ctx* ctx, void* stream);
/* Calls fn(user) once, on a runtime thread, after all work enqueued on `stream` before this call
 * has finished, without blocking the caller (hipLaunchHostFunc). Replaces the CUDA backend's
 * completion callback that hands an external-work module's holder back to the framework
 * (src/cuda/CUDACore/ScopedContext.cc:94-98, cudaStreamAddCallback). fn must not call this library
 * or HIP; it runs whether or not the enqueued kernels flagged an error (read those with clue_check). */
int32_t clue_stream_callback(clue_ctx* ctx, void* stream, void (*fn)(void* user), void* user);

/* Device / pinned-host memory owned by the caller, counted against the budget (device only). */
void* clue_malloc_device(clue_ctx* ctx, uint64_t bytes);
int32_t clue_free_device(clue_ctx* ctx, void* ptr);
void* clue_malloc_pinned(clue_ctx* ctx, uint64_t bytes);
int32_t clue_free_pinned(clue_ctx* ctx, void* ptr);
/* Stream-ordered caching allocator over the context's device memory (replaces the reference's
 * CachingDeviceAllocator, src/cuda/CUDACore/getCachingDeviceAllocator.h:14-94, used through
 * cms::cuda::make_device_unique). Sizes round up to power-of-two bins of 2^8 .. 2^34 bytes (larger
 * requests get exact blocks that are not cached). clue_cache_free gives a block back to the cache
 * with an event recorded on the stream it was allocated on; a later clue_cache_alloc of the same bin
 * reuses it at once on that stream, on another stream once the event has completed. Blocks count
 * against the HBM budget; idle blocks are released when a new block would not fit, when idle bytes
 * pass the limit (the budget, else 80 % of the free device memory at first use), by
 * clue_cache_release, and with the context. */
void* clue_cache_alloc(clue_ctx* ctx, uint64_t bytes, void* stream);
int32_t clue_cache_free(clue_ctx* ctx, void* ptr);
int32_t clue_cache_release(clue_ctx* ctx);
int32_t clue_cache_stats(clue_ctx* ctx, uint64_t* live_bytes, uint64_t* idle_bytes, int64_t* hits, int64_t* misses);
/* Asynchronous on `stream` (stream-ordered); pinned host memory makes them truly asynchronous. */
int32_t clue_memcpy_htod(clue_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);
int32_t clue_memcpy_dtoh(clue_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);
int32_t clue_memcpy_dtod(clue_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);

/* Synchronise `stream` and report errors the kernels flagged since the last check
 * (CLUE_ERR_LAYER, CLUE_ERR_BIN, CLUE_ERR_CAPACITY). */
int32_t clue_check(clue_ctx* ctx, void* stream);

/* Elapsed milliseconds between two points of `stream`: record with clue_mark, read with clue_elapsed_ms
 * after synchronising. Marks are HIP events owned by the context (index 0..15). */
int32_t clue_mark(clue_ctx* ctx, int32_t index, void* stream);
float clue_elapsed_ms(clue_ctx* ctx, int32_t from_index, int32_t to_index);

/* ---------------------------------------------------------------- CLUE2D */

/* Select the CLUE2D path (CLUE_PATH_*) for the context's later calls and plans. dim must be 2.
 * No reference counterpart: the reference has one implementation per backend. With the fused path
 * enabled, a call whose batch holds an event of more than CLUE_FUSED_CAPACITY points waits once for
 * the device (after grouping by layer) to learn whether every layer fits, and a plan decides at its
 * first launch and reports CLUE_ERR_CAPACITY (through clue_check) if a later launch's data no longer
 * fits. */
int32_t clue_set_path(clue_ctx* ctx, int32_t dim, int32_t path);

/* One event, device-resident SoA, asynchronous on `stream`.
 * Replaces CLUEAlgoCUDA::makeClusters (src/cuda/plugin-CLUEClusterizer/CLUEAlgoCUDA.cu:59-81) with the
 * serial backend's results (CLUEAlgoSerial::makeClusters, src/serial/plugin-CLUEClusterizer/CLUEAlgoSerial.cc:18-35). */
int32_t clue2d_run_device(clue_ctx* ctx, const clue2d_params* p, int32_t n, const float* x, const float* y,
                          const int32_t* layer, const float* weight, float* rho, float* delta, int32_t* nearestHigher,
                          int32_t* isSeed, int32_t* clusterIndex, void* stream);

/* clue2d_run_device without any allocation: no reference counterpart. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and stream semantics as clue2d_run_device; the workspace is a
 * fixed 64 MiB device array of the library (one per GPU, held until the process ends) instead of
 * the context's scratch, so events are limited to CLUE_NOALLOC_MAX_POINTS points (CLUE_ERR_CAPACITY
 * above). Every call on a GPU shares that GPU's array: calls on one stream run one after another,
 * but calls that overlap on one GPU (other streams or contexts) race. */
#define CLUE_NOALLOC_MAX_POINTS 1048576
int32_t clue2d_run_device_noalloc(clue_ctx* ctx, const clue2d_params* p, int32_t n, const float* x, const float* y,
                                  const int32_t* layer, const float* weight, float* rho, float* delta,
                                  int32_t* nearestHigher, int32_t* isSeed, int32_t* clusterIndex, void* stream);

/* One event, host SoA in and out, synchronous: the body of an EDProducer::produce
 * (src/serial/plugin-CLUEClusterizer/CLUESerialClusterizer.cc:29-36). */
int32_t clue2d_run_host(clue_ctx* ctx, const clue2d_params* p, int32_t n, const float* x, const float* y,
                        const int32_t* layer, const float* weight, float* rho, float* delta, int32_t* nearestHigher,
                        int32_t* isSeed, int32_t* clusterIndex);

/* A packed batch of events (event e = points offsets[e] .. offsets[e+1]-1; host offsets, device SoA),
 * asynchronous on `stream`, as the reference's CUDA backend enqueues its copies and launches
 * (src/cuda/plugin-CLUEClusterizer/CLUEAlgoCUDA.cu:31-38,53-79): the batch's block table is built on
 * the host into one of 4 pinned buffers of the context and copied with hipMemcpyAsync on `stream`.
 * The call waits for the device only when the context's workspace must grow (a larger batch than
 * any before) or when the pinned buffer it takes still awaits its copy (4 earlier batch calls not
 * yet started on the device). `offsets` may be reused as soon as the call returns. No reference
 * counterpart otherwise: the reference runs one event per produce call. */
int32_t clue2d_run_batch_device(clue_ctx* ctx, const clue2d_params* p, int32_t n_events, const int64_t* offsets,
                                const float* x, const float* y, const int32_t* layer, const float* weight, float* rho,
                                float* delta, int32_t* nearestHigher, int32_t* isSeed, int32_t* clusterIndex,
                                void* stream);

/* A reusable plan for a fixed batch (sizes and buffers): its own workspace, launched with
 * clue_plan_launch. use_graph is a set of flags: CLUE_PLAN_GRAPH records the launch sequence once into
 * a hipGraph and replays it; CLUE_PLAN_AUTOTUNE makes the first launch time every variant
 * (points per workgroup, LDS staging capacity) of the density and nearest-higher kernels on the
 * plan's own data and keep the fastest of each for the later launches (clue_plan_tuning reports the
 * choice). All variants give the same bits. No reference counterpart. */
#define CLUE_PLAN_GRAPH 1
#define CLUE_PLAN_AUTOTUNE 2
/* CLUE2D plans: CLUE_PLAN_INDEX_PASS selects the second output form of the nearest-higher step. Its
 * kernel writes one 16-byte record per point at the point's index instead of delta, nearestHigher and
 * isSeed separately, keeps the follower links in tile order, and the labels step writes the three
 * outputs in index order while it walks the links. That pays when hit indices are unrelated to hit
 * positions (every scattered 4-byte write costs a memory burst); hits that arrive grouped by
 * position are served better by the direct writes. With CLUE_PLAN_AUTOTUNE the first launch times
 * both forms and keeps the faster (clue_plan_index_pass reports it). Both forms give the same bits. */
#define CLUE_PLAN_INDEX_PASS 4
clue_plan* clue2d_plan_create(clue_ctx* ctx, const clue2d_params* p, int32_t n_events, const int64_t* offsets,
                              const float* x, const float* y, const int32_t* layer, const float* weight, float* rho,
                              float* delta, int32_t* nearestHigher, int32_t* isSeed, int32_t* clusterIndex,
                              int32_t use_graph);

/* ---------------------------------------------------------------- CLUE3D */

/* One event, device-resident SoA, asynchronous. Replaces CLUE3DAlgoSerial::makeTrackstersSoA
 * (src/serial/plugin-CLUETracksterizer/CLUE3DAlgoSerial.cc:42-50) with explicit kernel parameters.
 * ClusterCollection's x and y (ClusterCollection.h:9-10) are not read by the algorithm and not passed. */
int32_t clue3d_run_device(clue_ctx* ctx, const clue3d_params* p, int32_t n, const float* z, const float* eta,
                          const float* phi, const float* r_over_absz, const float* radius, const float* energy,
                          const int32_t* layer, const int32_t* isSilicon, float* rho, float* delta_first,
                          int32_t* delta_second, int32_t* nh_layer, int32_t* nh_index, int32_t* isSeed,
                          int32_t* tracksterIndex, void* stream);

/* Host SoA in/out, synchronous: the body of CLUESerialTracksterizer::produce (CLUESerialTracksterizer.cc:29-41). */
int32_t clue3d_run_host(clue_ctx* ctx, const clue3d_params* p, int32_t n, const float* z, const float* eta,
                        const float* phi, const float* r_over_absz, const float* radius, const float* energy,
                        const int32_t* layer, const int32_t* isSilicon, float* rho, float* delta_first,
                        int32_t* delta_second, int32_t* nh_layer, int32_t* nh_index, int32_t* isSeed,
                        int32_t* tracksterIndex);

/* A packed CLUE3D batch, asynchronous in the same way as clue2d_run_batch_device. */
int32_t clue3d_run_batch_device(clue_ctx* ctx, const clue3d_params* p, int32_t n_events, const int64_t* offsets,
                                const float* z, const float* eta, const float* phi, const float* r_over_absz,
                                const float* radius, const float* energy, const int32_t* layer,
                                const int32_t* isSilicon, float* rho, float* delta_first, int32_t* delta_second,
                                int32_t* nh_layer, int32_t* nh_index, int32_t* isSeed, int32_t* tracksterIndex,
                                void* stream);

clue_plan* clue3d_plan_create(clue_ctx* ctx, const clue3d_params* p, int32_t n_events, const int64_t* offsets,
                              const float* z, const float* eta, const float* phi, const float* r_over_absz,
                              const float* radius, const float* energy, const int32_t* layer,
                              const int32_t* isSilicon, float* rho, float* delta_first, int32_t* delta_second,
                              int32_t* nh_layer, int32_t* nh_index, int32_t* isSeed, int32_t* tracksterIndex,
                              int32_t use_graph);

/* ---------------------------------------------------------------- plans */

int32_t clue_plan_launch(clue_plan* plan, void* stream);
/* CLUE2D plan: the path it runs (CLUE_PATH_FUSED or CLUE_PATH_GLOBAL), or -1 before the first launch
 * has decided it; CLUE3D plans: CLUE_PATH_GLOBAL. */
int32_t clue_plan_path(const clue_plan* plan);
/* 1 when a CLUE2D plan uses the index-pass output form (CLUE_PLAN_INDEX_PASS, or the autotuner's
 * choice after an autotuned plan's first launch), else 0. No reference counterpart. */
int32_t clue_plan_index_pass(const clue_plan* plan);
/* Points per workgroup and LDS staging capacity (points) of the density and the
 * nearest-higher kernels it launches, as out[0..3] = {density points, density stage, nh points,
 * nh stage}. Before an autotuned plan's first launch these are the defaults. */
int32_t clue_plan_tuning(const clue_plan* plan, int32_t* out4);
void clue_plan_destroy(clue_plan* plan);
uint64_t clue_plan_workspace_bytes(const clue_plan* plan);
/* Measurement helper, no reference counterpart: mean milliseconds of one launch of the density
 * (which = 0) or nearest-higher (which = 1) kernel of a plan that has been launched, with the
 * variant it uses, re-run `reps` times on `stream` between two HIP events (both kernels only rewrite
 * their outputs with the same values). Negative on error. */
float clue_plan_kernel_ms(clue_plan* plan, int32_t which, int32_t reps, void* stream);

/* ---------------------------------------------------------------- serving loop
 * No reference counterpart: the reference processes one event per produce call. A server owns
 * `slots` batches in flight, each with its own HIP stream, pinned staging and device workspace for
 * up to `max_events` events and `max_points_2d` / `max_points_3d` points, allocated once against
 * the context's HBM budget and kept resident. Submitted events queue in arrival order; whenever a
 * slot is free, a worker thread takes as many queued events as fit into it, runs them as one batch
 * and copies each result into the caller's output buffers (continuous batching: events arriving
 * while batches run are admitted as slots free up, and batches on different slots overlap on the
 * device). Input and output buffers are the caller's and must stay valid until clue_server_wait
 * returns for the request. While a server exists, its context must not be used from other threads.
 * A result is bit-identical to the same event run alone. */
clue_server* clue_server_create(clue_ctx* ctx, const clue2d_params* p2, const clue3d_params* p3,
                                int64_t max_points_2d, int64_t max_points_3d, int32_t max_events, int32_t slots);
void clue_server_destroy(clue_server* server);
/* Returns a request id >= 0, or a negative error code (e.g. CLUE_ERR_LAYER, CLUE_ERR_BUDGET when the
 * event exceeds a slot's capacity). */
int64_t clue_server_submit2d(clue_server* server, int32_t n, const float* x, const float* y, const int32_t* layer,
                             const float* weight, float* rho, float* delta, int32_t* nearestHigher, int32_t* isSeed,
                             int32_t* clusterIndex);
int64_t clue_server_submit3d(clue_server* server, int32_t n, const float* z, const float* eta, const float* phi,
                             const float* r_over_absz, const float* radius, const float* energy, const int32_t* layer,
                             const int32_t* isSilicon, float* rho, float* delta_first, int32_t* delta_second,
                             int32_t* nh_layer, int32_t* nh_index, int32_t* isSeed, int32_t* tracksterIndex);
/* Blocks until request `id` is done (outputs written) or timeout_ms elapses (CLUE_ERR_TIMEOUT);
 * on success stores the request's latency (submission to results in the caller's buffers). */
int32_t clue_server_wait(clue_server* server, int64_t id, int32_t timeout_ms, double* latency_ms);
int32_t clue_server_stats(clue_server* server, int64_t* batches, int64_t* events, int64_t* points);
const char* clue_server_last_error(clue_server* server);

/* ---------------------------------------------------------------- arithmetic self-check
 * Evaluates the device binning and distance helpers on n host values (synchronous), so that the
 * device arithmetic can be compared with the CPU reference value by value. Any output may be NULL.
 * xbin/ybin: LayerTilesSerial::getXBin/getYBin; eta_bin/phi_bin: TICLLayerTileT::etaBin/phiBin;
 * reduce: reco::reduceRange<float>; dist: sqrtf(a*a + b*b) of CLUEAlgoKernels.h:10-12 with a = v[i],
 * b = w[i]; quot: a / b. */
int32_t clue_selftest_arith(clue_ctx* ctx, int32_t n, const float* v, const float* w, int32_t* xbin, int32_t* ybin,
                            int32_t* eta_bin, int32_t* phi_bin, float* reduce, float* dist, float* quot);

#ifdef __cplusplus
}
#endif
#endif /* CLUE_HIP_H */
