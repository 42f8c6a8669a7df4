// CLUE2D fused segment path: launch sequence and workspace (kernels in fused2d_kernels.h).
#include "fused2d_kernels.h"

namespace clue {

using namespace fused;

namespace {

template <int CAP, int NT>
void launch_class(const Fused2DBuffers& b, const Clue2DIO& io, const Clue2DParams& p, int cls, int perCU,
                  hipStream_t s) {
  const int segCap = b.E * k2dLayers;
  int grid = perCU * b.cus;
  grid = grid < segCap ? grid : segCap;
  if (grid <= 0) return;
  k_fused2d<CAP, NT><<<grid, NT, 0, s>>>(b.segs + (int64_t)cls * segCap, b.ctl + cls, b.ctl + 4 + cls, b.lg, p.dc,
                                         p.dc2, p.rhoc, p.outlierDeltaFactor, p.dm2, io.rho, io.delta,
                                         io.nearestHigher, io.isSeed, b.root);
}

}  // namespace

std::vector<int> make_event_blocks(const int64_t* offsets, int E) {
  std::vector<int> v((size_t)E + 1);
  int64_t nb = 0;
  for (int e = 0; e < E; ++e) {
    v[e] = (int)nb;
    nb += (offsets[e + 1] - offsets[e] + kSortBlock - 1) / kSortBlock;
  }
  v[E] = (int)nb;
  return v;
}

size_t fused2d_workspace_bytes(int64_t N, int E, int nBlocks) {
  const size_t n = (size_t)(N > 0 ? N : 1);
  const size_t e = (size_t)(E > 0 ? E : 1);
  const size_t nb = (size_t)(nBlocks > 0 ? nBlocks : 1);
  size_t bytes = 0;
  auto add = [&](size_t sz) { bytes += (sz + 255) & ~(size_t)255; };
  add(n * 16);                                              // lg
  add(n * 4);                                               // root
  add(n * 4);                                               // rank
  add(kRadix * nb * 4);                                     // hist
  const int64_t maxScan = (int64_t)(kRadix * nb) > (int64_t)n ? (int64_t)(kRadix * nb) : (int64_t)n;
  add((size_t)scan_partials(maxScan) * 4);                  // partials
  add((size_t)kFusedClasses * e * k2dLayers * sizeof(int4));  // segment lists
  add(64 * 4);                                              // ctl
  add((e + 1) * 4);                                         // evOff
  add((e + 1) * 4);                                         // evBlk
  add(nb * sizeof(BlockDesc));                              // blocks
  return bytes;
}

void fused2d_carve(void* base, int64_t N, int E, int nBlocks, Fused2DBuffers& b) {
  Carver cv{reinterpret_cast<char*>(base)};
  const size_t n = (size_t)(N > 0 ? N : 1);
  const size_t e = (size_t)(E > 0 ? E : 1);
  const size_t nb = (size_t)(nBlocks > 0 ? nBlocks : 1);
  b.E = E;
  b.N = N;
  b.nBlocks = nBlocks;
  b.lg = cv.take<float4>(n);
  b.root = cv.take<int>(n);
  b.rank = cv.take<int>(n);
  b.hist = cv.take<int>(kRadix * nb);
  const int64_t maxScan = (int64_t)(kRadix * nb) > (int64_t)n ? (int64_t)(kRadix * nb) : (int64_t)n;
  b.part = cv.take<int>((size_t)scan_partials(maxScan));
  b.segs = cv.take<int4>((size_t)kFusedClasses * e * k2dLayers);
  b.ctl = cv.take<int>(64);
  b.evOff = cv.take<int>(e + 1);
  b.evBlk = cv.take<int>(e + 1);
  b.blocks = cv.take<BlockDesc>(nb);
}

void fused2d_enqueue_segments(const Fused2DBuffers& b, const Clue2DIO& io, bool strict, hipStream_t s) {
  fill_i32(b.ctl, 0, 16, s);
  if (b.N <= 0 || b.nBlocks <= 0) return;
  const int nHist = kRadix * b.nBlocks;
  LayerKey key{io.layer, b.err};
  k_radix_hist<LayerKey, DigitLayer><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, key, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<LayerKey, DigitLayer, EmitPoint2D>
      <<<b.nBlocks, kThreads, 0, s>>>(b.blocks, key, b.hist, EmitPoint2D{io.x, io.y, io.weight, b.lg});
  k_segments2d<<<b.E, kRadix, 0, s>>>(b.evOff, b.evBlk, b.hist, b.segs, b.E * k2dLayers, b.ctl, strict ? 1 : 0,
                                           b.err);
}

void fused2d_enqueue_cluster(const Fused2DBuffers& b, const Clue2DIO& io, const Clue2DParams& p, hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  fill_i32(b.root, -1, b.N, s);
  launch_class<kFusedCap, 1024>(b, io, p, 2, 1, s);
  launch_class<4096, 512>(b, io, p, 1, 2, s);
  launch_class<1024, 256>(b, io, p, 0, 4, s);
  scan_exclusive(io.isSeed, b.rank, b.N, b.part, s);
  k_cluster2d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, b.evOff, b.root, b.rank, io.clusterIndex);
}

}  // namespace clue
