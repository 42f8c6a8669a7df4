// CLUE3D (SoA path) on gfx950: tile build, density over sibling layers, nearest-higher on the
// other layers (+ seed/outlier flags) and trackster labels for a packed batch of events. The
// kernels are in clue3d_kernels.h.
#include "clue3d_kernels.h"

namespace clue {

using namespace g3d;

namespace {

template <int V, class... A>
void density_v(int nBlocks, hipStream_t s, A... a) {
  constexpr Variant3D v = kVariants3D[V];
  k_density3d<v.items, v.stage><<<CompShape3D<v.items>::kPerSort * nBlocks, kThreads, 0, s>>>(a...);
}
template <int V, class... A>
void nh_v(int nBlocks, hipStream_t s, A... a) {
  constexpr Variant3D v = kVariants3D[V];
  k_nh3d<v.items, v.stage><<<CompShape3D<v.items>::kPerSort * nBlocks, kThreads, 0, s>>>(a...);
}

template <class... A>
void launch_density3d(int variant, int nBlocks, hipStream_t s, A... a) {
  switch (variant) {
    case 1: density_v<1>(nBlocks, s, a...); break;
    case 2: density_v<2>(nBlocks, s, a...); break;
    case 3: density_v<3>(nBlocks, s, a...); break;
    default: density_v<0>(nBlocks, s, a...); break;
  }
}
template <class... A>
void launch_nh3d(int variant, int nBlocks, hipStream_t s, A... a) {
  switch (variant) {
    case 1: nh_v<1>(nBlocks, s, a...); break;
    case 2: nh_v<2>(nBlocks, s, a...); break;
    case 3: nh_v<3>(nBlocks, s, a...); break;
    default: nh_v<0>(nBlocks, s, a...); break;
  }
}
static_assert(kNumVariants3D == 4, "launch_density3d / launch_nh3d dispatch four variants");

void density_step(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& prm, int variant, hipStream_t s) {
  launch_density3d(variant, b.nBlocks, s, b.blocks, b.i1, b.cre, io.eta, io.radius, io.layer, io.isSilicon, b.table,
                   b.tableStride, prm, b.crp, io.rho);
}
void nh_step(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& prm, int variant, hipStream_t s) {
  int* parent = reinterpret_cast<int*>(b.k2);
  launch_nh3d(variant, b.nBlocks, s, b.blocks, b.i1, b.crp, b.cre, io.eta, io.radius, io.layer, io.isSilicon,
              b.table, b.tableStride, b.evOff, prm, io.delta_f, io.delta_l, io.nh_layer, io.nh_idx, io.isSeed,
              parent);
}

}  // namespace

void clue3d_enqueue(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& prm, hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  const int nHist = kRadix * b.nBlocks;
  Keys3D keys{io.eta, io.phi, io.layer, b.k2, b.err, b.errStride};
  k_radix_hist<Keys3D, DigitPhi><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, keys, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<KeyArray, DigitPhi, EmitFirst>
      <<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k2}, b.hist, EmitFirst{b.k1, b.i1});
  k_radix_hist<KeyArray, DigitEta><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k1}, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<KeyArray, DigitEta, EmitMid>
      <<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k1}, b.hist, EmitMid{b.i1, b.k2, b.i2});
  k_radix_hist<KeyArray, DigitLay><<<b.nBlocks, kThreads, 0, s>>>(b.blocks, KeyArray{b.k2}, b.hist);
  scan_exclusive(b.hist, b.hist, nHist, b.part, s);
  k_radix_scatter<KeyArray, DigitLay, EmitLast3D><<<b.nBlocks, kThreads, 0, s>>>(
      b.blocks, KeyArray{b.k2}, b.hist,
      EmitLast3D{b.i2, b.k1, b.i1, io.r_over_absz, io.phi, io.energy, io.z, b.cre});
  // sorted keys in k1, permutation in i1
  // a window reaches S sibling layers past a point's layer, and its eta row 34 the layer after those
  const int S = prm.densitySiblingLayers > 0 ? prm.densitySiblingLayers : 0;
  build_table(b.k1, b.blocks, b.nBlocks, b.evOff, b.E, k3dKeys, k3dBinsPerLayer, (S < k3dLayers ? S : k3dLayers) + 1,
              b.tableStride, b.table, b.tileStart, s);
  density_step(b, io, prm, b.densityVariant, s);
  nh_step(b, io, prm, b.nhVariant, s);
  int* parent = reinterpret_cast<int*>(b.k2);
  launch_labels(b.blocks, b.nBlocks, parent, io.isSeed, b.N, b.evOff, io.tracksterIndex, b.seeds, b.part, s);
}

void clue3d_enqueue_step(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& prm, int which,
                         int variant, hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  if (which == 0)
    density_step(b, io, prm, variant, s);
  else
    nh_step(b, io, prm, variant, s);
}

void clue3d_variant_shape(int variant, int& items, int& stage) {
  const Variant3D v = kVariants3D[variant >= 0 && variant < kNumVariants3D ? variant : 0];
  items = v.items;
  stage = v.stage;
}

// Carve a workspace of clue3d_workspace_bytes(N, E, nBlocks) bytes into the buffers of a batch.
void clue3d_carve(void* base, int64_t N, int E, int nBlocks, Clue3DBuffers& b) {
  Carver cv{reinterpret_cast<char*>(base)};
  const size_t n = (size_t)(N > 0 ? N : 1);
  b.E = E;
  b.N = N;
  b.nBlocks = nBlocks;
  b.k1 = cv.take<uint32_t>(n);
  b.i1 = cv.take<int>(n);
  b.k2 = cv.take<uint32_t>(n);
  b.i2 = cv.take<int>(n);
  b.cre = cv.take<float4>(n);
  b.crp = cv.take<float4>(n);
  b.hist = cv.take<int>((size_t)kRadix * (nBlocks > 0 ? nBlocks : 1));
  const int64_t maxScan = (int64_t)kRadix * nBlocks > N ? (int64_t)kRadix * nBlocks : N;
  b.part = cv.take<int>((size_t)scan_partials(maxScan > 0 ? maxScan : 1));
  b.tableStride = clue3d_table_stride();
  b.table = cv.take<int>((size_t)b.tableStride * (E > 0 ? E : 1));
  b.tileStart = cv.take<int>((size_t)(table_tiles(k3dKeys) + 1) * (E > 0 ? E : 1));
  b.evOff = cv.take<int>((size_t)E + 1);
  b.blocks = cv.take<BlockDesc>((size_t)(nBlocks > 0 ? nBlocks : 1));
  b.seeds = seed_rank_carve(cv, N);
}

size_t clue3d_workspace_bytes(int64_t N, int nEvents, int nBlocks) {
  const size_t n = (size_t)(N > 0 ? N : 1);
  size_t bytes = 0;
  auto add = [&](size_t sz) { bytes += (sz + 255) & ~(size_t)255; };
  for (int k = 0; k < 4; ++k) add(n * 4);   // k1 i1 k2 i2
  for (int k = 0; k < 2; ++k) add(n * 16);  // cre crp
  add((size_t)kRadix * (size_t)(nBlocks > 0 ? nBlocks : 1) * 4);
  const int64_t maxScan = (int64_t)kRadix * nBlocks > N ? (int64_t)kRadix * nBlocks : N;
  add((size_t)scan_partials(maxScan > 0 ? maxScan : 1) * 4);
  add((size_t)clue3d_table_stride() * (size_t)(nEvents > 0 ? nEvents : 1) * 4);
  add((size_t)(table_tiles(k3dKeys) + 1) * (size_t)(nEvents > 0 ? nEvents : 1) * 4);  // tileStart
  add((size_t)(nEvents + 1) * 4);
  add((size_t)(nBlocks > 0 ? nBlocks : 1) * sizeof(BlockDesc));
  bytes += seed_rank_bytes(N);  // seed words
  add(256);
  return bytes;
}

}  // namespace clue
