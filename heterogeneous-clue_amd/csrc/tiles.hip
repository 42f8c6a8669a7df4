// Scan, bin table, label propagation and fill kernels shared by CLUE2D and CLUE3D.
#include "tiles_kernels.h"

namespace clue {

std::vector<BlockDesc> make_blocks(const int64_t* offsets, int nEvents, int blockSize) {
  std::vector<BlockDesc> out;
  for (int e = 0; e < nEvents; ++e) {
    const int64_t a = offsets[e], b = offsets[e + 1];
    const int nb = (int)((b - a + blockSize - 1) / blockSize);
    const int first = (int)out.size();
    for (int k = 0; k < nb; ++k) {
      BlockDesc d;
      d.event = e;
      d.begin = (int32_t)(a + (int64_t)k * blockSize);
      d.end = (int32_t)((a + (int64_t)(k + 1) * blockSize) < b ? (a + (int64_t)(k + 1) * blockSize) : b);
      d.histBase = kRadix * first;
      d.nbEvent = nb;
      d.bLocal = k;
      out.push_back(d);
    }
  }
  return out;
}

void scan_exclusive(const int* in, int* out, int64_t n, int* part, hipStream_t s) {
  if (n <= 0) return;
  const int np = scan_partials(n);
  if (np <= kScanTile) {  // one or two launches: the down pass sums the partials it needs itself
    if (np > 1) k_scan_reduce<<<np, kThreads, 0, s>>>(in, n, part);
    k_scan_down_direct<<<np, kThreads, 0, s>>>(in, out, n, part);
    return;
  }
  k_scan_reduce<<<np, kThreads, 0, s>>>(in, n, part);
  k_scan_partials<<<1, kThreads, 0, s>>>(part, np);
  k_scan_down<<<np, kThreads, 0, s>>>(in, out, n, part);
}

void build_table(const uint32_t* keySorted, const BlockDesc* blocksDev, int nBlocks, const int* evOffDev,
                 int nEvents, int nKeys, int binsPerLayer, int layerReach, int64_t stride, int* table,
                 int* tileStart, hipStream_t s) {
  if (nEvents <= 0) return;
  const int tiles = table_tiles(nKeys);
  if (nBlocks > 0) k_tile_starts<<<nBlocks, kThreads, 0, s>>>(blocksDev, keySorted, evOffDev, tiles, tileStart);
  k_table<<<nEvents * tiles, kThreads, 0, s>>>(keySorted, evOffDev, tileStart, tiles, nKeys, binsPerLayer, layerReach,
                                               stride, table);
}

namespace {
// Seed words (from isSeed, or already set when isSeed is null), word prefixes and scanned tile counts.
void seed_ranks(const int* isSeed, int64_t N, const SeedRank& sr, int* part, hipStream_t s) {
  const int tiles = seed_tiles(N);
  if (isSeed)
    k_seed_bits<<<tiles, kThreads, 0, s>>>(isSeed, N, sr);
  else
    k_seed_prefix<<<(tiles * 32 + kThreads - 1) / kThreads, kThreads, 0, s>>>((int64_t)tiles * 32, sr);
  scan_exclusive(sr.tilePre, sr.tilePre, tiles, part, s);
}
}  // namespace

void launch_labels(const BlockDesc* blocksDev, int nBlocks, const int* parent, const int* isSeed, int64_t N,
                   const int* evOffDev, int32_t* label, const SeedRank& sr, int* part, hipStream_t s) {
  if (nBlocks <= 0 || N <= 0) return;
  seed_ranks(isSeed, N, sr, part, s);
  k_labels<<<nBlocks, kThreads, 0, s>>>(blocksDev, parent, sr, evOffDev, label);
}

void launch_seed_ranks(int64_t N, const SeedRank& sr, int* part, hipStream_t s) {
  if (N <= 0) return;
  seed_ranks(nullptr, N, sr, part, s);
}

void fill_i32(int* p, int value, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, s>>>(p, value, n);
}

}  // namespace clue
