// Emulated counterparts of the library's host launch helpers (tiles.hip): the same kernels, grids
// and argument order, run through emu::launch. Test infrastructure only.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tiles_kernels.h"

namespace emu_host {

using namespace clue;

// scan_exclusive's launch sequence; the direct two-launch form up to `directMax` partials.
inline void scan_exclusive(const int* in, int* out, int64_t n, int* part, int directMax = kScanTile) {
  if (n <= 0) return;
  const int np = scan_partials(n);
  if (np <= directMax) {
    if (np > 1) emu::launch(np, kThreads, [&] { k_scan_reduce(in, n, part); });
    emu::launch(np, kThreads, [&] { k_scan_down_direct(in, out, n, part); });
    return;
  }
  emu::launch(np, kThreads, [&] { k_scan_reduce(in, n, part); });
  emu::launch(1, kThreads, [&] { k_scan_partials(part, np); });
  emu::launch(np, kThreads, [&] { k_scan_down(in, out, n, part); });
}

inline void build_table(const uint32_t* key, const BlockDesc* blocks, int nBlocks, const int* evOff, int nEvents,
                        int nKeys, int binsPerLayer, int layerReach, int64_t stride, int* table, int* tileStart) {
  if (nEvents <= 0) return;
  const int tiles = table_tiles(nKeys);
  if (nBlocks > 0) emu::launch(nBlocks, kThreads, [&] { k_tile_starts(blocks, key, evOff, tiles, tileStart); });
  emu::launch(nEvents * tiles, kThreads,
              [&] { k_table(key, evOff, tileStart, tiles, nKeys, binsPerLayer, layerReach, stride, table); });
}

inline void launch_labels(const BlockDesc* blocks, int nBlocks, const int* parent, const int* isSeed, int64_t N,
                          const int* evOff, int32_t* label, const SeedRank& sr, int* part) {
  if (nBlocks <= 0 || N <= 0) return;
  const int tiles = seed_tiles(N);
  if (isSeed)
    emu::launch(tiles, kThreads, [&] { k_seed_bits(isSeed, N, sr); });
  else
    emu::launch((tiles * 32 + kThreads - 1) / kThreads, kThreads, [&] { k_seed_prefix((int64_t)tiles * 32, sr); });
  scan_exclusive(sr.tilePre, sr.tilePre, tiles, part);
  emu::launch(nBlocks, kThreads, [&] { k_labels(blocks, parent, sr, evOff, label); });
}

inline void launch_seed_ranks(int64_t N, const SeedRank& sr, int* part) {
  if (N <= 0) return;
  const int tiles = seed_tiles(N);
  emu::launch((tiles * 32 + kThreads - 1) / kThreads, kThreads, [&] { k_seed_prefix((int64_t)tiles * 32, sr); });
  scan_exclusive(sr.tilePre, sr.tilePre, tiles, part);
}

// Seed-rank buffers of N points, poisoned like fresh device memory.
struct SeedRankHost {
  std::vector<uint64_t> word;
  std::vector<uint16_t> wordPre;
  std::vector<int> tilePre;
  explicit SeedRankHost(int64_t N) {
    const size_t t = (size_t)std::max(seed_tiles(N), 1);
    word.assign(t * 32, 0x5A5A5A5A5A5A5A5Aull);
    wordPre.assign(t * 32, 0x5A5A);
    tilePre.assign(t + 1, 0x5A5A5A5A);
  }
  SeedRank view() { return SeedRank{word.data(), wordPre.data(), tilePre.data()}; }
};

// Block table of a packed batch, as make_blocks builds it.
inline std::vector<BlockDesc> blocks_of(const std::vector<int64_t>& offs) {
  std::vector<BlockDesc> out;
  const int E = (int)offs.size() - 1;
  for (int e = 0; e < E; ++e) {
    const int nb = (int)((offs[e + 1] - offs[e] + kSortBlock - 1) / kSortBlock);
    const int first = (int)out.size();
    for (int k = 0; k < nb; ++k) {
      BlockDesc d;
      d.event = e;
      d.begin = (int)(offs[e] + (int64_t)k * kSortBlock);
      d.end = (int)std::min<int64_t>(offs[e] + (int64_t)(k + 1) * kSortBlock, offs[e + 1]);
      d.histBase = kRadix * first;
      d.nbEvent = nb;
      d.bLocal = k;
      out.push_back(d);
    }
  }
  return out;
}

inline float sqrt_le(float d) {  // largest v with sqrtf(v) <= d
  // as the library's sqrt_le_threshold: NaN passes nothing, a negative d nothing either
  if (d != d) return d;
  if (d < 0.f) return -1.f;
  if (d == INFINITY) return INFINITY;
  uint32_t lo = 0, hi = 0x7f800000u;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo + 1) / 2;
    float v;
    std::memcpy(&v, &mid, 4);
    if (std::sqrt(v) <= d)
      lo = mid;
    else
      hi = mid - 1;
  }
  float t;
  std::memcpy(&t, &lo, 4);
  return t;
}

// Input file of the emulation drivers: int32 E, float dc, rhoc, odf, int64 offsets[E+1],
// float x[N], y[N], int32 layer[N], float w[N].
struct Input2D {
  int E = 0;
  float dc = 0, rhoc = 0, odf = 0;
  std::vector<int64_t> offs;
  std::vector<float> x, y, w;
  std::vector<int> layer;
};

template <class T>
inline void rd(FILE* f, T* p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}

inline Input2D read_input(const char* path) {
  Input2D in;
  FILE* f = fopen(path, "rb");
  if (!f) exit(2);
  rd(f, &in.E, 1);
  rd(f, &in.dc, 1);
  rd(f, &in.rhoc, 1);
  rd(f, &in.odf, 1);
  in.offs.resize(in.E + 1);
  rd(f, in.offs.data(), in.offs.size());
  const size_t N = (size_t)in.offs[in.E];
  in.x.resize(N);
  in.y.resize(N);
  in.layer.resize(N);
  in.w.resize(N);
  rd(f, in.x.data(), N);
  rd(f, in.y.data(), N);
  rd(f, in.layer.data(), N);
  rd(f, in.w.data(), N);
  fclose(f);
  return in;
}

// Output file: int32 err, int32 extra, float rho[N], delta[N], int32 nh[N], isSeed[N], clusterIndex[N].
inline void write_output(const char* path, int err, int extra, const std::vector<float>& rho,
                         const std::vector<float>& delta, const std::vector<int>& nh, const std::vector<int>& seed,
                         const std::vector<int>& ci) {
  FILE* o = fopen(path, "wb");
  if (!o) exit(2);
  fwrite(&err, 4, 1, o);
  fwrite(&extra, 4, 1, o);
  fwrite(rho.data(), 4, rho.size(), o);
  fwrite(delta.data(), 4, delta.size(), o);
  fwrite(nh.data(), 4, nh.size(), o);
  fwrite(seed.data(), 4, seed.size(), o);
  fwrite(ci.data(), 4, ci.size(), o);
  fclose(o);
}

}  // namespace emu_host
