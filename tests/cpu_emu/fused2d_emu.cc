// Runs the CLUE2D fused-path kernels (layer bucket pass, segment lists, fused segment kernels, seed
// numbering) on the CPU through the HIP-model emulation in hip/hip_runtime.h, with the same launch
// sequence as fused2d_enqueue_segments + fused2d_enqueue_cluster. Test infrastructure only.
//
// usage: fused2d_emu <in> <out> [workgroups per class]
//   in : int32 E, float dc, rhoc, odf, int64 offsets[E+1], float x[N], y[N], int32 layer[N], float w[N]
//   out: int32 err, int32 largest segment, float rho[N], delta[N], int32 nh[N], isSeed[N], clusterIndex[N]
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "fused2d_kernels.h"

using namespace clue;
using namespace clue::fused;

namespace {

float sqrt_le(float d) {  // largest v with sqrtf(v) <= d
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

template <class T>
void rd(FILE* f, T* p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}

void excl_scan(std::vector<int>& v) {
  int run = 0;
  for (int& x : v) {
    const int t = x;
    x = run;
    run += t;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int wgsArg = argc > 3 ? atoi(argv[3]) : 0;  // 0: the library's grid (workgroups per CU x 256 CUs)
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int E;
  float dc, rhoc, odf;
  rd(f, &E, 1);
  rd(f, &dc, 1);
  rd(f, &rhoc, 1);
  rd(f, &odf, 1);
  std::vector<int64_t> offs(E + 1);
  rd(f, offs.data(), offs.size());
  const int N = (int)offs[E];
  std::vector<float> x(N), y(N), w(N);
  std::vector<int> layer(N);
  rd(f, x.data(), N);
  rd(f, y.data(), N);
  rd(f, layer.data(), N);
  rd(f, w.data(), N);
  fclose(f);

  // block table (one block per kSortBlock points of each event) and first block of each event
  std::vector<BlockDesc> blocks;
  std::vector<int> evBlk(E + 1), evOff(E + 1);
  for (int e = 0; e < E; ++e) {
    evBlk[e] = (int)blocks.size();
    evOff[e] = (int)offs[e];
    const int nb = (int)((offs[e + 1] - offs[e] + kSortBlock - 1) / kSortBlock);
    const int first = (int)blocks.size();
    for (int k = 0; k < nb; ++k) {
      BlockDesc d;
      d.event = e;
      d.begin = (int)(offs[e] + (int64_t)k * kSortBlock);
      d.end = (int)std::min<int64_t>(offs[e] + (int64_t)(k + 1) * kSortBlock, offs[e + 1]);
      d.histBase = kRadix * first;
      d.nbEvent = nb;
      d.bLocal = k;
      blocks.push_back(d);
    }
  }
  evBlk[E] = (int)blocks.size();
  evOff[E] = N;
  const int nb = (int)blocks.size();
  const int segCap = E * k2dLayers;

  // device scratch starts with arbitrary contents (only ctl is cleared, by a fill kernel)
  std::vector<int> hist((size_t)kRadix * (nb > 0 ? nb : 1), 0x5A5A5A5A), root(N, 0x5A5A5A5A), rank(N, 0x5A5A5A5A),
      ctl(64, 0);
  std::vector<float4> lg(N, float4{NAN, NAN, NAN, __int_as_float(0x5A5A5A5A)});
  std::vector<int4> segs((size_t)kFusedClasses * segCap, int4{0x5A5A5A5A, 0x5A5A5A5A, 0x5A5A5A5A, 0});
  std::vector<float> rho(N, -7.f), delta(N, -7.f);
  std::vector<int> nh(N, -7), isSeed(N, -7), ci(N, -7);
  int err = 0;
  const float dc2 = sqrt_le(dc), dm2 = sqrt_le(odf * dc);

  if (nb > 0) {
    LayerKey key{layer.data(), &err};
    emu::launch(nb, kThreads, [&] { k_radix_hist<LayerKey, DigitLayer>(blocks.data(), key, hist.data()); });
    excl_scan(hist);
    emu::launch(nb, kThreads, [&] {
      k_radix_scatter<LayerKey, DigitLayer, EmitPoint2D>(blocks.data(), key, hist.data(),
                                                         EmitPoint2D{x.data(), y.data(), w.data(), lg.data()});
    });
    emu::launch(E, kRadix, [&] {
      k_segments2d(evOff.data(), evBlk.data(), hist.data(), segs.data(), segCap, ctl.data(), 1, &err);
    });
    std::fill(root.begin(), root.end(), -1);
    auto grid = [&](int perCU) { return (unsigned)(wgsArg > 0 ? wgsArg : std::min(perCU * 256, segCap)); };
    emu::launch(grid(1), 1024, [&] {
      k_fused2d<kFusedCap, 1024>(segs.data() + 2 * (size_t)segCap, ctl.data() + 2, ctl.data() + 6, lg.data(), dc,
                                 dc2, rhoc, odf, dm2, rho.data(), delta.data(), nh.data(), isSeed.data(),
                                 root.data());
    });
    emu::launch(grid(2), 512, [&] {
      k_fused2d<4096, 512>(segs.data() + 1 * (size_t)segCap, ctl.data() + 1, ctl.data() + 5, lg.data(), dc, dc2,
                           rhoc, odf, dm2, rho.data(), delta.data(), nh.data(), isSeed.data(), root.data());
    });
    emu::launch(grid(4), 256, [&] {
      k_fused2d<1024, 256>(segs.data(), ctl.data(), ctl.data() + 4, lg.data(), dc, dc2, rhoc, odf, dm2, rho.data(),
                           delta.data(), nh.data(), isSeed.data(), root.data());
    });
    rank = isSeed;
    excl_scan(rank);
    emu::launch(nb, kThreads, [&] { k_cluster2d(blocks.data(), evOff.data(), root.data(), rank.data(), ci.data()); });
  }

  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  const int big = ctl[8];
  fwrite(&err, 4, 1, o);
  fwrite(&big, 4, 1, o);
  fwrite(rho.data(), 4, N, o);
  fwrite(delta.data(), 4, N, o);
  fwrite(nh.data(), 4, N, o);
  fwrite(isSeed.data(), 4, N, o);
  fwrite(ci.data(), 4, N, o);
  fclose(o);
  return 0;
}
