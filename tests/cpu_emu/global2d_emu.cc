// Runs the CLUE2D global-sort path (layer pass, per-layer bin sort with the bin table, density,
// nearest higher, seed ranks, labels) on the CPU through the HIP-model emulation, with clue2d_enqueue's launch sequence
// and grids. Test infrastructure only.
// usage: global2d_emu <in> <out> [density variant] [nearest-higher variant] [index pass 0/1]
// (formats: launch_emu.h; variants index kVariants2D; the last selects k_nh2d's output form)
#include <type_traits>

#include "clue2d_kernels.h"
#include "launch_emu.h"

using namespace clue;
using namespace clue::g2d;
using namespace emu_host;

template <int V, class F>
void with_variant(F&& f) {
  constexpr Variant2D v = kVariants2D[V];
  f(std::integral_constant<int, v.threads>{}, std::integral_constant<int, v.items>{},
    std::integral_constant<int, v.stage>{});
}
template <class F>
void dispatch(int variant, F&& f) {
  static_assert(kNumVariants2D == 6, "dispatch covers every variant");
  switch (variant) {
    case 1: with_variant<1>(f); break;
    case 2: with_variant<2>(f); break;
    case 3: with_variant<3>(f); break;
    case 4: with_variant<4>(f); break;
    case 5: with_variant<5>(f); break;
    default: with_variant<0>(f); break;
  }
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  Input2D in = read_input(argv[1]);
  const int E = in.E;
  const int N = (int)in.offs[E];
  std::vector<BlockDesc> blocks = blocks_of(in.offs);
  const int nb = (int)blocks.size();
  std::vector<int> evOff(E + 1);
  for (int e = 0; e <= E; ++e) evOff[e] = (int)in.offs[e];
  const int64_t stride = clue2d_table_stride();
  // device scratch starts with arbitrary contents
  const int P = 0x5A5A5A5A;
  std::vector<uint32_t> keyS(N, P);
  std::vector<int> scr(N, P), hist((size_t)kRadix * std::max(nb, 1), P), segB((size_t)std::max(E, 1) * (k2dLayers + 1), P);
  std::vector<float4> lg(N, float4{NAN, NAN, NAN, NAN});
  const int64_t maxScan = std::max<int64_t>((int64_t)kRadix * nb, N);
  std::vector<int> part(scan_partials(std::max<int64_t>(maxScan, 1)), P);
  std::vector<int> table((size_t)stride * std::max(E, 1), P);
  std::vector<uint16_t> table16((size_t)stride * std::max(E, 1), 0x5A5A);
  const BinTables2D tabs{table16.data(), table.data(), segB.data(), stride};
  std::vector<float4> xyw(N, float4{NAN, NAN, NAN, NAN});
  std::vector<uint32_t> rhoKey(N, 0x5A5A5A5Au);
  std::vector<float> rho(N, -7.f), delta(N, -7.f);
  std::vector<int> nh(N, -7), isSeed(N, -7), ci(N, -7);
  int err = 0;
  const float dc = in.dc, rhoc = in.rhoc, odf = in.odf;
  const float dc2 = sqrt_le(dc), dm2 = sqrt_le(odf * dc), dm2up = std::nextafter(dm2, INFINITY);
  if (nb > 0) {
    const int nHist = kRadix * nb;
    const LayerOf2D layerOf{in.layer.data(), &err, 0};
    const int nSeg = E * k2dLayers;
    std::vector<int> segList((size_t)std::max(nSeg, 1), P), segCnt(1, P);
    SeedRankHost seeds(N);  // poisoned: the histogram kernel's prologue clears the words
    const Prologue2D pro{seeds.word.data(), (int64_t)seed_tiles(N) * 32, N, segCnt.data()};
    emu::launch(nb, kThreads, [&] {
      k_radix_hist<LayerOf2D, DigitLayer2D, Prologue2D>(blocks.data(), layerOf, hist.data(), pro);
    });
    scan_exclusive(hist.data(), hist.data(), nHist, part.data());
    emu::launch(nb, kThreads, [&] {
      k_radix_scatter<LayerOf2D, DigitLayer2D, EmitSeg2D>(
          blocks.data(), layerOf, hist.data(),
          EmitSeg2D{in.x.data(), in.y.data(), in.w.data(), lg.data(), segB.data(), segList.data(), segCnt.data()});
    });
    // launch size of the largest class from the event sizes, as clue2d_segment_bounds
    int gridH = 0;
    for (int e = 0; e < E; ++e) gridH += (int)std::min<int64_t>(k2dLayers, (in.offs[e + 1] - in.offs[e]) / (kSegCap + 1));
    emu::launch(nSeg, kSegThreadsS, [&] {
      k_segsort2d<kSegThreadsS, 0, kSegSmall>(evOff.data(), segB.data(), lg.data(), scr.data(), keyS.data(), xyw.data(),
                                              table16.data(), table.data(), stride, nullptr, nullptr);
    });
    emu::launch(nSeg, kSegThreadsL, [&] {
      k_segsort2d<kSegThreadsL, kSegSmall, kSegCap>(evOff.data(), segB.data(), lg.data(), scr.data(), keyS.data(),
                                                    xyw.data(), table16.data(), table.data(), stride, nullptr, nullptr);
    });
    if (gridH > 0)
      emu::launch(gridH, kSegThreadsH, [&] {
        k_segsort2d<kSegThreadsH, kSegCap, kSegNone>(evOff.data(), segB.data(), lg.data(), scr.data(), keyS.data(),
                                                     xyw.data(), table16.data(), table.data(), stride, segList.data(),
                                                     segCnt.data());
      });
    const int nTiles = nb * kTilesPerSort;
    std::vector<int4> reachD(nTiles, int4{P, P, P, P}), reachN(nTiles, int4{P, P, P, P});
    emu::launch((nTiles + kReachThreads - 1) / kReachThreads, kReachThreads, [&] {
      k_reach2d(blocks.data(), nTiles, keyS.data(), tabs, dc, odf, reachD.data(), reachN.data());
    });
    const int dv = argc > 3 ? atoi(argv[3]) : 0, nv = argc > 4 ? atoi(argv[4]) : 0;
    const bool indexPass = argc > 5 && atoi(argv[5]) != 0;
    std::vector<int2> lab(N, int2{P, P});
    dispatch(dv, [&](auto T, auto I, auto S) {
      emu::launch(CompShape<T, I>::kPerSort * nb, T, [&] {
        if (indexPass)
          k_density2d<T, I, S, false>(blocks.data(), xyw.data(), tabs, reachD.data(), dc, dc2, rhoKey.data(),
                                      rho.data(), lab.data());
        else
          k_density2d<T, I, S, true>(blocks.data(), xyw.data(), tabs, reachD.data(), dc, dc2, rhoKey.data(),
                                     rho.data(), lab.data());
      });
    });
    // the links and records reuse the layer-order records and the bin sort's slots, as clue2d.hip does
    int4* rec = reinterpret_cast<int4*>(lg.data());
    int* parent = indexPass ? scr.data() : reinterpret_cast<int*>(lg.data());
    dispatch(nv, [&](auto T, auto I, auto S) {
      emu::launch(CompShape<T, I>::kPerSort * nb, T, [&] {
        if (indexPass)
          k_nh2d<T, I, S, true>(blocks.data(), xyw.data(), rhoKey.data(), tabs, reachN.data(), evOff.data(), dc, rhoc,
                                odf, dm2, dm2up, delta.data(), nh.data(), isSeed.data(), parent, rec,
                                seeds.word.data());
        else
          k_nh2d<T, I, S, false>(blocks.data(), xyw.data(), rhoKey.data(), tabs, reachN.data(), evOff.data(), dc, rhoc,
                                 odf, dm2, dm2up, delta.data(), nh.data(), isSeed.data(), parent, rec,
                                 seeds.word.data());
      });
    });
    if (indexPass) {
      launch_seed_ranks(N, seeds.view(), part.data());
      emu::launch(nb, kThreads, [&] { k_labels_tile2d(blocks.data(), parent, seeds.view(), evOff.data(), lab.data()); });
      emu::launch(nb, kThreads, [&] {
        k_outputs2d(blocks.data(), rec, lab.data(), rho.data(), delta.data(), nh.data(), isSeed.data(), ci.data());
      });
    } else {
      launch_labels(blocks.data(), nb, parent, nullptr, N, evOff.data(), ci.data(), seeds.view(), part.data());
    }
  }
  write_output(argv[2], err, 0, rho, delta, nh, isSeed, ci);
  return 0;
}
