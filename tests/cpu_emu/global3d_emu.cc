// Runs the CLUE3D path (three radix passes, bin table, density, nearest higher, seed ranks, labels)
// on the CPU through the HIP-model emulation, with clue3d_enqueue's launch sequence and grids. Test
// infrastructure only.
// usage: global3d_emu <in> <out> [density variant] [nearest-higher variant] (variants index
// kVariants3D)
//   in : int32 E, Clue3DParams (40 bytes), int64 offsets[E+1], float z, eta, phi, r_over_absz,
//        radius, energy [N], int32 layer, isSilicon [N]
//   out: int32 err, float rho[N], delta_f[N], int32 delta_l[N], nh_layer[N], nh_idx[N], isSeed[N],
//        tracksterIndex[N]
#include <type_traits>

#include "clue3d_kernels.h"
#include "launch_emu.h"

using namespace clue;
using namespace clue::g3d;
using namespace emu_host;

template <int V, class F>
void with_variant(F&& f) {
  constexpr Variant3D v = kVariants3D[V];
  f(std::integral_constant<int, v.items>{}, std::integral_constant<int, v.stage>{});
}
template <class F>
void dispatch(int variant, F&& f) {
  switch (variant) {
    case 1: with_variant<1>(f); break;
    case 2: with_variant<2>(f); break;
    case 3: with_variant<3>(f); break;
    default: with_variant<0>(f); break;
  }
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int E = 0;
  Clue3DParams prm;
  static_assert(sizeof(Clue3DParams) == 40, "parameter record");
  rd(f, &E, 1);
  rd(f, &prm, 1);
  std::vector<int64_t> offs(E + 1);
  rd(f, offs.data(), offs.size());
  const int N = (int)offs[E];
  std::vector<float> z(N), eta(N), phi(N), roz(N), radius(N), energy(N);
  std::vector<int> layer(N), isSi(N);
  for (auto* v : {&z, &eta, &phi, &roz, &radius, &energy}) rd(f, v->data(), N);
  rd(f, layer.data(), N);
  rd(f, isSi.data(), N);
  fclose(f);

  std::vector<BlockDesc> blocks = blocks_of(offs);
  const int nb = (int)blocks.size();
  std::vector<int> evOff(E + 1);
  for (int e = 0; e <= E; ++e) evOff[e] = (int)offs[e];
  const int64_t stride = clue3d_table_stride();
  const int tiles = table_tiles(k3dKeys);
  // device scratch starts with arbitrary contents
  const int P = 0x5A5A5A5A;
  std::vector<uint32_t> k1(N, P), k2(N, P);
  std::vector<int> i1(N, P), i2(N, P), hist((size_t)kRadix * std::max(nb, 1), P);
  const int64_t maxScan = std::max<int64_t>((int64_t)kRadix * nb, N);
  std::vector<int> part(scan_partials(std::max<int64_t>(maxScan, 1)), P);
  std::vector<int> table((size_t)stride * std::max(E, 1), P), tileStart((size_t)(tiles + 1) * std::max(E, 1), P);
  std::vector<float4> cre(N, float4{NAN, NAN, NAN, NAN}), crp(N, float4{NAN, NAN, NAN, NAN});
  std::vector<float> rho(N, -7.f), df(N, -7.f);
  std::vector<int> dl(N, -7), nhl(N, -7), nhi(N, -7), seed(N, -7), ti(N, -7);
  int err = 0;
  if (nb > 0) {
    const int nHist = kRadix * nb;
    Keys3D keys{eta.data(), phi.data(), layer.data(), k2.data(), &err, 0};
    emu::launch(nb, kThreads, [&] { k_radix_hist<Keys3D, DigitPhi>(blocks.data(), keys, hist.data()); });
    scan_exclusive(hist.data(), hist.data(), nHist, part.data());
    emu::launch(nb, kThreads, [&] {
      k_radix_scatter<KeyArray, DigitPhi, EmitFirst>(blocks.data(), KeyArray{k2.data()}, hist.data(),
                                                     EmitFirst{k1.data(), i1.data()});
    });
    emu::launch(nb, kThreads, [&] { k_radix_hist<KeyArray, DigitEta>(blocks.data(), KeyArray{k1.data()}, hist.data()); });
    scan_exclusive(hist.data(), hist.data(), nHist, part.data());
    emu::launch(nb, kThreads, [&] {
      k_radix_scatter<KeyArray, DigitEta, EmitMid>(blocks.data(), KeyArray{k1.data()}, hist.data(),
                                                   EmitMid{i1.data(), k2.data(), i2.data()});
    });
    emu::launch(nb, kThreads, [&] { k_radix_hist<KeyArray, DigitLay>(blocks.data(), KeyArray{k2.data()}, hist.data()); });
    scan_exclusive(hist.data(), hist.data(), nHist, part.data());
    emu::launch(nb, kThreads, [&] {
      k_radix_scatter<KeyArray, DigitLay, EmitLast3D>(
          blocks.data(), KeyArray{k2.data()}, hist.data(),
          EmitLast3D{i2.data(), k1.data(), i1.data(), roz.data(), phi.data(), energy.data(), z.data(), cre.data()});
    });
    {
      const int S = prm.densitySiblingLayers > 0 ? prm.densitySiblingLayers : 0;
      build_table(k1.data(), blocks.data(), nb, evOff.data(), E, k3dKeys, k3dBinsPerLayer,
                  (S < k3dLayers ? S : k3dLayers) + 1, stride, table.data(), tileStart.data());
    }
    const int dv = argc > 3 ? atoi(argv[3]) : 0, nv = argc > 4 ? atoi(argv[4]) : 0;
    dispatch(dv, [&](auto I, auto S) {
      emu::launch(CompShape3D<I>::kPerSort * nb, kThreads, [&] {
        k_density3d<I, S>(blocks.data(), i1.data(), cre.data(), eta.data(), radius.data(), layer.data(), isSi.data(),
                          table.data(), stride, prm, crp.data(), rho.data());
      });
    });
    int* parent = reinterpret_cast<int*>(k2.data());
    dispatch(nv, [&](auto I, auto S) {
      emu::launch(CompShape3D<I>::kPerSort * nb, kThreads, [&] {
        k_nh3d<I, S>(blocks.data(), i1.data(), crp.data(), cre.data(), eta.data(), radius.data(), layer.data(),
                     isSi.data(), table.data(), stride, evOff.data(), prm, df.data(), dl.data(), nhl.data(),
                     nhi.data(), seed.data(), parent);
      });
    });
    SeedRankHost seeds(N);
    launch_labels(blocks.data(), nb, parent, seed.data(), N, evOff.data(), ti.data(), seeds.view(), part.data());
  }
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(&err, 4, 1, o);
  fwrite(rho.data(), 4, N, o);
  fwrite(df.data(), 4, N, o);
  for (auto* v : {&dl, &nhl, &nhi, &seed, &ti}) fwrite(v->data(), 4, N, o);
  fclose(o);
  return 0;
}
