// CLUE3D (SoA path) on gfx950: tile build, density over sibling layers, nearest-higher on the
// other layers (+ seed/outlier flags) and trackster labels for a packed batch of events.
//
// Reference semantics: src/serial/plugin-CLUETracksterizer/CLUE3DAlgoKernels.h:19-24 (tiles),
// :26-150 (density), :286-392 (nearest higher), :509-580 (seeds and propagation).
// Serial scan order: sibling layer ascending, eta bin ascending, phi bin in window order
// (pb-w .. pb+w, wrapped modulo 126), SoA index ascending. With keys (layer, eta, phi) sorted
// x-major and bins in index order, each (layer, eta row) window is at most two contiguous runs
// of the sorted array, visited in that order, so the fp32 sums and the "last minimum wins"
// updates see candidates in exactly the serial order.
#include "clue3d.h"

#include <float.h>
#include <limits.h>

namespace clue {

namespace {

__device__ __forceinline__ int clamp_layer3d(int L) { return L < 0 ? 0 : (L >= k3dLayers ? k3dLayers - 1 : L); }
__device__ __forceinline__ int clamp_phibin(int p) { return p < 0 ? 0 : (p > k3dPhi ? k3dPhi : p); }

struct Keys3D {
  const float* eta;
  const float* phi;
  const int32_t* layer;
  uint32_t* keyOut;
  int* err;
  __device__ __forceinline__ uint32_t operator()(int i) const {
    int L = layer[i];
    if (L < 0 || L >= k3dLayers) {
      atomicOr(err, kErrLayerRange);
      L = clamp_layer3d(L);
    }
    int pb = phi_bin3d(phi[i]);
    if (pb < 0 || pb > k3dPhi) {
      atomicOr(err, kErrBinRange);
      pb = clamp_phibin(pb);
    }
    // TICLLayerTileT::globalBin = phiBin + etaBin * 126 (TICLLayerTile.h:52); a phiBin of 126 is
    // the first bin of the next row there too.
    const uint32_t k = (uint32_t)(L * k3dBinsPerLayer + eta_bin3d(eta[i]) * k3dPhi + pb);
    keyOut[i] = k;
    return k;
  }
};

// Last pass: also gathers r_over_absz, phi and energy into tile order.
struct EmitLast3D {
  const int* iIn;
  uint32_t* kOut;
  int* iOut;
  const float* roz;
  const float* phi;
  const float* en;
  float* rozS;
  float* phiS;
  float* enS;
  struct Payload {
    int idx;
    float roz, phi, en;
  };
  __device__ __forceinline__ Payload load(int i) const {
    const int idx = iIn[i];
    return Payload{idx, roz[idx], phi[idx], en[idx]};
  }
  __device__ __forceinline__ void store(int pos, uint32_t k, const Payload& v) const {
    kOut[pos] = k;
    iOut[pos] = v.idx;
    rozS[pos] = v.roz;
    phiS[pos] = v.phi;
    enS[pos] = v.en;
  }
};

using DigitPhi = Digit<1, k3dPhi>;
using DigitEta = Digit<k3dPhi, k3dEta>;
using DigitLay = Digit<k3dBinsPerLayer, kRadix>;

// The phi window pb-w .. pb+w (pb in [0, 126], w <= 2) as at most two runs of phi bins of one
// row, in the reference's iteration order ((it % 126) + 126) % 126.
__device__ __forceinline__ int phi_runs(int pb, int w, int* a, int* b) {
  const int lo = pb - w, hi = pb + w;
  if (lo < 0) {
    a[0] = lo + k3dPhi;
    b[0] = k3dPhi - 1;
    a[1] = 0;
    b[1] = hi;
    return 2;
  }
  if (hi > k3dPhi - 1) {
    a[0] = lo;
    b[0] = k3dPhi - 1;
    a[1] = 0;
    b[1] = hi - k3dPhi;
    return 2;
  }
  a[0] = lo;
  b[0] = hi;
  return 1;
}

__global__ __launch_bounds__(kThreads) void k_density3d(
    const BlockDesc* __restrict__ blocks, const int* __restrict__ perm, const float* __restrict__ rozS,
    const float* __restrict__ phiS, const float* __restrict__ enS, const float* __restrict__ z,
    const float* __restrict__ roz, const float* __restrict__ eta, const float* __restrict__ radius,
    const int32_t* __restrict__ layer, const int32_t* __restrict__ isSi, const int* __restrict__ table,
    int64_t stride, Clue3DParams prm, float* __restrict__ rhoS, float* __restrict__ rhoOut) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int* T = table + (int64_t)bd.event * stride;
  const int S = prm.densitySiblingLayers;
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int p = bd.begin + r * kThreads + (int)threadIdx.x;
    if (p >= bd.end) break;
    const int i = perm[p];
    const int L = clamp_layer3d(layer[i]);
    const float zi = z[i];
    const float r0 = roz[i] * zi;
    const float phii = phiS[p];
    const int eb = eta_bin3d(eta[i]);
    const int pb = clamp_phibin(phi_bin3d(phii));
    const float dsq = isSi[i] ? prm.densityXYDistanceSqr : radius[i] * radius[i];
    int minL, maxL;
    if (L < k3dLastPerSide) {
      minL = L - S > 0 ? L - S : 0;
      maxL = L + S < k3dLastPerSide - 1 ? L + S : k3dLastPerSide - 1;
    } else {
      minL = L - S > k3dLastPerSide ? L - S : k3dLastPerSide;
      maxL = L + S < 2 * k3dLastPerSide - 1 ? L + S : 2 * k3dLastPerSide - 1;
    }
    const int emin = eb - 2 > 0 ? eb - 2 : 0;
    const int emax = eb + 2 < k3dEta ? eb + 2 : k3dEta;
    int ra[2], rb[2];
    const int nr = phi_runs(pb, 2, ra, rb);
    float rho = 0.f;
    for (int cl = minL; cl <= maxL; ++cl) {
      const float fsame = (cl == L && !prm.densityOnSameLayer) ? 0.f : 1.f;
      for (int ie = emin; ie <= emax; ++ie) {
        const int row = cl * k3dBinsPerLayer + ie * k3dPhi;
        for (int k = 0; k < nr; ++k) {
          const int s = T[row + ra[k]];
          const int t = T[row + rb[k] + 1];
          for (int q = s; q < t; ++q) {
            const float r1 = rozS[q] * zi;
            const float dp = reduce_range(phii - phiS[q]);
            if ((r0 - r1) * (r0 - r1) + r1 * r1 * dp * dp < dsq)
              rho += ((q == p) ? 1.f : prm.kernelDensityFactor * fsame) * enS[q];
          }
        }
      }
    }
    rhoS[p] = rho;
    rhoOut[i] = rho;
  }
}

__global__ __launch_bounds__(kThreads) void k_nh3d(
    const BlockDesc* __restrict__ blocks, const int* __restrict__ perm, const float* __restrict__ rozS,
    const float* __restrict__ phiS, const float* __restrict__ rhoS, const float* __restrict__ z,
    const float* __restrict__ roz, const float* __restrict__ eta, const float* __restrict__ radius,
    const float* __restrict__ energy, const int32_t* __restrict__ layer, const int32_t* __restrict__ isSi,
    const int* __restrict__ table, int64_t stride, const int* __restrict__ evOff, Clue3DParams prm,
    float* __restrict__ deltaF, int32_t* __restrict__ deltaL, int32_t* __restrict__ nhLayer,
    int32_t* __restrict__ nhIdx, int32_t* __restrict__ isSeedOut, int* __restrict__ parent) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int* T = table + (int64_t)bd.event * stride;
  const int evBase = evOff[bd.event];
  const int S = prm.densitySiblingLayers;
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int p = bd.begin + r * kThreads + (int)threadIdx.x;
    if (p >= bd.end) break;
    const int i = perm[p];
    const int L = clamp_layer3d(layer[i]);
    const float zi = z[i];
    const float r0 = roz[i] * zi;
    const float phii = phiS[p];
    const float rhoi = rhoS[p];
    const int eb = eta_bin3d(eta[i]);
    const int pb = clamp_phibin(phi_bin3d(phii));
    int minL, maxL;
    if (L < k3dLastPerSide) {
      minL = L - S > 0 ? L - S : 0;
      maxL = L + S < k3dLastPerSide - 1 ? L + S : k3dLastPerSide - 1;
    } else {
      minL = L - S > k3dLastPerSide + 1 ? L - S : k3dLastPerSide + 1;
      maxL = L + S < 2 * k3dLastPerSide - 1 ? L + S : 2 * k3dLastPerSide - 1;
    }
    const int emin = eb - 1 > 0 ? eb - 1 : 0;
    const int emax = eb + 1 < k3dEta ? eb + 1 : k3dEta;
    int ra[2], rb[2];
    const int nr = phi_runs(pb, 1, ra, rb);
    float best = FLT_MAX;
    int bcl = -1, bq = -1;
    for (int cl = minL; cl <= maxL; ++cl) {
      if (!prm.nearestHigherOnSameLayer && cl == L) continue;
      for (int ie = emin; ie <= emax; ++ie) {
        const int row = cl * k3dBinsPerLayer + ie * k3dPhi;
        for (int k = 0; k < nr; ++k) {
          const int s = T[row + ra[k]];
          const int t = T[row + rb[k] + 1];
          for (int q = s; q < t; ++q) {
            const float r1 = rozS[q] * zi;
            const float dp = reduce_range(phii - phiS[q]);
            const float dist = (r0 - r1) * (r0 - r1) + r1 * r1 * dp * dp;
            if (rhoS[q] > rhoi && dist <= best) {
              best = dist;
              bcl = cl;
              bq = q;
            }
          }
        }
      }
    }
    float df;
    int dl, hl, hiG;
    if (best != FLT_MAX) {
      df = sqrtf(best);
      dl = bcl > L ? bcl - L : L - bcl;
      hl = bcl;
      hiG = perm[bq];
    } else {
      df = FLT_MAX;
      dl = INT_MAX;
      hl = -1;
      hiG = -1;
    }
    deltaF[i] = df;
    deltaL[i] = dl;
    nhLayer[i] = hl;
    nhIdx[i] = hiG >= 0 ? hiG - evBase : -1;
    // KernelFindAndAssignClustersSoA (CLUE3DAlgoKernels.h:522-543)
    const float cut = isSi[i] ? prm.criticalXYDistance : radius[i];
    const bool seed = (df > cut || (float)dl > prm.criticalZDistanceLyr) && (rhoi >= prm.criticalDensity) &&
                      (energy[i] / rhoi > prm.criticalSelfDensity);
    const bool outlier = (df > prm.outlierMultiplier * prm.criticalXYDistance) && (rhoi < prm.criticalDensity);
    isSeedOut[i] = seed ? 1 : 0;
    parent[i] = (seed || outlier || hl < 0) ? i : hiG;
  }
}

}  // namespace

void clue3d_enqueue(const Clue3DBuffers& b, const Clue3DIO& io, const Clue3DParams& prm, hipStream_t s) {
  if (b.N <= 0 || b.nBlocks <= 0) return;
  const int nHist = kRadix * b.nBlocks;
  Keys3D keys{io.eta, io.phi, io.layer, b.k2, b.err};
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
      EmitLast3D{b.i2, b.k1, b.i1, io.r_over_absz, io.phi, io.energy, b.rozS, b.phiS, b.enS});
  const uint32_t* keyS = b.k1;
  const int* perm = b.i1;
  build_table(keyS, b.blocks, b.nBlocks, b.evOff, b.E, k3dKeys, b.tableStride, b.table, b.tileStart, s);
  k_density3d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, perm, b.rozS, b.phiS, b.enS, io.z, io.r_over_absz, io.eta,
                                             io.radius, io.layer, io.isSilicon, b.table, b.tableStride, prm, b.rhoS,
                                             io.rho);
  int* parent = reinterpret_cast<int*>(b.k2);
  k_nh3d<<<b.nBlocks, kThreads, 0, s>>>(b.blocks, perm, b.rozS, b.phiS, b.rhoS, io.z, io.r_over_absz, io.eta,
                                        io.radius, io.energy, io.layer, io.isSilicon, b.table, b.tableStride, b.evOff,
                                        prm, io.delta_f, io.delta_l, io.nh_layer, io.nh_idx, io.isSeed, parent);
  launch_labels(b.blocks, b.nBlocks, parent, io.isSeed, b.N, b.evOff, io.tracksterIndex, b.seeds, b.part, s);
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
  b.rozS = cv.take<float>(n);
  b.phiS = cv.take<float>(n);
  b.enS = cv.take<float>(n);
  b.rhoS = cv.take<float>(n);
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
  for (int k = 0; k < 8; ++k) add(n * 4);
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
