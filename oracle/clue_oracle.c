/*
 * CPU restatement of the reference serial CLUE / CLUE3D algorithms.
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Build: gcc -std=gnu11 -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * Every floating-point expression below is written in the operand order of the reference
 * so that, without FMA contraction, each rounding step matches it.
 */
#include "clue_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ helpers */

/* float -> int32 as x86 cvttss2si does it (the reference's `int v = float_expr;`):
 * truncation toward zero, INT_MIN when the value is NaN or outside the int32 range. */
int oracle_f2i(float v) {
  if (v >= -2147483648.0f && v < 2147483648.0f) return (int)v;
  return INT_MIN;
}

static int d2i(double v) {
  if (v >= -2147483648.0 && v < 2147483648.0) return (int)v;
  return INT_MIN;
}

/* LayerTilesConstants.h:7-27: minX=-250, maxX=250, tileSize=5, nColumns=nRows=100,
 * rX = nColumns / (maxX - minX) evaluated in float. */
#define L2D_NLAYERS 100
#define L2D_NCOLS 100
#define L2D_NROWS 100
#define L2D_NBINS_LAYER (L2D_NCOLS * L2D_NROWS)

static const float L2D_RX = (float)L2D_NCOLS / (250.f - (-250.f));

/* LayerTilesSerial.h:24-31 */
int oracle_xbin(float x) {
  int b = oracle_f2i((x - (-250.f)) * L2D_RX);
  if (b > L2D_NCOLS - 1) b = L2D_NCOLS - 1;
  if (b < 0) b = 0;
  return b;
}

/* LayerTilesSerial.h:33-40 (same constants in y) */
int oracle_ybin(float y) {
  int b = oracle_f2i((y - (-250.f)) * L2D_RX);
  if (b > L2D_NROWS - 1) b = L2D_NROWS - 1;
  if (b < 0) b = 0;
  return b;
}

/* Math/deltaPhi.h:17-24, T = float: o2pi = float(1/(2 pi)); |x| <= float(pi) -> x;
 * else x - roundf(x*o2pi) * float(2 pi). */
float oracle_reduce_range(float x) {
  const float o2pi = (float)(1. / (2. * M_PI));
  if (fabsf(x) <= (float)M_PI) return x;
  float n = roundf(x * o2pi);
  return x - n * (float)(2. * M_PI);
}

/* Common.h:9-16 + TICLLayerTile.h:15-22: r = nEtaBins / (maxEta - minEta) in float,
 * bin = int((|eta| - minEta) * r) clamped to [0, 33]. */
#define T3D_NETA 34
#define T3D_NPHI 126
#define T3D_NLAYERS 94
#define T3D_NBINS_LAYER (T3D_NETA * T3D_NPHI)

int oracle_eta_bin(float eta) {
  const float r = (float)T3D_NETA / (3.2f - 1.5f);
  int b = oracle_f2i((fabsf(eta) - 1.5f) * r);
  if (b < 0) b = 0;
  if (b > T3D_NETA - 1) b = T3D_NETA - 1;
  return b;
}

/* TICLLayerTile.h:24-30: normPhi = reduceRange(phi) (float);
 * r = float(126 * M_1_PI * 0.5f) (the product is formed in double);
 * bin = int((double(normPhi) + M_PI) * double(r)). */
int oracle_phi_bin(float phi) {
  float normPhi = oracle_reduce_range(phi);
  float r = (float)(T3D_NPHI * M_1_PI * 0.5f);
  return d2i(((double)normPhi + M_PI) * (double)r);
}

/* CSR tile: bins listed in ascending point index (a stable counting sort over i). */
typedef struct {
  int32_t* start; /* nbins + 1 */
  int32_t* items; /* n */
} csr_t;

static int csr_build(csr_t* t, int32_t n, int32_t nbins, const int32_t* key) {
  t->start = (int32_t*)calloc((size_t)nbins + 1, sizeof(int32_t));
  t->items = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!t->start || !t->items) return -1;
  for (int32_t i = 0; i < n; ++i) t->start[key[i] + 1]++;
  for (int32_t b = 0; b < nbins; ++b) t->start[b + 1] += t->start[b];
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)nbins);
  if (!cur) return -1;
  memcpy(cur, t->start, sizeof(int32_t) * (size_t)nbins);
  for (int32_t i = 0; i < n; ++i) t->items[cur[key[i]]++] = i;
  free(cur);
  return 0;
}

static void csr_free(csr_t* t) {
  free(t->start);
  free(t->items);
}

/* Seed numbering + DFS label propagation shared by both algorithms
 * (CLUEAlgoKernels.h:108-153, CLUE3DAlgoKernels.h:509-580): followers[p] lists i with
 * parent p (registered in ascending i); seeds are numbered in ascending index; labels flow
 * from each seed down its follower tree. follow[i] < 0 means "not registered". */
static int propagate(int32_t n, const int32_t* isSeedFlag, const int32_t* follow, int32_t* label) {
  int32_t* cnt = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  int32_t* lst = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!cnt || !lst || !stack) return -1;
  for (int32_t i = 0; i < n; ++i)
    if (follow[i] >= 0) cnt[follow[i] + 1]++;
  for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!cur) return -1;
  memcpy(cur, cnt, sizeof(int32_t) * (size_t)n);
  for (int32_t i = 0; i < n; ++i)
    if (follow[i] >= 0) lst[cur[follow[i]]++] = i;
  free(cur);
  int32_t sp = 0, nClusters = 0;
  for (int32_t i = 0; i < n; ++i) {
    label[i] = -1;
    if (isSeedFlag[i]) {
      label[i] = nClusters++;
      stack[sp++] = i;
    }
  }
  while (sp > 0) {
    int32_t i = stack[--sp];
    for (int32_t k = cnt[i]; k < cnt[i + 1]; ++k) {
      int32_t j = lst[k];
      label[j] = label[i];
      stack[sp++] = j;
    }
  }
  free(cnt);
  free(lst);
  free(stack);
  return 0;
}

/* ------------------------------------------------------------------ CLUE2D */

/* CLUEAlgoKernels.h:7-16 */
static inline float dist2d(const float* x, const float* y, const int32_t* layer, int32_t i, int32_t j) {
  if (layer[i] == layer[j]) {
    const float dx = x[i] - x[j];
    const float dy = y[i] - y[j];
    return sqrtf(dx * dx + dy * dy);
  }
  return FLT_MAX;
}

int oracle_clue2d(int32_t n, const float* x, const float* y, const int32_t* layer, const float* weight,
                  float dc, float rhoc, float outlierDeltaFactor,
                  float* rho, float* delta, int32_t* nearestHigher, int32_t* isSeed, int32_t* clusterIndex) {
  for (int32_t i = 0; i < n; ++i)
    if (layer[i] < 0 || layer[i] >= L2D_NLAYERS) return -1;

  /* kernel_compute_histogram (CLUEAlgoKernels.h:18-23): bin = (layer, xBin, yBin) */
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!key) return -2;
  for (int32_t i = 0; i < n; ++i)
    key[i] = layer[i] * L2D_NBINS_LAYER + oracle_xbin(x[i]) * L2D_NROWS + oracle_ybin(y[i]);
  csr_t t;
  if (csr_build(&t, n, L2D_NLAYERS * L2D_NBINS_LAYER, key)) return -2;

  /* kernel_calculate_density (CLUEAlgoKernels.h:25-55) */
  for (int32_t i = 0; i < n; ++i) {
    float r = 0.f;
    const int32_t base = layer[i] * L2D_NBINS_LAYER;
    const int xlo = oracle_xbin(x[i] - dc), xhi = oracle_xbin(x[i] + dc);
    const int ylo = oracle_ybin(y[i] - dc), yhi = oracle_ybin(y[i] + dc);
    for (int xb = xlo; xb < xhi + 1; ++xb)
      for (int yb = ylo; yb < yhi + 1; ++yb) {
        const int32_t b = base + xb * L2D_NROWS + yb;
        for (int32_t k = t.start[b]; k < t.start[b + 1]; ++k) {
          const int32_t j = t.items[k];
          if (dist2d(x, y, layer, i, j) <= dc) r += (i == j ? 1.f : 0.5f) * weight[j];
        }
      }
    rho[i] = r;
  }

  /* kernel_calculate_distanceToHigher (CLUEAlgoKernels.h:57-106) */
  const float dm = outlierDeltaFactor * dc;
  for (int32_t i = 0; i < n; ++i) {
    float delta_i = FLT_MAX;
    int32_t nh_i = -1;
    const float xi = x[i], yi = y[i], rho_i = rho[i];
    const int32_t base = layer[i] * L2D_NBINS_LAYER;
    const int xlo = oracle_xbin(xi - dm), xhi = oracle_xbin(xi + dm);
    const int ylo = oracle_ybin(yi - dm), yhi = oracle_ybin(yi + dm);
    for (int xb = xlo; xb < xhi + 1; ++xb)
      for (int yb = ylo; yb < yhi + 1; ++yb) {
        const int32_t b = base + xb * L2D_NROWS + yb;
        for (int32_t k = t.start[b]; k < t.start[b + 1]; ++k) {
          const int32_t j = t.items[k];
          int higher = rho[j] > rho_i;
          higher = higher || ((rho[j] == rho_i) && (j > i));
          const float d = dist2d(x, y, layer, i, j);
          if (higher && d <= dm) {
            if (d < delta_i) {
              delta_i = d;
              nh_i = j;
            }
          }
        }
      }
    delta[i] = delta_i;
    nearestHigher[i] = nh_i;
  }
  csr_free(&t);

  /* kernel_findAndAssign_clusters (CLUEAlgoKernels.h:108-153) */
  int32_t* follow = key; /* reuse */
  for (int32_t i = 0; i < n; ++i) {
    const int seed = (delta[i] > dc) && (rho[i] >= rhoc);
    const int outlier = (delta[i] > outlierDeltaFactor * dc) && (rho[i] < rhoc);
    isSeed[i] = seed ? 1 : 0;
    /* a non-seed, non-outlier point always has nh >= 0 for finite dc*odf; a negative nh
     * (reference: out-of-range followers[-1]) is treated as unregistered. */
    follow[i] = (!seed && !outlier) ? nearestHigher[i] : -1;
  }
  int rc = propagate(n, isSeed, follow, clusterIndex);
  free(key);
  return rc ? -2 : 0;
}

/* ------------------------------------------------------------------ CLUE3D */

static inline float dphi3d(float a, float b) { return oracle_reduce_range(a - b); }

int oracle_clue3d(int32_t n, const float* z, const float* eta, const float* phi, const float* r_over_absz,
                  const float* radius, const float* energy, const int32_t* layer, const int32_t* isSilicon,
                  const oracle_clue3d_params* p,
                  float* rho, float* delta_f, int32_t* delta_l, int32_t* nh_layer, int32_t* nh_idx,
                  int32_t* isSeed, int32_t* tracksterIndex) {
  for (int32_t i = 0; i < n; ++i)
    if (layer[i] < 0 || layer[i] >= T3D_NLAYERS) return -1;
  const int S = p->densitySiblingLayers;
  const int last = T3D_NLAYERS / 2; /* lastLayerPerSide = 47 */

  /* KernelComputeHistogramSoA (CLUE3DAlgoKernels.h:19-24). Keys are flat
   * (layer * 4284 + etaBin * 126 + phiBin); one spare row of 126 empty bins follows the
   * last layer. A window row etaBin == 34 (CLUE3DAlgoKernels.h:73,329 clamp to nEtaBins,
   * not nEtaBins-1) therefore reads bins (layer+1, eta 0, phi), which is the memory the
   * reference's contiguous tile array holds there; after layer 93 it reads empty bins. */
  const int32_t nbins = T3D_NLAYERS * T3D_NBINS_LAYER + T3D_NPHI;
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* eb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* pb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!key || !eb || !pb) return -2;
  for (int32_t i = 0; i < n; ++i) {
    eb[i] = oracle_eta_bin(eta[i]);
    pb[i] = oracle_phi_bin(phi[i]);
    key[i] = layer[i] * T3D_NBINS_LAYER + eb[i] * T3D_NPHI + pb[i];
  }
  csr_t t;
  if (csr_build(&t, n, nbins, key)) return -2;

  /* KernelCalculateDensitySoA (CLUE3DAlgoKernels.h:26-150) */
  for (int32_t i = 0; i < n; ++i) {
    const int L = layer[i];
    int minL, maxL;
    if (L < last) {
      minL = L - S > 0 ? L - S : 0;
      maxL = L + S < last - 1 ? L + S : last - 1;
    } else {
      minL = L - S > last ? L - S : last;
      maxL = L + S < 2 * last - 1 ? L + S : 2 * last - 1;
    }
    const float r0 = r_over_absz[i] * z[i];
    const float dsq = isSilicon[i] ? p->densityXYDistanceSqr : radius[i] * radius[i];
    float r = 0.f;
    for (int cl = minL; cl <= maxL; ++cl) {
      const int onSame = (cl == L);
      const int emin = eb[i] - 2 > 0 ? eb[i] - 2 : 0;
      const int emax = eb[i] + 2 < T3D_NETA ? eb[i] + 2 : T3D_NETA;
      for (int ie = emin; ie <= emax; ++ie)
        for (int it = pb[i] - 2; it <= pb[i] + 2; ++it) {
          const int ip = ((it % T3D_NPHI) + T3D_NPHI) % T3D_NPHI;
          const int32_t b = cl * T3D_NBINS_LAYER + ie * T3D_NPHI + ip;
          if (b >= nbins) continue;
          for (int32_t k = t.start[b]; k < t.start[b + 1]; ++k) {
            const int32_t j = t.items[k];
            const float r1 = r_over_absz[j] * z[i];
            const float dp = dphi3d(phi[i], phi[j]);
            if ((r0 - r1) * (r0 - r1) + r1 * r1 * dp * dp < dsq) {
              const float fsame = (onSame && !p->densityOnSameLayer) ? 0.f : 1.f;
              r += ((i == j) ? 1.f : p->kernelDensityFactor * fsame) * energy[j];
            }
          }
        }
    }
    rho[i] = r;
  }

  /* KernelComputeDistanceToHigherSoA (CLUE3DAlgoKernels.h:286-392) */
  for (int32_t i = 0; i < n; ++i) {
    const int L = layer[i];
    int minL, maxL;
    if (L < last) {
      minL = L - S > 0 ? L - S : 0;
      maxL = L + S < last - 1 ? L + S : last - 1;
    } else {
      minL = L - S > last + 1 ? L - S : last + 1;
      maxL = L + S < 2 * last - 1 ? L + S : 2 * last - 1;
    }
    const float r0 = r_over_absz[i] * z[i];
    float best = FLT_MAX;
    float bf = FLT_MAX;
    int32_t bl = INT_MAX, hl = -1, hi = -1;
    for (int cl = minL; cl <= maxL; ++cl) {
      if (!p->nearestHigherOnSameLayer && cl == L) continue;
      const int emin = eb[i] - 1 > 0 ? eb[i] - 1 : 0;
      const int emax = eb[i] + 1 < T3D_NETA ? eb[i] + 1 : T3D_NETA;
      for (int ie = emin; ie <= emax; ++ie)
        for (int it = pb[i] - 1; it <= pb[i] + 1; ++it) {
          const int ip = ((it % T3D_NPHI) + T3D_NPHI) % T3D_NPHI;
          const int32_t b = cl * T3D_NBINS_LAYER + ie * T3D_NPHI + ip;
          if (b >= nbins) continue;
          for (int32_t k = t.start[b]; k < t.start[b + 1]; ++k) {
            const int32_t j = t.items[k];
            const float r1 = r_over_absz[j] * z[i];
            const float dp = dphi3d(phi[i], phi[j]);
            const float dist = (r0 - r1) * (r0 - r1) + r1 * r1 * dp * dp;
            if (rho[j] > rho[i] && dist <= best) {
              best = dist;
              bf = sqrtf(dist);
              bl = abs(cl - L);
              hl = cl;
              hi = j;
            }
          }
        }
    }
    if (best != FLT_MAX) {
      delta_f[i] = bf;
      delta_l[i] = bl;
      nh_layer[i] = hl;
      nh_idx[i] = hi;
    } else {
      delta_f[i] = FLT_MAX;
      delta_l[i] = INT_MAX;
      nh_layer[i] = -1;
      nh_idx[i] = -1;
    }
  }
  csr_free(&t);

  /* KernelFindAndAssignClustersSoA (CLUE3DAlgoKernels.h:509-580) */
  int32_t* follow = key;
  const float outlierDelta = p->outlierMultiplier * p->criticalXYDistance;
  for (int32_t i = 0; i < n; ++i) {
    const float cut = isSilicon[i] ? p->criticalXYDistance : radius[i];
    const int seed = (delta_f[i] > cut || (float)delta_l[i] > p->criticalZDistanceLyr) &&
                     (rho[i] >= p->criticalDensity) && (energy[i] / rho[i] > p->criticalSelfDensity);
    const int outlier = (delta_f[i] > outlierDelta) && (rho[i] < p->criticalDensity);
    isSeed[i] = seed ? 1 : 0;
    follow[i] = (!seed && !outlier && nh_layer[i] >= 0) ? nh_idx[i] : -1;
  }
  int rc = propagate(n, isSeed, follow, tracksterIndex);
  free(key);
  free(eb);
  free(pb);
  return rc ? -2 : 0;
}
