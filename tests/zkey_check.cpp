// Host check of the z sub-slab key (zkey, desi_amd/csrc/pcf_internal.h) and of
// the step-window rule the work list builds on it (desi_amd/csrc/worklist.h):
//   1. zkey is non-decreasing in z, also across slab and cell edges and
//      outside the grid (clamped);
//   2. zkey / zq is the z cell the grid build assigns (floor((z - lo) / cs),
//      clamped to [0, P));
//   3. a block whose first and last points have keys kf <= kl holds points
//      with keys in [kf, kl] only (the column is sorted by key); every point
//      within hz of one of them along z has a key in [kf - Kf, kl + Kf],
//      Kf = ceil(hz * 1.00001 * zq / cs) + 1, the window the fill takes.
// Prints "zkey_check ok" and the number of cases, or the first violation.
//   zkey_check
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "pcf_internal.h"

namespace {

int cell_of(float z, const pcf::GridGeom& g) {
  int c = (int)std::floor((z - g.lo[2]) * g.inv_cs[2]);
  return c < 0 ? 0 : (c >= g.P[2] ? g.P[2] - 1 : c);
}

pcf::GridGeom geom(int P, float cs, float lo) {
  pcf::GridGeom g{};
  for (int a = 0; a < 3; ++a) {
    g.P[a] = P;
    g.cs[a] = cs;
    g.inv_cs[a] = 1.0f / cs;
    g.lo[a] = lo;
  }
  return g;
}

}  // namespace

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long cases = 0;
  const int zqs[] = {1, 2, 4, 16};
  // 1 and 2: a dense sweep with the neighbours of every slab edge.
  for (int gi = 0; gi < 24; ++gi) {
    const int P = 3 + (int)(U(rng) * 60);
    const float cs = (float)(0.5 + U(rng) * 60.0);
    const float lo = gi % 3 == 0 ? 0.0f : (float)((U(rng) - 0.5) * 500.0);
    const pcf::GridGeom g = geom(P, cs, lo);
    for (int zq : zqs) {
      int prev = -1;
      float zprev = 0.f;
      for (int k = -2 * zq; k <= (P + 2) * zq; ++k) {
        const float e = (float)((double)lo + (double)k * cs / zq);
        const float zs[3] = {std::nextafter(e, -INFINITY), e, std::nextafter(e, INFINITY)};
        for (float z : zs) {
          const int key = pcf::zkey(z, g, zq);
          ++cases;
          if (key < prev) {
            std::printf("FAIL monotonic: P=%d cs=%g lo=%g zq=%d z=%.9g key=%d after z=%.9g key=%d\n",
                        P, cs, lo, zq, z, key, zprev, prev);
            return 1;
          }
          if (key < 0 || key >= P * zq || key / zq != cell_of(z, g)) {
            std::printf("FAIL cell: P=%d cs=%g lo=%g zq=%d z=%.9g key=%d cell=%d\n", P, cs, lo,
                        zq, z, key, cell_of(z, g));
            return 1;
          }
          prev = key;
          zprev = z;
        }
      }
    }
  }
  // 3: the window rule on random blocks, reaches and separations.
  for (int t = 0; t < 400000; ++t) {
    const int P = 3 + (int)(U(rng) * 60);
    const float cs = (float)(0.5 + U(rng) * 60.0);
    const float lo = (float)((U(rng) - 0.5) * 200.0);
    const int zq = zqs[t & 3];
    const pcf::GridGeom g = geom(P, cs, lo);
    const double span = (double)P * cs;
    // A block: its first and last points, their keys, and a point of the
    // block anywhere in the slabs between (its key lies in [kf, kl]).
    const float z1 = (float)(lo + U(rng) * span);
    const float z2 = (float)(z1 + U(rng) * U(rng) * 3.0 * cs);
    const int kf = pcf::zkey(z1, g, zq), kl = pcf::zkey(z2, g, zq);
    const double slab = (double)cs / zq;
    const double zlo = (double)lo + (double)kf * slab, zhi = (double)lo + (double)(kl + 1) * slab;
    float zi = (float)(zlo + U(rng) * (zhi - zlo));
    const int ki = pcf::zkey(zi, g, zq);
    if (ki < kf || ki > kl) continue;  // rounded onto a neighbour slab: not a block point
    const float hz = (float)(U(rng) < 0.1 ? U(rng) * 0.01 * cs : U(rng) * 2.5 * cs);
    const int Kf = (int)std::ceil(hz * 1.00001f * g.inv_cs[2] * (float)zq) + 1;
    // A second point within hz of zi along z (just inside, or anywhere inside).
    const double frac = U(rng) < 0.5 ? 1.0 - 1e-7 : U(rng);
    const double d = (U(rng) < 0.5 ? -1.0 : 1.0) * frac * (double)hz;
    const float zj = (float)((double)zi + d);
    if (!(std::fabs((double)zj - (double)zi) < (double)hz)) continue;
    const int kj = pcf::zkey(zj, g, zq);
    ++cases;
    const int wlo = std::max(0, kf - Kf), whi = std::min(P * zq - 1, kl + Kf);
    if (kj < wlo || kj > whi) {
      std::printf("FAIL window: P=%d cs=%.9g lo=%.9g zq=%d z1=%.9g z2=%.9g zi=%.9g zj=%.9g hz=%.9g "
                  "kf=%d kl=%d kj=%d Kf=%d\n",
                  P, cs, lo, zq, z1, z2, zi, zj, hz, kf, kl, kj, Kf);
      return 1;
    }
  }
  std::printf("zkey_check ok %ld\n", cases);
  return 0;
}
