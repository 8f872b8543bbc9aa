// pcf_oracle.cpp -- CPU restatement of 3PCFGPU's pair/triplet histogram semantics.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
// check and bench.py's cpu_baseline leg may load it.  The product path
// (desi_amd/libpcf_hip.so) never calls into it.
//
// Two arithmetic modes are restated:
//
//   * FP64 canonical ("fp64"): the reference GPU kernels' rules (cuda/kernels/*)
//     with every distance formed in IEEE double from the FP32 inputs,
//     (x2-x1)^2 + (y2-y1)^2 + (z2-z1)^2 evaluated left to right without
//     contraction, sqrt correctly rounded, bin = (int)(sqrt(d2) * ds).  Weight
//     products are formed in double (exact for two FP32 factors).
//   * cpp32: the CPU programs' rules (cpp/2PCFisotropic/normal/2PCF.h) with FP32
//     distances and no FMA contraction, unordered pairs counted twice.  It is
//     compared bit-for-bit with the compiled reference programs (oracle/_ref).
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).
//
// Pinning status of each entry point is listed in the header comment of the
// function; functions whose pin is "formula" only are marked parity unpinned
// where that applies.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Pt {
  float x, y, z, w;
};

// ---------------------------------------------------------------------------
// Rules shared by the pair kernels.
// ---------------------------------------------------------------------------

enum ZeroRule {
  ZERO_NONE = 0,      // XY kernels: d = 0 counted (cuda/kernels/2iso_k.cuh:107)
  ZERO_ISO = 1,       // XX: 0 < d2 (cuda/kernels/2iso_k.cuh:50)
  ZERO_ANI_BOTH = 2,  // 0 < dz2 and 0 < dort2 (cuda/kernels/2ani_k.cuh:52)
};

struct PairRule {
  int ani;
  int bpc;
  int bins;
  double R;   // (double)(float)(dmax*dmax)   (cuda/kernels/2iso_k.cuh:35)
  double ds;  // bin scale
  int le;     // range test d2 <= R (cuda/kernels/2aniBPC_k.cuh:63) instead of <
  int zero;   // ZeroRule
  double L;   // periodic box side
};

// ds = floor(bn/dmax * 1e6) / 1e6 in double (cuda/kernels/2iso_k.cuh:37).
double ds_floored(int bins, float dmax) {
  return std::floor(((double)bins / (double)dmax) * 1000000.0) / 1000000.0;
}

PairRule make_pair_rule(int ani, int bpc, int auto_, int bins, float dmax, float box) {
  PairRule r;
  r.ani = ani;
  r.bpc = bpc;
  r.bins = bins;
  float dd_max = dmax * dmax;  // FP32 product, as the kernels compute it
  r.R = (double)dd_max;
  r.L = (double)box;
  if (!ani) {
    // 2iso-BPC uses the unfloored scale bn/dmax (cuda/kernels/2isoBPC_k.cuh:36).
    r.ds = bpc ? (double)bins / (double)dmax : ds_floored(bins, dmax);
    r.le = 0;
    r.zero = auto_ ? ZERO_ISO : ZERO_NONE;
  } else {
    r.ds = ds_floored(bins, dmax);  // cuda/kernels/2aniBPC_k.cuh:39 floors too
    r.le = bpc ? 1 : 0;             // cuda/kernels/2aniBPC_k.cuh:63 vs 2ani_k.cuh:52
    // Non-periodic 2ani excludes zero components in XX and XY alike
    // (cuda/kernels/2ani_k.cuh:52,123); the periodic XY kernel does not
    // (cuda/kernels/2aniBPC_k.cuh:401).
    if (!bpc) r.zero = ZERO_ANI_BOTH;
    else r.zero = auto_ ? ZERO_ANI_BOTH : ZERO_NONE;
  }
  return r;
}

inline bool in_range(double d2, const PairRule& r) { return r.le ? (d2 <= r.R) : (d2 < r.R); }

// Per-axis squared separation for image k of point j relative to point i.
// k == 0: (a_j - a_i)^2.  k != 0: the reference's projected separation
// size_box - |a_j - a_i| (cuda/kernels/2isoBPC_k.cuh:87,115,143).  The
// reference only ever evaluates the image that brings the two points closer,
// i.e. k = -sign(a_j - a_i); the opposite image (separation L + |delta|) is
// not a reference term and is rejected.  For delta == 0 the two images are
// the same term; it is kept for k = -1 only.
inline bool axis_sq(float ai, float aj, int k, double L, double* out) {
  double d = (double)aj - (double)ai;
  if (k == 0) {
    *out = d * d;
    return true;
  }
  if ((k > 0 && d >= 0.0) || (k < 0 && d < 0.0)) return false;
  double t = L - std::fabs(d);
  *out = t * t;
  return true;
}

// One (ordered pair, image) evaluation in the FP64 canonical mode.  Returns the
// flat bin or -1.
inline int pair_bin(const Pt& pi, const Pt& pj, int kx, int ky, int kz, const PairRule& r) {
  double tx, ty, tz;
  if (!axis_sq(pi.x, pj.x, kx, r.L, &tx)) return -1;
  if (!axis_sq(pi.y, pj.y, ky, r.L, &ty)) return -1;
  if (!axis_sq(pi.z, pj.z, kz, r.L, &tz)) return -1;
  const bool regular = (kx == 0 && ky == 0 && kz == 0);
  if (!r.ani) {
    double d2 = tx + ty + tz;  // left to right, as (x2-x1)^2+(y2-y1)^2+(z2-z1)^2
    if (!in_range(d2, r)) return -1;
    if (r.zero == ZERO_ISO && regular && !(d2 > 0.0)) return -1;
    int b = (int)(std::sqrt(d2) * r.ds);
    if (b > r.bins - 1) return -1;
    return b;
  }
  double dz2 = tz;
  double do2 = tx + ty;
  if (!in_range(dz2, r) || !in_range(do2, r)) return -1;
  if (r.zero == ZERO_ANI_BOTH && !(dz2 > 0.0 && do2 > 0.0)) return -1;
  int bz = (int)(std::sqrt(dz2) * r.ds);  // cuda/kernels/2ani_k.cuh:54-57
  if (bz > r.bins - 1) return -1;
  int bo = (int)(std::sqrt(do2) * r.ds);
  if (bo > r.bins - 1) return -1;
  return bz * r.bins + bo;
}

// ---------------------------------------------------------------------------
// Uniform cell grid used only to skip far pairs; the cell edge is >= dmax so
// every pair within range lies in adjacent cells (27 neighbour images).
// ---------------------------------------------------------------------------
struct Grid {
  int P[3];
  double lo[3];
  double cs[3];
  std::vector<int64_t> start;  // P0*P1*P2 + 1
  std::vector<Pt> pts;

  int cell_of(float v, int a) const {
    double t = ((double)v - lo[a]) / cs[a];
    int c = (int)std::floor(t);
    if (c < 0) c = 0;
    if (c >= P[a]) c = P[a] - 1;
    return c;
  }
  int64_t lin(int cx, int cy, int cz) const { return ((int64_t)cx * P[1] + cy) * P[2] + cz; }

  void build(const Pt* p, int64_t n) {
    int64_t nc = (int64_t)P[0] * P[1] * P[2];
    std::vector<int64_t> cnt(nc + 1, 0);
    std::vector<int64_t> cid(n);
    for (int64_t i = 0; i < n; ++i) {
      cid[i] = lin(cell_of(p[i].x, 0), cell_of(p[i].y, 1), cell_of(p[i].z, 2));
      cnt[cid[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    start = cnt;
    pts.resize(n);
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; ++i) pts[cur[cid[i]]++] = p[i];
  }
};

// Geometry shared by the two catalogs of a job.
void grid_geometry(const Pt* a, int64_t na, const Pt* b, int64_t nb, double reach, int bpc,
                   double L, Grid* g) {
  for (int ax = 0; ax < 3; ++ax) {
    if (bpc) {
      int P = (int)std::floor(L / reach);
      if (P < 1) P = 1;
      if (P > 256) P = 256;
      g->P[ax] = P;
      g->lo[ax] = 0.0;
      g->cs[ax] = L / P;
    } else {
      double mn = 1e300, mx = -1e300;
      auto scan = [&](const Pt* p, int64_t n) {
        for (int64_t i = 0; i < n; ++i) {
          double v = ax == 0 ? p[i].x : (ax == 1 ? p[i].y : p[i].z);
          mn = std::min(mn, v);
          mx = std::max(mx, v);
        }
      };
      scan(a, na);
      if (b) scan(b, nb);
      if (!(mx >= mn)) { mn = 0; mx = 1; }
      double ext = mx - mn;
      int P = (int)std::floor(ext / reach);
      if (P < 1) P = 1;
      if (P > 256) P = 256;
      g->P[ax] = P;
      g->lo[ax] = mn;
      g->cs[ax] = std::max(ext / P, reach);
      if (g->cs[ax] * P < ext) g->cs[ax] = ext / P * (1 + 1e-12);
    }
  }
}

// Adjacent-image enumeration: calls f(cell_b_linear, kx, ky, kz) for every
// neighbour cell image of cell (cx,cy,cz).
template <class F>
void for_neighbours(const Grid& g, int bpc, int cx, int cy, int cz, F f) {
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        int c[3] = {cx + dx, cy + dy, cz + dz};
        int k[3] = {0, 0, 0};
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
          if (c[a] < 0) {
            if (!bpc) { ok = false; break; }
            c[a] += g.P[a];
            k[a] = -1;
          } else if (c[a] >= g.P[a]) {
            if (!bpc) { ok = false; break; }
            c[a] -= g.P[a];
            k[a] = 1;
          }
        }
        if (!ok) continue;
        f(g.lin(c[0], c[1], c[2]), k[0], k[1], k[2]);
      }
}

int nthreads_or(int nt) {
#ifdef _OPENMP
  return nt > 0 ? nt : omp_get_max_threads();
#else
  (void)nt;
  return 1;
#endif
}

}  // namespace

// The value of a 64-bit integer product or sum as the reference's 32-bit
// signed int holds it (two's complement wrap, what nvcc and g++ produce for
// the reference's overflowing int expressions).
inline int32_t wrap_i32(uint64_t v) { return (int32_t)(uint32_t)(v & 0xffffffffull); }

// 3i^2 + 3i + 1, the shell count of the analytic RR, exact or wrapped.
inline int64_t shell_int(int64_t i, bool wrap32) {
  if (!wrap32) return 3 * i * i + 3 * i + 1;
  const uint32_t u = (uint32_t)i;
  return (int64_t)(int32_t)(3u * u * u + 3u * u + 1u);
}

extern "C" {

// Ordered-pair 2PCF histogram in the FP64 canonical mode.
//   ani = 0: 2iso, flat bin b(d)       (cuda/kernels/2iso_k.cuh:8-120,
//                                        cuda/kernels/2isoBPC_k.cuh:8-557)
//   ani = 1: 2ani, bin b(|dz|)*bn + b(d_perp) (cuda/kernels/2ani_k.cuh:7-135,
//                                        cuda/kernels/2aniBPC_k.cuh:8-682)
//   auto_ = 1: XX over catalog a (b ignored), else XY from a to b.
// Weighted sums use w_i*w_j in double; the reference's XY2ani weight-array
// defect (cuda/kernels/2ani_k.cuh:128) is not reproduced.
// Pinned by: brute force (pcf_oracle_pairs_brute), dyadic-lattice catalogs
// with exact integer distances, the cpp32 comparison against compiled cpp/,
// translation invariance under BPC.
int pcf_oracle_pairs(const float* a_, int64_t na, const float* b_, int64_t nb, int auto_, int ani,
                     int bpc, float box, int bins, float dmax, int nthreads, uint64_t* counts,
                     double* wsum) {
  if (bins <= 0 || !(dmax > 0) || na < 0 || (!auto_ && nb < 0)) return -1;
  if (bpc && !(dmax < box)) return -2;
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  PairRule r = make_pair_rule(ani, bpc, auto_, bins, dmax, box);
  int nout = ani ? bins * bins : bins;
  std::memset(counts, 0, sizeof(uint64_t) * nout);
  std::memset(wsum, 0, sizeof(double) * nout);
  if (na == 0 || nb == 0) return 0;

  Grid g;
  grid_geometry(a, na, auto_ ? nullptr : b, nb, (double)dmax * 1.0000001 + 1e-6, bpc, r.L, &g);
  Grid gb = g;
  g.build(a, na);
  gb.build(b, nb);
  const int nt = nthreads_or(nthreads);
  const int64_t ncell = (int64_t)g.P[0] * g.P[1] * g.P[2];
  std::vector<uint64_t> c_all((size_t)nt * nout, 0);
  std::vector<double> w_all((size_t)nt * nout, 0.0);
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    uint64_t* C = &c_all[(size_t)tid * nout];
    double* W = &w_all[(size_t)tid * nout];
#pragma omp for schedule(dynamic, 1)
    for (int64_t ca = 0; ca < ncell; ++ca) {
      int cx = (int)(ca / ((int64_t)g.P[1] * g.P[2]));
      int cy = (int)((ca / g.P[2]) % g.P[1]);
      int cz = (int)(ca % g.P[2]);
      for_neighbours(gb, bpc, cx, cy, cz, [&](int64_t cb, int kx, int ky, int kz) {
        for (int64_t i = g.start[ca]; i < g.start[ca + 1]; ++i) {
          const Pt& pi = g.pts[i];
          for (int64_t j = gb.start[cb]; j < gb.start[cb + 1]; ++j) {
            const Pt& pj = gb.pts[j];
            int bin = pair_bin(pi, pj, kx, ky, kz, r);
            if (bin < 0) continue;
            C[bin] += 1;
            W[bin] += (double)pi.w * (double)pj.w;
          }
        }
      });
    }
  }
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < nout; ++k) {
      counts[k] += c_all[(size_t)t * nout + k];
      wsum[k] += w_all[(size_t)t * nout + k];
    }
  return 0;
}

// Same histogram by brute force over all ordered pairs and all 8 image
// subsets (the reference's regular term plus its z, y, x, xy, xz, yz, xyz
// projections, cuda/kernels/2isoBPC_k.cuh:50-280).  O(n^2); small inputs only.
// Used to pin the cell-grid enumeration of pcf_oracle_pairs.
int pcf_oracle_pairs_brute(const float* a_, int64_t na, const float* b_, int64_t nb, int auto_,
                           int ani, int bpc, float box, int bins, float dmax, uint64_t* counts,
                           double* wsum) {
  if (bins <= 0 || !(dmax > 0)) return -1;
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  PairRule r = make_pair_rule(ani, bpc, auto_, bins, dmax, box);
  int nout = ani ? bins * bins : bins;
  std::memset(counts, 0, sizeof(uint64_t) * nout);
  std::memset(wsum, 0, sizeof(double) * nout);
  for (int64_t i = 0; i < na; ++i)
    for (int64_t j = 0; j < nb; ++j) {
      const Pt& pi = a[i];
      const Pt& pj = b[j];
      int nsub = bpc ? 8 : 1;
      for (int s = 0; s < nsub; ++s) {
        // Subset s selects the projected axes; the projection of an axis is
        // the image k = -sign(delta) (k = -1 for delta == 0).
        int k[3] = {0, 0, 0};
        float ai[3] = {pi.x, pi.y, pi.z}, aj[3] = {pj.x, pj.y, pj.z};
        for (int ax = 0; ax < 3; ++ax)
          if (s & (1 << ax)) k[ax] = ((double)aj[ax] - (double)ai[ax]) > 0 ? -1 : 1;
        for (int ax = 0; ax < 3; ++ax)
          if ((s & (1 << ax)) && ((double)aj[ax] - (double)ai[ax]) == 0.0) k[ax] = -1;
        int bin = pair_bin(pi, pj, k[0], k[1], k[2], r);
        if (bin < 0) continue;
        counts[bin] += 1;
        wsum[bin] += (double)pi.w * (double)pj.w;
      }
    }
  return 0;
}

// cpp32 mode: the CPU program cpp/2PCFisotropic/normal/2PCF.h.
//   XX (make_histoXX, 2PCF.h:151-305): unordered pairs i<j, FP32
//   dx = x_i - x_j etc., dis = dx*dx+dy*dy+dz*dz in FP32 without contraction,
//   dis < dd_max (FP32 dd_max), bin = (int)(sqrt((double)dis)*ds) with floored
//   ds (2PCF.h:75), deposit 2*w1*w2 in FP32 added to a double (2PCF.h:198).
//   d = 0 pairs are counted (no d > 0 test, 2PCF.h:197).
//   XY (make_histoXY, 2PCF.h:308-384): all pairs, deposit w1*w2.
// The node prune of the program is conservative, so the histogram equals
// this direct enumeration.  Pinned bit-for-bit against the compiled program.
int pcf_oracle_cpp32_2iso(const float* a_, int64_t na, const float* b_, int64_t nb, int auto_,
                          int bins, float dmax, int nthreads, double* hist) {
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  float dd_max = dmax * dmax;
  double ds = ds_floored(bins, dmax);
  std::memset(hist, 0, sizeof(double) * bins);
  const int nt = nthreads_or(nthreads);
  std::vector<double> h_all((size_t)nt * bins, 0.0);
  int err = 0;
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* H = &h_all[(size_t)tid * bins];
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < na; ++i) {
      const float x = a[i].x, y = a[i].y, z = a[i].z, w1 = a[i].w;
      for (int64_t j = auto_ ? i + 1 : 0; j < nb; ++j) {
        float dx = x - b[j].x;
        float dy = y - b[j].y;
        float dz = z - b[j].z;
        float disf = dx * dx + dy * dy + dz * dz;
        double dis = disf;
        if (dis < dd_max) {
          int bin = (int)(std::sqrt(dis) * ds);
          if (bin < 0 || bin >= bins) {
#pragma omp atomic write
            err = 1;
            continue;
          }
          float v = auto_ ? 2 * w1 * b[j].w : w1 * b[j].w;
          H[bin] += v;
        }
      }
    }
  }
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < bins; ++k) hist[k] += h_all[(size_t)t * bins + k];
  return err ? -3 : 0;
}

// Analytic RR of 2iso-BPC (cuda/pcf2isoBPC.cuh:199-202, kernel
// cuda/kernels/2isoBPC_k.cuh:559-575): RR_i = alpha * (3 i^2 + 3 i + 1) with
// dr = dmax/bn (FP32 division), beta = np^2 / L^3 (FP32 division of an FP32
// numerator by the FP32 cube), alpha = 8 dr^3 acos(0) beta / 3 in double.
// The reference's int overflows (np*np, 3i^2+3i+1) are replaced by 64-bit
// arithmetic.  Pinned by the shell-volume closed form (4pi/3) N^2 / L^3
// ((i+1)^3 - i^3) dr^3 and by the sum over bins.
static void analytic_rr(int bins, int64_t np, float dmax, float box, bool wrap32, double* rr) {
  float drf = dmax / (float)bins;
  double dr = drf;
  float np2 = wrap32 ? (float)wrap_i32((uint64_t)np * (uint64_t)np)
                     : (float)((double)np * (double)np);
  float L3 = box * box * box;
  float betaf = np2 / L3;
  double beta = betaf;
  double alpha = 8 * dr * dr * dr * std::acos(0.0) * beta / 3;
  for (int i = 0; i < bins; ++i) rr[i] = alpha * (double)shell_int(i, wrap32);
}

void pcf_oracle_analytic_rr(int bins, int64_t np, float dmax, float box, double* rr) {
  analytic_rr(bins, np, dmax, box, false, rr);
}

// The same with the reference's 32-bit int arithmetic (its defects): np*np
// (cuda/pcf2isoBPC.cuh:201) and 3i^2+3i+1 (cuda/kernels/2isoBPC_k.cuh:572)
// wrap modulo 2^32 into a signed int.  Pinned by hand-computed wrapped values.
void pcf_oracle_analytic_rr_refbugs(int bins, int64_t np, float dmax, float box, double* rr) {
  analytic_rr(bins, np, dmax, box, true, rr);
}


}  // extern "C"

// ---------------------------------------------------------------------------
// Triplets (3iso, 3ani)
// ---------------------------------------------------------------------------
namespace {

struct TripRule {
  int ani, bpc, auto_, bins;
  double R, ds, L, ds_th;
};

// Periodic per-axis term of the reference (cuda/kernels/3isoBPC_k.cuh:56-75):
// |delta| - L*f, where f says whether the axis is projected.  Squared by the
// caller.  For the 2-3 side f = |f2 - f3|.
inline double proj(double d, double L, int f) { return std::fabs(d) - L * (double)f; }

// One ordered triplet (i, j, k) seen through projection flags f2 (node 2) and
// f3 (node 3), each a 3-bit set of axes.  Node 1 is never projected
// (cuda/kernels/3isoBPC_k.cuh:56).  Non-periodic: f2 = f3 = 0 and the terms
// are the plain differences (cuda/kernels/3iso_k.cuh:57-70).  Returns the flat
// bin or -1.  Iso layout [b(d12)][b(d23)][b(d31)] (3iso_k.cuh:71-78); ani
// layout [b(d12)][b(d31)][b(d23)][t][p] (3ani_k.cuh:75-92).
inline int64_t triplet_bin(const Pt& p1, const Pt& p2, const Pt& p3, int f2, int f3,
                           const TripRule& r) {
  const double L = r.L;
  double t12[3], t23[3], t31[3];
  const float a1[3] = {p1.x, p1.y, p1.z}, a2[3] = {p2.x, p2.y, p2.z}, a3[3] = {p3.x, p3.y, p3.z};
  for (int a = 0; a < 3; ++a) {
    const double d12 = (double)a2[a] - (double)a1[a];
    const double d23 = (double)a3[a] - (double)a2[a];
    const double d31 = (double)a3[a] - (double)a1[a];
    if (r.bpc) {
      const int g2 = (f2 >> a) & 1, g3 = (f3 >> a) & 1;
      t12[a] = proj(d12, L, g2);
      t23[a] = proj(d23, L, g2 != g3);
      t31[a] = proj(d31, L, g3);
    } else {
      t12[a] = d12;
      t23[a] = d23;
      t31[a] = d31;
    }
  }
  const double s12 = t12[0] * t12[0] + t12[1] * t12[1] + t12[2] * t12[2];
  if (!(s12 < r.R && s12 > 0)) return -1;
  const double s23 = t23[0] * t23[0] + t23[1] * t23[1] + t23[2] * t23[2];
  const double s31 = t31[0] * t31[0] + t31[1] * t31[1] + t31[2] * t31[2];
  if (r.auto_) {
    // XXX: all three sides > 0 (3iso_k.cuh:67,70)
    if (!(s23 < r.R && s23 > 0 && s31 < r.R && s31 > 0)) return -1;
  } else {
    // XXY: only d12 > 0 (3iso_k.cuh:146,153,155)
    if (!(s23 < r.R && s31 < r.R)) return -1;
  }
  const double d12 = std::sqrt(s12), d23 = std::sqrt(s23), d31 = std::sqrt(s31);
  const int64_t bn = r.bins;
  const int a = (int)(d12 * r.ds);
  if (a > bn - 1) return -1;
  const int b23 = (int)(d23 * r.ds);
  if (b23 > bn - 1) return -1;
  const int b31 = (int)(d31 * r.ds);
  if (b31 > bn - 1) return -1;
  if (!r.ani) return ((int64_t)a * bn + b23) * bn + b31;
  // cos of the 1-2 and 1-3 directions to the line of sight z
  // (3ani_k.cuh:75-76; periodic: signed by the projected z term,
  // 3aniBPC_k.cuh:57,73).
  // XXY allows point 3 on top of point 1 (d31 = 0): the cosine is then 0/0.
  // The reference GPU's float-to-int conversion maps NaN to 0 (PTX cvt), so
  // such a triplet lands in p = 0; x86's conversion would not.
  auto to_bin = [](double x) { return std::isnan(x) ? 0 : (int)x; };
  const double cth12 = 1 + t12[2] / d12;
  const double cth31 = 1 + t31[2] / d31;
  const int t = to_bin(cth12 * r.ds_th);
  if (t > bn - 1) return -1;
  const int p = to_bin(cth31 * r.ds_th);
  if (p > bn - 1) return -1;
  return (((((int64_t)a * bn + b31) * bn + b23) * bn + t) * bn) + p;
}

// The projection of node 2 (or 3) relative to node 1 that the image k
// corresponds to, or -1 if k is the far image (the reference projects an axis
// only toward the close image: |delta| - L).  For delta == 0 the two images
// are one term, kept for k = -1.
inline int flags_of_image(const Pt& p1, const Pt& p2, const int k[3]) {
  int f = 0;
  const float a1[3] = {p1.x, p1.y, p1.z}, a2[3] = {p2.x, p2.y, p2.z};
  for (int a = 0; a < 3; ++a) {
    if (k[a] == 0) continue;
    const double d = (double)a2[a] - (double)a1[a];
    if ((k[a] > 0 && d >= 0.0) || (k[a] < 0 && d < 0.0)) return -1;
    f |= 1 << a;
  }
  return f;
}

TripRule make_trip_rule(int ani, int bpc, int auto_, int bins, float dmax, float box) {
  TripRule r;
  r.ani = ani;
  r.bpc = bpc;
  r.auto_ = auto_;
  r.bins = bins;
  float dd_max = dmax * dmax;
  r.R = (double)dd_max;
  r.ds = ds_floored(bins, dmax);  // 3iso_k.cuh:44, 3isoBPC_k.cuh:218, 3ani_k.cuh:45
  r.L = (double)box;
  float ds_th = (float)bins / 2;  // 3ani_k.cuh:42
  r.ds_th = (double)ds_th;
  return r;
}

}  // namespace

extern "C" {

// Ordered-triplet histogram in the FP64 canonical mode (GPU layout).
//   auto_ = 1: XXX over a (DDD/RRR).  auto_ = 0: XXY with points 1, 2 from a
//   and point 3 from b (DDR with a = D, b = R; DRR with a = R, b = D,
//   cuda/pcf3iso.cuh:111-114).
// Periodic: every image combination of nodes 2 and 3 relative to node 1 that
// lies within range (the reference's 64 projection combinations,
// cuda/kernels/3isoBPC_k.cuh:191-938).  Weighted sums use w1*w2*w3 formed in
// double per triplet (the reference multiplies cumulatively across the k loop,
// cuda/kernels/3iso_k.cuh:79 -- not reproduced).
// Pinned by: brute force over all 64 projection combinations
// (pcf_oracle_triplets_brute: the enumeration), the cpp32 comparison with the
// compiled reference CPU program, the reference's results/*.example,
// permutation symmetry, the 3iso marginal of 3ani, and -- independently of
// triplet_bin -- hand-built 3ani triplets of known angles (t, p, the cos = 1
// exclusion, the periodic |dz| - L*f sign rule), the 2 <-> 3 exchange and
// z-reflection identities of the angle bins, a minimum-image brute force of
// periodic 3iso/3ani on integer lattices (the |f2 - f3| rule of the 2-3 side)
// and translation invariance; a mutation test shows each planted mistake in
// these rules fails one of them (tests/test_oracle_triplet_pins.py).
int pcf_oracle_triplets(const float* a_, int64_t na, const float* b_, int64_t nb, int auto_,
                        int ani, int bpc, float box, int bins, float dmax, int nthreads,
                        uint64_t* counts, double* wsum) {
  if (bins <= 0 || !(dmax > 0)) return -1;
  // Periodic: dmax < box (2*dmax >= box included: a point may then count
  // under its regular and its projected image on an axis, as the reference's
  // 64 (f2, f3) combinations allow; the cells are then the whole box and the
  // neighbour images enumerate both).
  if (bpc && !(dmax < box)) return -2;
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  TripRule r = make_trip_rule(ani, bpc, auto_, bins, dmax, box);
  const int64_t bn = bins;
  const int64_t nout = ani ? bn * bn * bn * bn * bn : bn * bn * bn;
  std::memset(counts, 0, sizeof(uint64_t) * nout);
  std::memset(wsum, 0, sizeof(double) * nout);
  if (na == 0 || nb == 0) return 0;
  Grid g;
  grid_geometry(a, na, auto_ ? nullptr : b, nb, (double)dmax * 1.0000001 + 1e-6, bpc,
                (double)box, &g);
  Grid gb = g;
  g.build(a, na);
  gb.build(b, nb);
  const int nt = nthreads_or(nthreads);
  const int64_t ncell = (int64_t)g.P[0] * g.P[1] * g.P[2];
  // Thread-private histograms are large for 3ani; use atomics into shared ones.
  std::vector<uint64_t> C(nout, 0);
  std::vector<double> W(nout, 0.0);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t ca = 0; ca < ncell; ++ca) {
    int cx = (int)(ca / ((int64_t)g.P[1] * g.P[2]));
    int cy = (int)((ca / g.P[2]) % g.P[1]);
    int cz = (int)(ca % g.P[2]);
    std::vector<std::pair<int64_t, std::array<int, 3>>> nbrs;
    for_neighbours(g, bpc, cx, cy, cz, [&](int64_t cb, int kx, int ky, int kz) {
      nbrs.push_back({cb, {kx, ky, kz}});
    });
    for (int64_t i = g.start[ca]; i < g.start[ca + 1]; ++i) {
      const Pt& p1 = g.pts[i];
      for (auto& n2 : nbrs) {
        for (int64_t j = g.start[n2.first]; j < g.start[n2.first + 1]; ++j) {
          const Pt& p2 = g.pts[j];
          const int f2 = flags_of_image(p1, p2, n2.second.data());
          if (f2 < 0) continue;
          for (auto& n3 : nbrs) {
            for (int64_t k = gb.start[n3.first]; k < gb.start[n3.first + 1]; ++k) {
              const Pt& p3 = gb.pts[k];
              const int f3 = flags_of_image(p1, p3, n3.second.data());
              if (f3 < 0) continue;
              const int64_t bin = triplet_bin(p1, p2, p3, f2, f3, r);
              if (bin < 0) continue;
              const double w = (double)p1.w * (double)p2.w * (double)p3.w;
#pragma omp atomic
              C[bin] += 1;
#pragma omp atomic
              W[bin] += w;
            }
          }
        }
      }
    }
  }
  for (int64_t k = 0; k < nout; ++k) {
    counts[k] = C[k];
    wsum[k] = W[k];
  }
  return 0;
}

// Literal restatement for tiny inputs: every ordered triplet and, periodic,
// every one of the 64 (f2, f3) projection combinations of
// cuda/kernels/3isoBPC_k.cuh:191-938 with the per-axis terms of count123
// (3isoBPC_k.cuh:56-75).  O(n^3 * 64).  Pins pcf_oracle_triplets.
int pcf_oracle_triplets_brute(const float* a_, int64_t na, const float* b_, int64_t nb,
                              int auto_, int ani, int bpc, float box, int bins, float dmax,
                              int nthreads, uint64_t* counts, double* wsum) {
  (void)nthreads;
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  TripRule r = make_trip_rule(ani, bpc, auto_, bins, dmax, box);
  const int64_t bn = bins;
  const int64_t nout = ani ? bn * bn * bn * bn * bn : bn * bn * bn;
  std::memset(counts, 0, sizeof(uint64_t) * nout);
  std::memset(wsum, 0, sizeof(double) * nout);
  const int ncomb = bpc ? 8 : 1;
  for (int64_t i = 0; i < na; ++i)
    for (int64_t j = 0; j < na; ++j)
      for (int64_t k = 0; k < nb; ++k)
        for (int f2 = 0; f2 < ncomb; ++f2)
          for (int f3 = 0; f3 < ncomb; ++f3) {
            const int64_t bin = triplet_bin(a[i], a[j], b[k], f2, f3, r);
            if (bin < 0) continue;
            counts[bin] += 1;
            wsum[bin] += (double)a[i].w * (double)a[j].w * (double)b[k].w;
          }
  return 0;
}

// cpp32 mode of cpp/3PCFisotropic/normal (3PCFiso.h): unique triangles with
// FP32 differences and squares (no contraction) converted to double, only
// d < dd_max tested (no d > 0 test, 3PCFiso.h:534-550), deposit at
// [(int)(d12 ds)][(int)(d13 ds)][(int)(d23 ds)]; XXX deposits w1*w2*w3
// (FP32 product), XXY deposits w1*w2*w3/3 (FP32, 3PCFiso.h:949,1011) for each
// unordered X pair and Y point; then symmetrize (3PCFiso.h:705-727): the sum
// of the six permutations, rounded to FP32, stored in all six.
// Output: the program's histogram (double), bins^3 values.
int pcf_oracle_cpp32_3iso(const float* a_, int64_t na, const float* b_, int64_t nb, int auto_,
                          int bins, float dmax, int nthreads, double* hist) {
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  const float dd_max = dmax * dmax;
  const double ds = ds_floored(bins, dmax);
  const int64_t bn = bins;
  std::vector<double> U(bn * bn * bn, 0.0);
  Grid g;
  grid_geometry(a, na, auto_ ? nullptr : b, nb, (double)dmax * 1.0000001 + 1e-6, 0, 0.0, &g);
  Grid gb = g;
  g.build(a, na);
  gb.build(b, nb);
  const int nt = nthreads_or(nthreads);
  const int64_t ncell = (int64_t)g.P[0] * g.P[1] * g.P[2];
  auto fdist = [](const Pt& p, const Pt& q) {
    float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    float d = dx * dx + dy * dy + dz * dz;
    return (double)d;
  };
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t ca = 0; ca < ncell; ++ca) {
    int cx = (int)(ca / ((int64_t)g.P[1] * g.P[2]));
    int cy = (int)((ca / g.P[2]) % g.P[1]);
    int cz = (int)(ca % g.P[2]);
    std::vector<int64_t> nb_cells;
    for_neighbours(g, 0, cx, cy, cz, [&](int64_t cb, int, int, int) { nb_cells.push_back(cb); });
    std::vector<double> Uloc;
    std::vector<std::pair<int64_t, double>> dep;
    for (int64_t ii = g.start[ca]; ii < g.start[ca + 1]; ++ii) {
      // point i is the one with the smallest sorted index of its triangle
      const Pt& p1 = g.pts[ii];
      for (int64_t cj : nb_cells)
        for (int64_t jj = g.start[cj]; jj < g.start[cj + 1]; ++jj) {
          if (auto_ && jj <= ii) continue;
          const Pt& p2 = g.pts[jj];
          double d12 = fdist(p1, p2);
          if (!(d12 < dd_max)) continue;
          if (!auto_ && jj <= ii) continue;  // unordered X pair
          for (int64_t ck : nb_cells)
            for (int64_t kk = gb.start[ck]; kk < gb.start[ck + 1]; ++kk) {
              if (auto_ && kk <= jj) continue;
              const Pt& p3 = gb.pts[kk];
              double d13 = fdist(p1, p3);
              if (!(d13 < dd_max)) continue;
              double d23 = fdist(p2, p3);
              if (!(d23 < dd_max)) continue;
              int x = (int)(std::sqrt(d12) * ds), y = (int)(std::sqrt(d13) * ds),
                  z = (int)(std::sqrt(d23) * ds);
              if (x >= bins || y >= bins || z >= bins) { err = 1; continue; }
              float w = p1.w * p2.w * p3.w;
              if (!auto_) w = w / 3;
              dep.push_back({(x * bn + y) * bn + z, (double)w});
            }
        }
    }
    if (!dep.empty()) {
#pragma omp critical
      for (auto& d : dep) U[d.first] += d.second;
    }
  }
  // symmetrize in FP32 (3PCFiso.h:705-727)
  auto at = [&](int64_t i, int64_t j, int64_t k) -> double& { return U[(i * bn + j) * bn + k]; };
  for (int64_t i = 0; i < bn; i++)
    for (int64_t j = i; j < bn; j++)
      for (int64_t k = j; k < bn; k++) {
        float elem = at(i, j, k) + at(k, i, j) + at(j, k, i) + at(j, i, k) + at(k, j, i) + at(i, k, j);
        at(i, j, k) = elem;
        at(k, i, j) = elem;
        at(j, k, i) = elem;
        at(j, i, k) = elem;
        at(k, j, i) = elem;
        at(i, k, j) = elem;
      }
  std::memcpy(hist, U.data(), sizeof(double) * bn * bn * bn);
  return err ? -3 : 0;
}

}  // extern "C"

extern "C" {

// ---------------------------------------------------------------------------
// Analytic periodic 3iso (-bpc -a): cuda/pcf3isoBPC.cuh:177-375 with the
// kernels DD2iso_forBPC, RR2iso_BPC_analytic, make_ff_av, make_ff_av_ref and
// make_histo_analitic (cuda/kernels/3isoBPC_k.cuh:1694-2119,
// cuda/kernels/2isoBPC_k.cuh:559-575).
//   fine DD: periodic ordered pairs binned twice, at nb1 = ptt*bn and
//     nb2 = ptt*bn_ref*bn bins with the FP32 scale (float)nb/dmax
//     (3isoBPC_k.cuh:1724); FP64 distances here (the reference uses FP32);
//     weights w_i*w_j summed in double.
//   fine RR: alpha*(3i^2+3i+1), dr = dmax/nb (FP32), beta = np^2/L^3 (FP32),
//     64-bit (the reference's int overflows at i >= 26755, defect fixed).
//   ff_av[i] = sum_j (ri + rj)(DD/RR - 1)/ptt, ff_av_ref likewise
//     (3isoBPC_k.cuh:1984-2019), summed in index order.
//   RRR, DDR per bin cube: closed form when all 8 corners satisfy the triangle
//     inequality, else a bn_ref^3 sub-cube sum (3isoBPC_k.cuh:2021-2119),
//     summed in (u, v, w) order.  DRR := DDR (cuda/pcf3isoBPC.cuh:340).
// Pinned by: the closed form of full cubes; the reference CPU program's
// ff_av (cpp/3PCFisotropic/BPC, FP32 distances: agreement to the fine-bin
// edge-flip noise); GPU parity at reduced ptt/bn_ref.  The sub-cube refinement
// itself is formula-pinned (no runnable reference output exists for it).
static int analytic_3iso(const float* a_, int64_t na, int bins, float dmax, float box,
                         int64_t np, int ptt, int bn_ref, int nthreads, double* ddr, double* rrr,
                         double* ff_av_out, double* ff_av_ref_out, bool wrap32) {
  if (bins <= 0 || ptt <= 0 || bn_ref <= 0 || !(dmax < box)) return -1;
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const int64_t nb1 = (int64_t)ptt * bins, nb2 = (int64_t)ptt * bn_ref * bins;
  const float dsf1 = (float)nb1 / dmax, dsf2 = (float)nb2 / dmax;
  const double ds1 = dsf1, ds2 = dsf2;
  PairRule r = make_pair_rule(0, 1, 1, bins, dmax, box);
  std::vector<double> DD1(nb1, 0.0), DD2(nb2, 0.0);
  Grid g;
  grid_geometry(a, na, nullptr, 0, (double)dmax * 1.0000001 + 1e-6, 1, (double)box, &g);
  g.build(a, na);
  const int nt = nthreads_or(nthreads);
  const int64_t ncell = (int64_t)g.P[0] * g.P[1] * g.P[2];
#pragma omp parallel num_threads(nt)
  {
    std::vector<double> L1(nb1, 0.0), L2(nb2, 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int64_t ca = 0; ca < ncell; ++ca) {
      int cx = (int)(ca / ((int64_t)g.P[1] * g.P[2]));
      int cy = (int)((ca / g.P[2]) % g.P[1]);
      int cz = (int)(ca % g.P[2]);
      for_neighbours(g, 1, cx, cy, cz, [&](int64_t cb, int kx, int ky, int kz) {
        for (int64_t i = g.start[ca]; i < g.start[ca + 1]; ++i)
          for (int64_t j = g.start[cb]; j < g.start[cb + 1]; ++j) {
            double tx, ty, tz;
            if (!axis_sq(g.pts[i].x, g.pts[j].x, kx, r.L, &tx)) continue;
            if (!axis_sq(g.pts[i].y, g.pts[j].y, ky, r.L, &ty)) continue;
            if (!axis_sq(g.pts[i].z, g.pts[j].z, kz, r.L, &tz)) continue;
            const double d2 = tx + ty + tz;
            if (!(d2 < r.R && d2 > 0)) continue;
            const double d = std::sqrt(d2);
            const double w = (double)g.pts[i].w * (double)g.pts[j].w;
            const int64_t b1 = (int64_t)(d * ds1), b2 = (int64_t)(d * ds2);
            if (b1 <= nb1 - 1) L1[b1] += w;
            if (b2 <= nb2 - 1) L2[b2] += w;
          }
      });
    }
#pragma omp critical
    {
      for (int64_t k = 0; k < nb1; ++k) DD1[k] += L1[k];
      for (int64_t k = 0; k < nb2; ++k) DD2[k] += L2[k];
    }
  }
  // Fine analytic RR (cuda/pcf3isoBPC.cuh:233-238); wrap32: the
  // reference's 32-bit np*np and 3i^2+3i+1.
  const float np2 = wrap32 ? (float)wrap_i32((uint64_t)np * (uint64_t)np)
                           : (float)((double)np * (double)np);
  const float L3 = box * box * box;
  const double beta = (double)(float)(np2 / L3);
  auto rr_of = [&](int64_t nb, int64_t i) {
    const double dr = (double)(float)(dmax / (float)nb);
    const double alpha = 8 * dr * dr * dr * std::acos(0.0) * beta / 3;
    return alpha * (double)shell_int(i, wrap32);
  };
  std::vector<double> ff(bins, 0.0), ffr((size_t)bins * bn_ref, 0.0);
  for (int i = 0; i < bins; ++i) {
    const double ri = (double)(float)(i * dmax) / (double)bins;
    for (int j = 0; j < ptt; ++j) {
      const double rj = (j + 0.5) * (double)dmax / (double)nb1;
      const int64_t q = (int64_t)i * ptt + j;
      ff[i] += (ri + rj) * (DD1[q] / rr_of(nb1, q) - 1) / (double)ptt;
    }
  }
  {
    const double dr = (double)dmax / (double)bins;
    const double dr_ref = dr / bn_ref;
    const double dr_ptt_ref = (double)dmax / (double)nb2;
    for (int i = 0; i < bins; ++i)
      for (int j = 0; j < bn_ref; ++j)
        for (int k = 0; k < ptt; ++k) {
          const double ri = i * dr, rj = j * dr_ref, rk = (k + 0.5) * dr_ptt_ref;
          const int64_t q = ((int64_t)i * bn_ref) * ptt + (int64_t)j * ptt + k;
          ffr[(size_t)i * bn_ref + j] += (ri + rj + rk) * ((DD2[q] / rr_of(nb2, q)) - 1) / (double)ptt;
        }
  }
  // Cubes (3isoBPC_k.cuh:2021-2119; constants cuda/pcf3isoBPC.cuh:240-245).
  const double dr = (double)dmax / (double)bins, dr2 = dr * 0.5;
  const double V = (double)(box * box);
  const double beta3 = (double)np / V;
  double gama = 8 * (4 * std::acos(0.0) * std::acos(0.0)) * ((double)np * beta3 * beta3);
  gama /= V;
  const double dr_ref = dr / bn_ref, dr_ref2 = dr_ref * 0.5;
  const double alpha = gama * dr * dr * dr, alpha_ref = gama * dr_ref * dr_ref * dr_ref;
  if (ff_av_out) std::memcpy(ff_av_out, ff.data(), sizeof(double) * bins);
  if (ff_av_ref_out) std::memcpy(ff_av_ref_out, ffr.data(), sizeof(double) * bins * bn_ref);
  if (!ddr || !rrr) return 0;  // ff_av / ff_av_ref only
  const int64_t b3 = (int64_t)bins * bins * bins;
  std::memset(ddr, 0, sizeof(double) * b3);
  std::memset(rrr, 0, sizeof(double) * b3);
  auto corners = [](double r1, double r2, double r3, double h) {
    int v = 0;
    for (int aa = 0; aa < 2; ++aa)
      for (int bb = 0; bb < 2; ++bb)
        for (int cc = 0; cc < 2; ++cc) {
          const double s1 = r1 + (double)aa * h, s2 = r2 + (double)bb * h, s3 = r3 + (double)cc * h;
          if (s1 + s2 >= s3 && s1 + s3 >= s2 && s2 + s3 >= s1) ++v;
        }
    return v;
  };
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t t = 0; t < b3; ++t) {
    const int i = (int)(t / ((int64_t)bins * bins)), j = (int)((t / bins) % bins), k = (int)(t % bins);
    const double ri = (double)i * dr, rj = (double)j * dr, rk = (double)k * dr;
    const int vin = corners(ri, rj, rk, dr);
    if (vin == 8) {
      const double sv = alpha * (ri + dr2) * (rj + dr2) * (rk + dr2);
      double f = 1;
      f += ff[i] / (3 * (ri + dr2));
      f += ff[j] / (3 * (rj + dr2));
      f += ff[k] / (3 * (rk + dr2));
      f *= sv;
      rrr[t] = sv;
      ddr[t] = f;
    } else if (vin > 0) {
      double S = 0.0, F = 0.0;
      bool con = false;
      for (int u = 0; u < bn_ref; ++u) {
        const double ru = ri + (double)u * dr_ref;
        for (int v = 0; v < bn_ref; ++v) {
          const double rv = rj + (double)v * dr_ref;
          for (int w = 0; w < bn_ref; ++w) {
            const double rw = rk + (double)w * dr_ref;
            if (corners(ru, rv, rw, dr_ref) == 8) {
              const double c = (ru + dr_ref2) * (rv + dr_ref2) * (rw + dr_ref2);
              S += c;
              double f = 1;
              f += ffr[(size_t)i * bn_ref + u] / (3 * (ru + dr_ref2));
              f += ffr[(size_t)j * bn_ref + v] / (3 * (rv + dr_ref2));
              f += ffr[(size_t)k * bn_ref + w] / (3 * (rw + dr_ref2));
              f *= c;
              F += f;
              con = true;
            }
          }
        }
      }
      if (con) {
        rrr[t] = S * alpha_ref;
        ddr[t] = F * alpha_ref;
      }
    }
  }
  return 0;
}


int pcf_oracle_analytic_3iso(const float* a_, int64_t na, int bins, float dmax, float box,
                             int64_t np, int ptt, int bn_ref, int nthreads, double* ddr,
                             double* rrr, double* ff_av_out, double* ff_av_ref_out) {
  return analytic_3iso(a_, na, bins, dmax, box, np, ptt, bn_ref, nthreads, ddr, rrr, ff_av_out,
                       ff_av_ref_out, false);
}

// The same with the reference's 32-bit np*np (cuda/pcf3isoBPC.cuh:236) and
// 3i^2+3i+1 of its fine RR (cuda/kernels/2isoBPC_k.cuh:572).
int pcf_oracle_analytic_3iso_refbugs(const float* a_, int64_t na, int bins, float dmax,
                                     float box, int64_t np, int ptt, int bn_ref, int nthreads,
                                     double* ddr, double* rrr, double* ff_av_out,
                                     double* ff_av_ref_out) {
  return analytic_3iso(a_, na, bins, dmax, box, np, ptt, bn_ref, nthreads, ddr, rrr, ff_av_out,
                       ff_av_ref_out, true);
}

// ---------------------------------------------------------------------------
// Reference defects reproduced on request (--ref-bugs).  Both depend on the
// reference's node grid: partitions^3 nodes of side size_box / partitions
// (FP32, cuda/main.cu:216), a point in node ((int)(x / size_node), ...)
// (cuda/create_grid.cuh:133-135; clamped to the last node where FP32 rounding
// of a coordinate just below the box side would reach past it), the nodes in
// row-major (x, y, z) order and a node's points in file order (add() appends,
// create_grid.cuh:81-102; cuda/main.cu:276-300 copies node by node).
// ---------------------------------------------------------------------------
}  // extern "C"
namespace {
int64_t ref_node_of(const Pt& p, float size_node, int P) {
  auto ax = [&](float v) {
    int c = (int)(v / size_node);
    return c < 0 ? 0 : (c >= P ? P - 1 : c);
  };
  return ((int64_t)ax(p.x) * P + ax(p.y)) * P + ax(p.z);
}
// Indices of the catalog in the reference's ordered array (node order, file
// order inside a node).
std::vector<int64_t> ref_order(const Pt* p, int64_t n, float box, int partitions) {
  const float size_node = box / (float)partitions;
  std::vector<int64_t> node(n), ord(n);
  for (int64_t i = 0; i < n; ++i) node[i] = ref_node_of(p[i], size_node, partitions);
  for (int64_t i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return node[a] < node[b]; });
  return ord;
}
}  // namespace
extern "C" {

// Triplet histogram with the reference GPU kernels' cumulative weight: in
// count123 and the XXX/XXY kernels (cuda/kernels/3iso_k.cuh:61,79,168,
// 3isoBPC_k.cuh:60,88,179, 3ani_k.cuh:94,197, 3aniBPC_k.cuh:98) v = w1*w2 (an
// FP32 product) is set for each (point 1, point 2) and multiplied by w3 of
// every third point deposited so far in the same node 3 under the same
// projection pair (f2, f3), in the node's point order, before each deposit;
// the deposit is v, not w1*w2*w3.  Counts are those of pcf_oracle_triplets.
// Literal form: every ordered (i, j), every (f2, f3) and every node of the
// third-point catalog's reference grid walked in order.  O(na^2 nb 64); small
// inputs.  Pinned by unit weights (equal to pcf_oracle_triplets) and
// hand-worked chains (tests/test_oracle_refbugs.py).
int pcf_oracle_triplets_refbugs(const float* a_, int64_t na, const float* b_, int64_t nb,
                                int auto_, int ani, int bpc, float box, int partitions,
                                int bins, float dmax, uint64_t* counts, double* wsum) {
  if (bins <= 0 || !(dmax > 0) || !(box > 0) || partitions <= 0) return -1;
  if (bpc && !(dmax < box)) return -2;
  const Pt* a = reinterpret_cast<const Pt*>(a_);
  const Pt* b = auto_ ? a : reinterpret_cast<const Pt*>(b_);
  if (auto_) nb = na;
  TripRule r = make_trip_rule(ani, bpc, auto_, bins, dmax, box);
  const int64_t bn = bins;
  const int64_t nout = ani ? bn * bn * bn * bn * bn : bn * bn * bn;
  std::memset(counts, 0, sizeof(uint64_t) * nout);
  std::memset(wsum, 0, sizeof(double) * nout);
  const std::vector<int64_t> ord = ref_order(b, nb, box, partitions);
  const float size_node = box / (float)partitions;
  // Node boundaries in the ordered array.
  std::vector<int64_t> bounds;
  for (int64_t q = 0; q < nb; ++q)
    if (q == 0 || ref_node_of(b[ord[q]], size_node, partitions) !=
                      ref_node_of(b[ord[q - 1]], size_node, partitions))
      bounds.push_back(q);
  bounds.push_back(nb);
  const int ncomb = bpc ? 8 : 1;
  for (int64_t i = 0; i < na; ++i)
    for (int64_t j = 0; j < na; ++j)
      for (int f2 = 0; f2 < ncomb; ++f2) {
        const float w12 = a[i].w * a[j].w;  // FP32 product, as the reference forms it
        for (size_t s = 0; s + 1 < bounds.size(); ++s)
          for (int f3 = 0; f3 < ncomb; ++f3) {
            double v = (double)w12;
            for (int64_t q = bounds[s]; q < bounds[s + 1]; ++q) {
              const Pt& p3 = b[ord[q]];
              const int64_t bin = triplet_bin(a[i], a[j], p3, f2, f3, r);
              if (bin < 0) continue;
              v *= (double)p3.w;
              counts[bin] += 1;
              wsum[bin] += v;
            }
          }
      }
  return 0;
}

// Weights the reference's XY2ani reads for the random points
// (cuda/kernels/2ani_k.cuh:128: v = w1 * elementsD[j].w, j an index of the
// random catalog's ordered array): out[j] = weight of the data point at the
// random point j's position in the data catalog's ordered array, or NaN where
// that position lies past the data catalog (the reference reads out of
// bounds there).
void pcf_oracle_xy2ani_alias_weights(const float* d_, int64_t nd, const float* r_, int64_t nr,
                                     float box, int partitions, float* out) {
  const Pt* d = reinterpret_cast<const Pt*>(d_);
  const Pt* rp = reinterpret_cast<const Pt*>(r_);
  const std::vector<int64_t> od = ref_order(d, nd, box, partitions);
  const std::vector<int64_t> orr = ref_order(rp, nr, box, partitions);
  for (int64_t q = 0; q < nr; ++q)
    out[orr[q]] = q < nd ? d[od[q]].w : std::numeric_limits<float>::quiet_NaN();
}

}  // extern "C"
