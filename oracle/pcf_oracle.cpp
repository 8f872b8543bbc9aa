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
#include <cmath>
#include <cstdint>
#include <cstring>
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
void pcf_oracle_analytic_rr(int bins, int64_t np, float dmax, float box, double* rr) {
  float drf = dmax / (float)bins;
  double dr = drf;
  float np2 = (float)((double)np * (double)np);
  float L3 = box * box * box;
  float betaf = np2 / L3;
  double beta = betaf;
  double alpha = 8 * dr * dr * dr * std::acos(0.0) * beta / 3;
  for (int i = 0; i < bins; ++i) {
    int64_t k = 3 * (int64_t)i * i + 3 * (int64_t)i + 1;
    rr[i] = alpha * (double)k;
  }
}

}  // extern "C"
