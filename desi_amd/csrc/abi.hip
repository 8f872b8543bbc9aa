// abi.hip -- host side of libpcf_hip.so: contexts, resident catalogs, job
// set-up (geometry, exact bin thresholds, FP32 bin table) and the C ABI
// declared in include/pcf_abi.h.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>
#include <tuple>
#include <map>

#include "../../include/pcf_abi.h"
#include "../../include/pcf_selftest.h"
#include "pcf_internal.h"
#include "pcf_ctx.h"

using namespace pcf;

static thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// Exact FP64 bin thresholds and the FP32 bulk table
// ---------------------------------------------------------------------------
namespace pcf {

double ds_floored(int bins, float dmax) {
  // ds = floor(bn/dmax*1e6)/1e6 in double (cuda/kernels/2iso_k.cuh:37).
  return std::floor(((double)bins / (double)dmax) * 1000000.0) / 1000000.0;
}

static inline uint64_t dbits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}
static inline double dfrom(uint64_t u) {
  double x;
  std::memcpy(&x, &u, 8);
  return x;
}

// T[k] = smallest double d2 >= 0 with (int)(sqrt(d2)*ds) >= k, by bisection
// over the ordered bit patterns of non-negative doubles (host sqrt is IEEE
// correctly rounded).  The map d2 -> (int)(sqrt(d2)*ds) is monotone, so
// bin(d2) = #{k >= 1 : T[k] <= d2} reproduces it exactly without a device sqrt.
void bin_thresholds(int bins, double ds, double* T) {
  T[0] = 0.0;
  for (int k = 1; k <= bins; ++k) {
    auto f = [&](double x) { return (int)(std::sqrt(x) * ds) >= k; };
    uint64_t lo = 0;  // f(0) false for k >= 1
    double h = ((double)k / ds) * ((double)k / ds) * 4.0 + 1.0;
    while (!f(h)) h *= 2;
    uint64_t hi = dbits(h);
    while (hi - lo > 1) {
      uint64_t mid = lo + (hi - lo) / 2;
      if (f(dfrom(mid))) hi = mid;
      else lo = mid;
    }
    T[k] = dfrom(hi);
  }
}

void make_rules(int type, int bpc, int auto_, int bins, float dmax, float box, BinRules* r) {
  std::memset(r, 0, sizeof(*r));
  const bool ani = (type == PCF_2ANI);
  float dd_max = dmax * dmax;  // FP32 product as in the kernels
  r->R = (double)dd_max;
  r->L = (double)box;
  r->bins = bins;
  r->ani = ani ? 1 : 0;
  if (!ani) {
    r->ds = bpc ? (double)bins / (double)dmax : ds_floored(bins, dmax);
    r->le = 0;
    r->zero = auto_ ? 1 : 0;
  } else {
    r->ds = ds_floored(bins, dmax);
    r->le = bpc ? 1 : 0;
    r->zero = (!bpc || auto_) ? 2 : 0;
  }
  // Thresholds serve the mixed path's exact re-evaluation (bins <= kMaxBinsIso);
  // the plain FP64 kernels bin with sqrt directly.
  if (bins <= kMaxBinsIso) bin_thresholds(bins, r->ds, r->T);
}

// Exact class of a double d2 for one component: bin in [0,bins) or -1 (out).
static int host_class(double d2, const BinRules& r) {
  if (r.le ? !(d2 <= r.R) : !(d2 < r.R)) return -1;
  if (d2 >= r.T[r.bins]) return -1;
  int b = 0;
  for (int k = 1; k < r.bins; ++k)
    if (r.T[k] <= d2) b = k;
  return b;
}

// Bound on |d2_fp32 - d2_exact| for the bulk arithmetic of the pair kernels
// (per axis one FP32 subtraction after an optional image shift xi - k*L, then
// dx*dx, fma, fma).  u = 2^-24.  A = max|coordinate| + L when periodic images
// are applied, else 0 (a plain FP32 difference has relative error <= u).
//   E(d2) = 5u d2 + 3.5u A sqrt(d2) + 6u^2 A^2      (doubled for margin)
static double err_bound(double d2, double A) {
  const double u = std::ldexp(1.0, -24);
  return 2.0 * (5 * u * d2 + 3.5 * u * A * std::sqrt(d2) + 6 * u * u * A * A) + 1e-300;
}

// Bound on |d2_fp32 - d2_exact| for the Gram form of the 2iso kernel (matrix
// cores), for tiles whose coordinates relative to the tile's box centre are at
// most rho in length.  With d the true separation, |x_i'| <= rho and
// |x_j'| <= rho + d:
//   |x_i'|^2 and |x_j'|^2 (mul, fma, fma):          3u rho^2 + 3u (rho + d)^2
//   the four chained fmas of v_mfma_f32_16x16x4f32: 4u (2 rho + d)^2
//   rounding of the shifted coordinates (image shift, origin), per axis at
//   most u (A + 2 rho + d), moving d2 by 2 sqrt(3) d u (A + 2 rho + d)
// plus second-order terms, doubled for margin.
static double err_bound_gram(double d2, double A, double rho) {
  const double u = std::ldexp(1.0, -24);
  const double d = std::sqrt(d2);
  const double e = 3 * rho * rho + 3 * (rho + d) * (rho + d) + 4 * (2 * rho + d) * (2 * rho + d) +
                   2 * std::sqrt(3.0) * d * (A + 2 * rho + d);
  const double w = A + 2 * rho + d;
  return 2.0 * (u * e + 12 * u * u * w * w) + 1e-300;
}

// The table scale inv_h is a power of four, so the bulk path can pre-scale
// coordinates by sigma = sqrt(inv_h) (a power of two, exact in FP32) and
// convert d2 straight to a table index.  *nt_out <= nt_max entries cover
// [0, Vmax) where everything beyond Vmax is certainly out of range.
static int build_table_uncached(const BinRules& r, double A, double rho, int nt_max, bool flip,
                                std::vector<uint8_t>* table, float* inv_h_out, int* nt_out,
                                float* top_out);

// The table depends only on (bins, ani, R, ds, le, zero, A, nt_max, rho,
// flip) and takes ~1 ms of host time to build: built tables are kept
// (process-wide, bounded).
int build_table(const BinRules& r, double A, int nt_max, std::vector<uint8_t>* table,
                float* inv_h_out, int* nt_out, double rho, bool flip, float* top_out) {
  struct Key {
    int bins, ani, le, zero, nt_max, flip;
    double R, ds, A, rho;
    bool operator<(const Key& o) const {
      return std::tie(bins, ani, le, zero, nt_max, flip, R, ds, A, rho) <
             std::tie(o.bins, o.ani, o.le, o.zero, o.nt_max, o.flip, o.R, o.ds, o.A, o.rho);
    }
  };
  struct Val {
    std::vector<uint8_t> table;
    float inv_h;
    int nt;
    float top;
    int rc;
  };
  static std::mutex mu;
  static std::map<Key, Val> cache;
  const Key key{r.bins, r.ani, r.le, r.zero, nt_max, flip ? 1 : 0, r.R, r.ds, A, rho};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *table = it->second.table;
      *inv_h_out = it->second.inv_h;
      *nt_out = it->second.nt;
      if (top_out) *top_out = it->second.top;
      return it->second.rc;
    }
  }
  Val v;
  v.inv_h = 1.f;
  v.nt = 2;
  v.top = 0.f;
  v.rc = build_table_uncached(r, A, rho, nt_max, flip, &v.table, &v.inv_h, &v.nt, &v.top);
  *table = v.table;
  *inv_h_out = v.inv_h;
  *nt_out = v.nt;
  if (top_out) *top_out = v.top;
  const int rc = v.rc;
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 64) cache.erase(cache.begin());
  cache.emplace(key, std::move(v));
  return rc;
}

static int build_table_uncached(const BinRules& r, double A, double rho, int nt_max, bool flip,
                                std::vector<uint8_t>* table, float* inv_h_out, int* nt_out,
                                float* top_out) {
  const int bins = r.bins;
  const bool ani = r.ani != 0;
  const double u = std::ldexp(1.0, -24);
  // Flipped tables carry two more cells (see below).
  const int budget = flip ? nt_max - 2 : nt_max;
  // Error margin of the paths sharing the table: the direct FP32 difference
  // form, and with rho >= 0 also the Gram form.  Flipped, both also round
  // Vtop - d2 (Vtop: the table top in unscaled units, set below): the direct
  // form's final subtraction adds u (Vtop + d2); in the Gram form every
  // partial sum of the product carries Vtop - |x_i|^2, so its four roundings
  // and that of the accumulator's start add at most 5 u Vtop (both doubled).
  double vtop = 0.0;
  auto err = [&](double d2) {
    double e = err_bound(d2, A);
    if (flip) e += 2.0 * u * (vtop + d2);
    if (rho >= 0) e = std::max(e, err_bound_gram(d2, A, rho) + (flip ? 10.0 * u * vtop : 0.0));
    return e;
  };
  double U = std::min(r.R, r.T[bins]);
  double vmax = U * (1.0 + 1e-3) + 4.0 * err(2 * U + 1) + 1e-30;
  // The u * Vtop terms are far below a cell; a 1e-3 margin on vmax covers them.
  if (flip) vmax *= 1.0 + 1e-3;
  int e2 = (int)std::floor(std::log2((double)(budget - 2) / vmax) / 2.0);
  if (e2 > 60) e2 = 60;
  if (e2 < -60) e2 = -60;
  double ih = std::ldexp(1.0, 2 * e2);
  while (vmax * ih > (double)(budget - 2)) ih *= 0.25;
  const int nt = (int)std::ceil(vmax * ih) + 2;
  const float inv_h = (float)ih;
  if (bins > (ani ? 200 : 253)) return -1;
  const uint8_t out_code = ani ? kAniOut8 : (uint8_t)bins;
  const uint8_t amb_code = ani ? kAniAmb8 : (uint8_t)(bins + 1);
  if (flip) {
    // Flipped table: the kernels index it by cvt_rz(Tf - d2) (table units,
    // Tf = nt), so everything beyond the top saturates to cell 0 and no clamp
    // is needed.  Cell t holds the FP32 values D = Tf - d2 with
    // cvt_rz(D) == t: t = 0 is D < 1 (d2 beyond the top, and the NaN-free
    // negatives), t >= 1 is D in [t, t + 1); d2 >= 0 bounds D by Tf plus the
    // rounding, hence the two cells above Tf.
    const double Tf = (double)nt;
    vtop = Tf / ih;
    const int ntf = nt + 2;
    table->assign(ntf, 0);
    for (int t = 0; t < ntf; ++t) {
      const double D_lo = t == 0 ? -std::numeric_limits<double>::infinity()
                                 : (double)t * (1.0 - 4 * u);
      const double D_hi = (double)(t + 1) * (1.0 + 4 * u);
      const double v_lo = (Tf - D_hi) / ih;  // smallest d2 of the cell (may be < 0)
      const double v_hi = (Tf - D_lo) / ih;  // largest (+inf for cell 0)
      const double d_lo = v_lo - 2.0 * err(2 * std::fabs(v_lo) + 1);
      const double d_hi =
          std::isinf(v_hi) ? v_hi : v_hi + 2.0 * err(2 * std::max(v_hi, 0.0) + 1);
      uint8_t code;
      if (r.zero != 0 && d_lo <= 0.0) {
        code = amb_code;
      } else {
        const int c_lo = host_class(std::max(d_lo, 0.0), r);
        const int c_hi = std::isinf(d_hi) ? -1 : host_class(std::max(d_hi, 0.0), r);
        if (c_lo == c_hi) code = (c_lo < 0) ? out_code : (uint8_t)c_lo;
        else code = amb_code;
      }
      (*table)[t] = code;
    }
    if ((*table)[0] != out_code) return -1;
    *inv_h_out = inv_h;
    *nt_out = ntf;
    if (top_out) *top_out = (float)Tf;
    return 0;
  }
  table->assign(nt, 0);
  for (int t = 0; t < nt; ++t) {
    // FP32 values v with min(cvt_rz(fl(v*inv_h)), nt-1) == t.
    double v_lo = (double)t * (1.0 - 4 * u) / ih;
    double v_hi = (t == nt - 1) ? std::numeric_limits<double>::infinity()
                                : (double)(t + 1) * (1.0 + 4 * u) / ih;
    double d_lo = v_lo - 2.0 * err(2 * v_lo + 1);
    double d_hi = std::isinf(v_hi) ? v_hi : v_hi + 2.0 * err(2 * v_hi + 1);
    uint8_t code;
    if (r.zero != 0 && d_lo <= 0.0) {
      code = amb_code;  // exact zero is excluded, anything above is bin 0
    } else {
      int c_lo = host_class(std::max(d_lo, 0.0), r);
      int c_hi = std::isinf(d_hi) ? -1 : host_class(d_hi, r);
      if (c_lo == c_hi) code = (c_lo < 0) ? out_code : (uint8_t)c_lo;
      else code = amb_code;
    }
    (*table)[t] = code;
  }
  if ((*table)[nt - 1] != out_code) return -1;
  *inv_h_out = inv_h;
  *nt_out = nt;
  return 0;
}

}  // namespace pcf

namespace {

int finish_upload(pcf_ctx* c, Catalog* cat) {
  const int nb = 512;
  HIPCHK(c, c->stats_partials.ensure(sizeof(float) * nb * 10));
  float* part = c->stats_partials.as<float>();
  if (cat->n > 0) {
    HIPCHK(c, launch_catalog_stats(cat->raw, cat->n, part, nb, c->stream));
    std::vector<float> h(nb * 10);
    HIPCHK(c, hipMemcpyAsync(h.data(), part, sizeof(float) * nb * 10, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) {
      cat->lo[a] = INFINITY;
      cat->hi[a] = -INFINITY;
    }
    cat->maxabs = 0;
    cat->all_w1 = true;
    cat->finite = true;
    for (int b = 0; b < nb; ++b) {
      const float* p = &h[b * 10];
      for (int a = 0; a < 3; ++a) {
        cat->lo[a] = std::min(cat->lo[a], p[a]);
        cat->hi[a] = std::max(cat->hi[a], p[3 + a]);
      }
      cat->maxabs = std::max(cat->maxabs, p[6]);
      if (p[7] < 1.0f) cat->all_w1 = false;
      if (p[8] < 1.0f) cat->finite = false;
    }
  }
  if (!cat->finite) return fail(c, PCF_ERR_RANGE, "catalog holds non-finite values");
  return PCF_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Geometry and grids
// ---------------------------------------------------------------------------
namespace pcf {

// Neighbour reach of pair counting up to dmax on a grid over [lo, hi] (or the
// periodic box): dmax plus the slack of the FP32 cell assignment.  A point may
// sit ~3u*(|coordinate| + extent) outside its nominal cell after the FP32
// cell assignment (per axis, both points of a pair), and the FP32 pruning
// arithmetic rounds at ~1e-7 relative.
double pair_reach(const double lo[3], const double hi[3], float dmax, int bpc, float box) {
  const double u = std::ldexp(1.0, -24);
  double ext_max = 0;
  for (int a = 0; a < 3; ++a) ext_max = std::max(ext_max, bpc ? (double)box : hi[a] - lo[a]);
  const double delta = 4 * u * (2 * ext_max + std::fabs(lo[0]) + std::fabs(lo[1]) +
                                std::fabs(lo[2]) + 2 * (double)dmax) + 1e-7;
  return (double)dmax * (1.0 + 1e-5) + 4 * delta;
}

int choose_geometry(const Catalog* X, const Catalog* Y, float dmax, int bpc, float box,
                    int cells_per_rmax, GridGeom* g, int K[3], float* reach_out, int* m_out,
                    int zsub) {
  if (zsub < 1) zsub = 1;
  const double u = std::ldexp(1.0, -24);
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = X->n ? X->lo[a] : 0;
    hi[a] = X->n ? X->hi[a] : 1;
    if (Y && Y->n) {
      lo[a] = std::min(lo[a], (double)Y->lo[a]);
      hi[a] = std::max(hi[a], (double)Y->hi[a]);
    }
  }
  int64_t ny = Y ? Y->n : X->n;
  int m = cells_per_rmax;
  if (m <= 0) {
    // Largest m (cells per rmax) whose mean occupancy still fills most of a
    // 64-lane tile.
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) vol *= bpc ? (double)box : std::max(hi[a] - lo[a], (double)dmax);
    double dens = (double)std::max<int64_t>(ny, 1) / vol;
    m = 1;
    for (int t = 4; t >= 1; --t) {
      double cs = dmax / t;
      if (dens * cs * cs * cs >= 32.0) { m = t; break; }
    }
  }
  if (m > 8) m = 8;
  double reach = pair_reach(lo, hi, dmax, bpc, box);
  // Cell edge a hair above reach/m, so the neighbour reach is exactly m cells.
  double cs_target = reach / m * (1.0 + 1e-6);
  g->bpc = bpc;
  g->L = box;
  for (int iter = 0; iter < 64; ++iter) {
    int64_t ncell = 1;
    for (int a = 0; a < 3; ++a) {
      const double cst = a == 2 ? cs_target / zsub : cs_target;
      if (bpc) {
        int P = (int)std::floor((double)box / cst);
        if (P < 3) P = 3;
        g->P[a] = P;
        g->lo[a] = 0.f;
        g->cs[a] = (float)((double)box / P);
      } else {
        // The origin snaps to a multiple of the cell edge, so catalogs with
        // about the same bounds (DD, RR and DR of one survey) share the
        // geometry and hence their resident grids.
        const double lo_s = std::floor(lo[a] / cst) * cst;
        double ext = hi[a] - lo_s;
        int P = (int)std::floor(ext / cst) + 1;
        if (P < 1) P = 1;
        g->P[a] = P;
        g->lo[a] = (float)lo_s;
        g->cs[a] = (float)cst;
      }
      g->inv_cs[a] = 1.0f / g->cs[a];
      ncell *= g->P[a];
    }
    if (ncell <= (int64_t)1 << 22) break;
    cs_target *= 1.25;
  }
  for (int a = 0; a < 3; ++a) {
    K[a] = (int)std::ceil(reach / g->cs[a]);
    // Periodic: offsets up to one period reach every reference term (the
    // regular separation and the close image, both within L); larger ones
    // would only add far images.
    if (bpc && K[a] > g->P[a]) K[a] = g->P[a];
  }
  *reach_out = (float)(reach * (1.0 + 1e-6));
  if (m_out) *m_out = m;
  return 0;
}

uint64_t next_grid_build_id() {
  static std::atomic<uint64_t> ids{0};
  return ++ids;
}

static void free_grid(DevGrid* gr) {
  if (gr->pts) (void)hipFree(gr->pts);
  if (gr->pts64) (void)hipFree(gr->pts64);
  if (gr->cell_start) (void)hipFree(gr->cell_start);
  if (gr->fine_start) (void)hipFree(gr->fine_start);
  if (gr->ready) (void)hipEventDestroy(gr->ready);
  gr->pts = nullptr;
  gr->pts64 = nullptr;
  gr->cell_start = nullptr;
  gr->fine_start = nullptr;
  gr->ready = nullptr;
}

// The catalog's grid for geometry g, built on this context's stream if it is
// not resident (or rebuilt when asked and no other job holds it).  The grid is
// pinned for the caller (release_grid / GridPins); if another stream built it,
// this stream waits for that build.
DevGrid* get_grid(pcf_ctx* c, Catalog* cat, const GridGeom& g, bool force_rebuild, int* rc) {
  std::lock_guard<std::mutex> lk(c->store->mu);
  for (auto& gr : cat->grids) {
    if (gr->valid && gr->g.same(g)) {
      hipError_t e = hipSuccess;
      if (force_rebuild && gr->users == 0) {
        e = build_grid(cat->raw, cat->n, g, gr.get(), c->scratch_cells, c->scratch_perm,
                       c->grid_stream);
        gr->build_id = next_grid_build_id();
        if (e == hipSuccess) e = hipEventRecord(gr->ready, c->grid_stream);
      }
      if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, gr->ready, 0);
      if (e != hipSuccess) {
        *rc = hipfail(c, e, "build_grid");
        return nullptr;
      }
      gr->users++;
      return gr.get();
    }
  }
  // Keep a handful of geometries per catalog (DR/RR reuse, serving): evict the
  // oldest grid nobody holds.
  if (cat->grids.size() >= 4) {
    for (auto it = cat->grids.begin(); it != cat->grids.end(); ++it) {
      if ((*it)->users == 0) {
        (void)hipEventSynchronize((*it)->ready);
        free_grid(it->get());
        cat->grids.erase(it);
        break;
      }
    }
  }
  auto gr = std::make_unique<DevGrid>();
  hipError_t e = hipEventCreateWithFlags(&gr->ready, hipEventDisableTiming);
  if (e == hipSuccess)
    e = build_grid(cat->raw, cat->n, g, gr.get(), c->scratch_cells, c->scratch_perm,
                   c->grid_stream);
  gr->build_id = next_grid_build_id();
  if (e == hipSuccess) e = hipEventRecord(gr->ready, c->grid_stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, gr->ready, 0);
  if (e != hipSuccess) {
    free_grid(gr.get());
    *rc = (e == hipErrorOutOfMemory) ? fail(c, PCF_ERR_NOMEM, "grid allocation")
                                     : hipfail(c, e, "build_grid");
    (void)hipGetLastError();
    return nullptr;
  }
  gr->users = 1;
  cat->grids.push_back(std::move(gr));
  return cat->grids.back().get();
}

void release_grid(pcf_ctx* c, DevGrid* g) {
  std::lock_guard<std::mutex> lk(c->store->mu);
  if (g->users > 0) g->users--;
}

hipError_t ensure_grid64(pcf_ctx* c, DevGrid* g) {
  std::lock_guard<std::mutex> lk(c->store->mu);
  hipError_t e;
  if (!g->pts64) {
    if ((e = hipStreamWaitEvent(c->grid_stream, g->ready, 0)) != hipSuccess) return e;
    if ((e = build_grid64(g, c->grid_stream)) != hipSuccess) return e;
    if ((e = hipEventRecord(g->ready, c->grid_stream)) != hipSuccess) return e;
  }
  return hipStreamWaitEvent(c->stream, g->ready, 0);
}

void free_catalog(Catalog* cat) {
  if (cat->raw) (void)hipFree(cat->raw);
  cat->raw = nullptr;
  for (auto& gr : cat->grids) free_grid(gr.get());
  cat->grids.clear();
}

}  // namespace pcf

// ---------------------------------------------------------------------------
// Pair histograms (2iso / 2ani)
// ---------------------------------------------------------------------------
namespace pcf {

// Largest tile box half-diagonal the Gram path's error bound is built for:
// a 64-point tile spans about 64 / (points per cell) cells of its column
// (the end cells count whole: the bound does not rely on the z sub-slab order
// inside a cell), plus one cell either way; tiles beyond it take the direct
// path.
double gram_rho(const GridGeom& g, int64_t n) {
  const double pc = (double)std::max<int64_t>(n, 1) / (double)g.ncell();
  const double zsub = std::max(1.0, std::round((double)g.cs[0] / (double)g.cs[2]));
  double cells = std::ceil(64.0 / std::max(pc, 1e-9)) + 2.0;
  cells = std::min(8.0 * zsub, std::max(3.0, cells));
  const double cz = cells * g.cs[2];
  const double r2 = 0.25 * ((double)g.cs[0] * g.cs[0] + (double)g.cs[1] * g.cs[1] + cz * cz);
  return std::sqrt(r2) * 1.02;
}

}  // namespace pcf

namespace {

int64_t gx_points(const Catalog* X) { return X->n; }

// Carves aligned pieces out of a fixed block of device memory.  With base ==
// nullptr it only adds up the bytes a plan needs.
struct Carve {
  char* base;
  size_t used = 0;
  explicit Carve(void* b) : base(reinterpret_cast<char*>(b)) {}
  template <class T>
  T* take(size_t bytes) {
    const size_t off = (used + 255) & ~(size_t)255;
    used = off + std::max<size_t>(bytes, 1);
    return base ? reinterpret_cast<T*>(base + off) : nullptr;
  }
};

}  // namespace

namespace pcf {

// One pair histogram (2iso / 2ani, XX / XY).  Device buffers come from the
// context (growable buffers, resident grid cache) or, with arena != nullptr,
// from the arena_bytes of device memory at arena: nothing is allocated then
// and nothing outlives the call.
int run_pairs(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights, void* arena,
              size_t arena_bytes) {
  const bool auto_ = job->hist == PCF_XX;
  Catalog* X = get_cat(c, job->cat_x);
  Catalog* Y = auto_ ? X : get_cat(c, job->cat_y);
  if (!X || !Y) return fail(c, PCF_ERR_ARG, "unknown catalog id");
  const int bins = job->bins;
  const bool ani = job->type == PCF_2ANI;
  const int64_t nflat64 = ani ? (int64_t)bins * bins : (int64_t)bins;
  if (bins <= 0 || nflat64 > kMaxFlatBins) return fail(c, PCF_ERR_ARG, "bins out of range");
  if (!(job->dmax > 0) || !std::isfinite(job->dmax)) return fail(c, PCF_ERR_ARG, "dmax must be > 0");
  const int bpc = job->bpc ? 1 : 0;
  const float L = job->box;
  if (bpc) {
    if (!(L > 0)) return fail(c, PCF_ERR_ARG, "box side required for periodic boundaries");
    // 2*dmax >= L is handled (a pair may count under its regular and its
    // projected image, as the reference's per-axis flags allow); dmax >= L
    // would also reach the far image of every axis, which no reference term is.
    if (!(job->dmax < L))
      return fail(c, PCF_ERR_UNSUPPORTED, "periodic boundaries need dmax < box side");
    for (Catalog* cc : {X, Y})
      for (int a = 0; a < 3; ++a)
        if (cc->n && (cc->lo[a] < 0.f || !(cc->hi[a] < L)))
          return fail(c, PCF_ERR_RANGE, "coordinates outside [0, box) with periodic boundaries");
  }
  const int nflat = (int)nflat64;
  const bool weighted = !(X->all_w1 && Y->all_w1);
  // The mixed (FP32 + FP64 recheck) kernel needs its u8 d2 -> bin table
  // codes, its thresholds and LDS histogram to fit, and j-side points that its
  // recheck queue can index (27 bits); everything else runs the plain FP64
  // kernel, which gives the same results.
  const bool mixed_fits = pair_mixed_fits(bins, ani, weighted);
  const bool plain = job->precision == PCF_PREC_FP64 || X->n >= kRecheckMaxPoints ||
                     Y->n >= kRecheckMaxPoints || !mixed_fits;
  // Histograms larger than an LDS copy deposit straight into global memory.
  const bool global_hist =
      plain && (size_t)(nflat + 1) * (weighted ? 12 : 4) > kPlainLdsHistMax;
  // Gram-form path (matrix cores): unweighted, mixed precision.  Its cells
  // are cut kGramZsub times thinner along z: the j range of a tile then ends
  // within a thin cell of what the tile can reach.
  const bool gram = !plain && !weighted && job->pair_kernel != PCF_PAIR_DIRECT;

  GridGeom g;
  int K[3];
  float reach;
  int m = 0;
  if (choose_geometry(X, auto_ ? nullptr : Y, job->dmax, bpc, L, job->cells_per_rmax, &g, K,
                      &reach, &m, gram ? kGramZsub : 1) != 0)
    return fail(c, PCF_ERR_UNSUPPORTED, "geometry");

  // Rules, thresholds and the FP32 table.
  static thread_local BinRules rules;
  make_rules(job->type, bpc, auto_, bins, job->dmax, L, &rules);
  double A = bpc ? (double)std::max(X->maxabs, Y->maxabs) + (double)L : 0.0;
  const int nt_max = kMixedTableMax;
  // The Gram path's error bound covers tiles whose bounding box has
  // half-diagonal <= rho.
  const double rho = gram ? gram_rho(g, gx_points(X)) : -1.0;
  static thread_local std::vector<uint8_t> table;
  float inv_h = 1.f;
  int nt = 2;
  float tflip = 0.f;  // flipped table: index cvt_rz(tflip - d2) (per component for 2ani)
  if (!plain && build_table(rules, A, nt_max, &table, &inv_h, &nt, rho, true, &tflip) != 0)
    return fail(c, PCF_ERR_UNSUPPORTED, "bin table");
  const float sigma = std::sqrt(inv_h);  // exact power of two
  const size_t rt_bytes = sizeof(BinRules) + (plain ? 0 : (size_t)nt);
  const size_t hist_bytes = sizeof(unsigned long long) * nflat;
  const size_t w_bytes = sizeof(double) * nflat;
  const size_t rd_bytes = hist_bytes + w_bytes + 32;  // histogram, weights, statistics
  // tail: 4 u64 statistics, then the work cursors (8 ranges x 64 bytes)
  const size_t tail = 32 + 8 * 64;

  WorkListArgs wa;
  wa.g = g;
  for (int a = 0; a < 3; ++a) wa.K[a] = K[a];
  wa.reach = reach;
  wa.auto_ = auto_ ? 1 : 0;
  wa.cylinder = ani ? 1 : 0;
  wa.map = nullptr;
  wa.y_fine_start = nullptr;
  wa.x_pts = nullptr;
  wa.zq = 1;
  const int64_t cap = worklist_bound(wa, X->n);

  const bool rebuild = job->rebuild_grid != 0;
  int rc = PCF_OK;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  GridPins pins(c);
  DevGrid* gx = nullptr;
  DevGrid* gy = nullptr;
  DevGrid ax, ay;  // arena grids (not cached)
  const uint8_t* rt_dev = nullptr;
  void* hist_dev = nullptr;
  float4* scaled_x = nullptr;
  float4* scaled_y = nullptr;
  PairItem* items = nullptr;
  int nitems = 0;
  const int* nitems_dev = nullptr;
  std::vector<uint8_t> host_stage;  // arena mode: pageable staging
  uint8_t* rd = nullptr;

  if (!arena) {
    gx = pins.add(get_grid(c, X, g, rebuild, &rc));
    if (!gx) return rc;
    gy = auto_ ? gx : pins.add(get_grid(c, Y, g, rebuild, &rc));
    if (!gy) return rc;
    if (plain) {
      HIPCHK(c, ensure_grid64(c, gx));
      if (!auto_) HIPCHK(c, ensure_grid64(c, gy));
    }
    // Rules and table in one device buffer, uploaded from page-locked staging
    // (asynchronous) and only when they differ from the last upload.  The
    // staging bytes are not touched again before this call's final sync.
    const size_t rd_off = (rt_bytes + 255) & ~(size_t)255;
    HIPCHK(c, c->pin.ensure(rd_off + rd_bytes));
    HIPCHK(c, c->pair_rt.ensure(rt_bytes));
    uint8_t* st = c->pin.as<uint8_t>();
    std::memcpy(st, &rules, sizeof(BinRules));
    if (!plain) std::memcpy(st + sizeof(BinRules), table.data(), nt);
    if (c->pair_rt_host.size() != rt_bytes ||
        std::memcmp(c->pair_rt_host.data(), st, rt_bytes) != 0) {
      HIPCHK(c, hipMemcpyAsync(c->pair_rt.p, st, rt_bytes, hipMemcpyHostToDevice, c->stream));
      c->pair_rt_host.assign(st, st + rt_bytes);
    }
    rt_dev = c->pair_rt.as<uint8_t>();
    rd = st + rd_off;
    // The list is sized by its bound and its length stays on the device (no
    // host round trip), unless the bound is too large to allocate casually.
    wa.x_cell_start = gx->cell_start;
    wa.y_cell_start = gy->cell_start;
    if (gx->zq > 1 && gy->zq == gx->zq) {
      wa.y_fine_start = gy->fine_start;
      wa.x_pts = gx->pts;
      wa.zq = gx->zq;
    }
    if (cap <= ((int64_t)1 << 25)) {
      HIPCHK(c, build_pair_worklist_async(wa, cap, c->wl_counts, c->wl_items, &nitems_dev,
                                          c->stream));
    } else {
      HIPCHK(c, build_pair_worklist(wa, c->wl_counts, c->wl_items, &nitems, c->stream));
    }
    items = c->wl_items.as<PairItem>();
    HIPCHK(c, c->hist_buf.ensure(hist_bytes + w_bytes + tail));
    hist_dev = c->hist_buf.p;
    if (!plain) {
      HIPCHK(c, c->scaled_x.ensure(sizeof(float4) * (size_t)std::max<int64_t>(gx->n, 1)));
      scaled_x = c->scaled_x.as<float4>();
      if (!auto_) {
        HIPCHK(c, c->scaled_y.ensure(sizeof(float4) * (size_t)std::max<int64_t>(gy->n, 1)));
        scaled_y = c->scaled_y.as<float4>();
      }
    }
  } else {
    // Plan the arena (sizes only), check it, then carve it for real.
    const int64_t ncell = g.ncell();
    const int64_t nmax = std::max<int64_t>(X->n, Y->n);
    auto plan = [&](Carve& cv) {
      ax.pts = cv.take<float4>(sizeof(float4) * (size_t)X->n);
      ax.cell_start = cv.take<int>(sizeof(int) * (size_t)(ncell + 1));
      if (plain) ax.pts64 = cv.take<double4>(sizeof(double4) * (size_t)X->n);
      if (!auto_) {
        ay.pts = cv.take<float4>(sizeof(float4) * (size_t)Y->n);
        ay.cell_start = cv.take<int>(sizeof(int) * (size_t)(ncell + 1));
        if (plain) ay.pts64 = cv.take<double4>(sizeof(double4) * (size_t)Y->n);
      }
      void* cells_tmp = cv.take<void>(grid_cells_tmp_bytes(ncell));
      int* perm_tmp = cv.take<int>(2 * sizeof(int) * (size_t)std::max<int64_t>(nmax, 1));
      rt_dev = cv.take<uint8_t>(rt_bytes);
      void* wl_counts = cv.take<void>(worklist_count_bytes(wa));
      items = cv.take<PairItem>(sizeof(PairItem) * (size_t)std::max<int64_t>(cap, 1));
      hist_dev = cv.take<void>(hist_bytes + w_bytes + tail);
      if (!plain) {
        scaled_x = cv.take<float4>(sizeof(float4) * (size_t)X->n);
        if (!auto_) scaled_y = cv.take<float4>(sizeof(float4) * (size_t)Y->n);
      }
      return std::make_pair(cells_tmp, std::make_pair(perm_tmp, wl_counts));
    };
    Carve sizing(nullptr);
    plan(sizing);
    if (sizing.used > arena_bytes)
      return fail(c, PCF_ERR_NOMEM,
                  "job needs " + std::to_string(sizing.used) +
                      " bytes of scratch; the no-allocation entry point has " +
                      std::to_string(arena_bytes));
    Carve cv(arena);
    auto tmp = plan(cv);
    void* cells_tmp = tmp.first;
    int* perm_tmp = tmp.second.first;
    void* wl_counts = tmp.second.second;
    ax.g = g;
    ax.n = X->n;
    HIPCHK(c, build_grid_into(X->raw, X->n, g, ax.pts, ax.cell_start, nullptr, 1, cells_tmp,
                              perm_tmp, c->stream));
    if (plain) HIPCHK(c, convert_grid64(ax.pts, ax.n, ax.pts64, c->stream));
    gx = &ax;
    gy = gx;
    if (!auto_) {
      ay.g = g;
      ay.n = Y->n;
      HIPCHK(c, build_grid_into(Y->raw, Y->n, g, ay.pts, ay.cell_start, nullptr, 1, cells_tmp,
                                perm_tmp, c->stream));
      if (plain) HIPCHK(c, convert_grid64(ay.pts, ay.n, ay.pts64, c->stream));
      gy = &ay;
    }
    host_stage.resize(std::max(rt_bytes, rd_bytes));
    std::memcpy(host_stage.data(), &rules, sizeof(BinRules));
    if (!plain) std::memcpy(host_stage.data() + sizeof(BinRules), table.data(), nt);
    // Pageable staging, alive until the final synchronisation below.
    HIPCHK(c, hipMemcpyAsync(const_cast<uint8_t*>(rt_dev), host_stage.data(), rt_bytes,
                             hipMemcpyHostToDevice, c->stream));
    rd = host_stage.data();
    wa.x_cell_start = gx->cell_start;
    wa.y_cell_start = gy->cell_start;
    HIPCHK(c, build_pair_worklist_into(wa, cap, wl_counts, items, &nitems_dev, c->stream));
  }

  // Histogram + counters.
  unsigned long long* d_hist = reinterpret_cast<unsigned long long*>(hist_dev);
  double* d_w = reinterpret_cast<double*>(reinterpret_cast<char*>(hist_dev) + hist_bytes);
  unsigned long long* d_stats = reinterpret_cast<unsigned long long*>(
      reinterpret_cast<char*>(hist_dev) + hist_bytes + w_bytes);
  int* d_counter = reinterpret_cast<int*>(d_stats + 4);
  HIPCHK(c, launch_zero(hist_dev, hist_bytes + w_bytes + tail, c->stream));

  PairLaunch pl{};
  pl.xp = gx->pts;
  pl.yp = gy->pts;
  if (!plain) {
    // Scaled copies for the bulk path (x, y, z times sigma, exact).
    HIPCHK(c, launch_scale_points(gx->pts, gx->n, sigma, scaled_x, c->stream));
    pl.xs = scaled_x;
    if (auto_) {
      pl.ys = pl.xs;
    } else {
      HIPCHK(c, launch_scale_points(gy->pts, gy->n, sigma, scaled_y, c->stream));
      pl.ys = scaled_y;
    }
  }
  pl.xp64 = gx->pts64;
  pl.yp64 = gy->pts64;
  pl.items = items;
  pl.nitems = nitems;
  pl.nitems_dev = nitems_dev;
  pl.nitems_cap = (int)std::min<int64_t>(cap, INT32_MAX);
  pl.counter = d_counter;
  pl.table = rt_dev + sizeof(BinRules);
  pl.nt = nt;
  pl.inv_h = inv_h;
  pl.L = L * sigma;
  pl.bins = bins;
  pl.nflat = nflat;
  const PairShape sh =
      pair_shape(nflat, bins, weighted, plain, job->lds_copies_log2, job->blocks_per_cu);
  pl.cshift = global_hist ? 0 : sh.cshift;
  pl.global_hist = global_hist ? 1 : 0;
  pl.hist = d_hist;
  pl.whist = d_w;
  pl.stats = d_stats;
  pl.rules = reinterpret_cast<const BinRules*>(rt_dev);
  pl.ani = ani ? 1 : 0;
  pl.weighted = weighted ? 1 : 0;
  pl.plain64 = plain ? 1 : 0;
  pl.R = rules.R;
  pl.Lx = (double)L;
  pl.le = rules.le;
  pl.zero = rules.zero;
  pl.nblocks = std::max(1, c->ncu * (global_hist ? 8 : sh.per_cu));
  pl.mfma = gram ? 1 : 0;
  // The kernel's own test on the tile box, 2% inside the table's margin.
  pl.rho2 = gram ? (float)(rho * rho / (1.02 * 1.02) * (double)inv_h) : 0.f;
  pl.tflip = tflip;
  pl.geo = g;
  pl.sigma = sigma;
  const bool launched = nitems_dev != nullptr || nitems > 0;
  if (launched) {
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, launch_pairs(pl, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  }
  // Histogram, weights and statistics are contiguous: one read-back.
  HIPCHK(c, hipMemcpyAsync(rd, d_hist, rd_bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const unsigned long long* hc = reinterpret_cast<const unsigned long long*>(rd);
  const double* hw = reinterpret_cast<const double*>(rd + hist_bytes);
  const unsigned long long* hs = reinterpret_cast<const unsigned long long*>(rd + hist_bytes + w_bytes);
  if (nitems_dev && (int64_t)hs[2] > cap) {
    // More items than the allocation: the bound did not hold and items were
    // dropped.  Never expected; an error rather than a short histogram.
    return fail(c, PCF_ERR_HIP, "work list exceeded its bound");
  }
  const uint64_t mult = auto_ ? 2 : 1;
  for (int b = 0; b < nflat; ++b) {
    if (counts) counts[b] = (uint64_t)hc[b] * mult;
    if (weights) weights[b] = weighted ? hw[b] * (double)mult : (double)((uint64_t)hc[b] * mult);
  }
  pcf_stats st{};
  st.evaluations = hs[0];
  st.rechecks = hs[1];
  st.work_items = nitems_dev ? hs[2] : (uint64_t)nitems;
  st.direct_items = (int32_t)std::min<uint64_t>(hs[3], INT32_MAX);
  float ms = 0.f;
  if (launched) {
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
    st.kernel_ms = ms;
  }
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
  st.total_ms = ms;
  st.cells_per_rmax = m;
  c->last = st;
  return PCF_OK;
}

}  // namespace pcf

namespace {

// wrap32: the reference's int shell count, which wraps from i = 26755 on
// (pcf_set_ref_bugs); else 64-bit, with no overflow for large bin counts.
__global__ void k_analytic_rr(double alpha, int bins, double* rr, int wrap32) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < bins) {
    const unsigned u = (unsigned)i;
    long long k = wrap32 ? (long long)(int)(3u * u * u + 3u * u + 1u) : 3LL * i * i + 3LL * i + 1;
    rr[i] = alpha * (double)k;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int64_t pcf_hist_size(int32_t type, int32_t bins) {
  int64_t b = bins;
  switch (type) {
    case PCF_2ISO: return b;
    case PCF_2ANI: return b * b;
    case PCF_3ISO: return b * b * b;
    case PCF_3ANI: return b * b * b * b * b;
  }
  return -1;
}

const char* pcf_create_error(void) { return g_create_error.c_str(); }

int pcf_create(pcf_ctx** out, int32_t device) {
  if (!out) return PCF_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
    (void)hipGetLastError();
    return PCF_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    g_create_error = "device index out of range";
    return PCF_ERR_ARG;
  }
  auto c = new pcf_ctx();
  c->device = device;
  c->store = std::make_shared<CatalogStore>();
  if ((e = hipSetDevice(device)) != hipSuccess) goto fail;
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) goto fail;
  c->ncu = prop.multiProcessorCount;
  // The no-allocation entry point's scratch array on this device.
  if ((e = resolve_noalloc_scratch(device)) != hipSuccess) goto fail;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) goto fail;
  if ((e = hipStreamCreateWithFlags(&c->grid_stream, hipStreamNonBlocking)) != hipSuccess)
    goto fail;
  for (auto& ev : c->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) goto fail;
  *out = c;
  return PCF_OK;
fail:
  g_create_error = std::string("HIP init: ") + hipGetErrorString(e);
  (void)hipGetLastError();
  delete c;
  return PCF_ERR_HIP;
}

void pcf_destroy(pcf_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  destroy_engine(c->engine);
  c->engine = nullptr;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_batch_plan(c->batch);
  c->batch = nullptr;
  if (!c->lane && c->store) {
    std::lock_guard<std::mutex> lk(c->store->mu);
    for (auto& cat : c->store->cats)
      if (cat) free_catalog(cat.get());
    c->store->cats.clear();
  }
  for (DevBuf* b : {&c->scratch_cells, &c->scratch_perm, &c->wl_counts, &c->wl_items,
                    &c->table_buf, &c->rules_buf, &c->hist_buf, &c->misc_buf, &c->stats_partials,
                    &c->scaled_x, &c->scaled_y, &c->trip_idx, &c->trip_pairs, &c->trip_hist,
                    &c->wl_items2, &c->trip_qidx, &c->trip_q, &c->pair_rt})
    b->release();
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->pin.release();
  for (auto& ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->grid_stream) (void)hipStreamSynchronize(c->grid_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->grid_stream) (void)hipStreamDestroy(c->grid_stream);
  delete c;
}

const char* pcf_last_error(const pcf_ctx* c) { return c ? c->err.c_str() : "null context"; }

static int upload_common(pcf_ctx* c, const void* src, int64_t n, hipMemcpyKind kind,
                         pcf_catalog* id) {
  if (!c || !id || n < 0 || (n > 0 && !src)) return fail(c, PCF_ERR_ARG, "bad upload arguments");
  if (n > (int64_t)INT32_MAX - 64) return fail(c, PCF_ERR_ARG, "catalog too large");
  HIPCHK(c, hipSetDevice(c->device));
  auto cat = std::make_unique<Catalog>();
  cat->n = n;
  {
    std::lock_guard<std::mutex> lk(c->store->mu);
    cat->serial = c->store->next_serial++;
  }
  HIPCHK(c, hipMalloc(&cat->raw, sizeof(float4) * (n > 0 ? n : 1)));
  if (n > 0) {
    hipError_t e = hipMemcpyAsync(cat->raw, src, sizeof(float4) * n, kind, c->stream);
    if (e != hipSuccess) {
      (void)hipFree(cat->raw);
      return hipfail(c, e, "catalog copy");
    }
  }
  int rc = finish_upload(c, cat.get());
  if (rc != PCF_OK) {
    free_catalog(cat.get());
    return rc;
  }
  std::lock_guard<std::mutex> lk(c->store->mu);
  auto& cats = c->store->cats;
  for (size_t i = 0; i < cats.size(); ++i)
    if (!cats[i]) {
      cats[i] = std::move(cat);
      *id = (pcf_catalog)i;
      return PCF_OK;
    }
  cats.push_back(std::move(cat));
  *id = (pcf_catalog)(cats.size() - 1);
  return PCF_OK;
}

int pcf_catalog_upload(pcf_ctx* c, const float* xyzw, int64_t n, pcf_catalog* id) {
  return upload_common(c, xyzw, n, hipMemcpyHostToDevice, id);
}

int pcf_catalog_upload_device(pcf_ctx* c, const void* d_xyzw, int64_t n, pcf_catalog* id) {
  return upload_common(c, d_xyzw, n, hipMemcpyDeviceToDevice, id);
}

int pcf_catalog_release(pcf_ctx* c, pcf_catalog id) {
  if (!c) return PCF_ERR_ARG;
  Catalog* cat = get_cat(c, id);
  if (!cat) return fail(c, PCF_ERR_ARG, "unknown catalog id");
  (void)hipStreamSynchronize(c->stream);
  if (c->engine) engine_drain(c->engine);
  std::lock_guard<std::mutex> lk(c->store->mu);
  free_catalog(cat);
  c->store->cats[id].reset();
  return PCF_OK;
}

int64_t pcf_catalog_size(pcf_ctx* c, pcf_catalog id) {
  if (!c) return -1;
  Catalog* cat = get_cat(c, id);
  return cat ? cat->n : -1;
}

int pcf_histogram(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights) {
  if (!c || !job) return fail(c, PCF_ERR_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  c->err.clear();
  return dispatch_histogram(c, job, counts, weights);
}

}  // extern "C"

namespace pcf {

// The reference's XY2ani weight defect applies: 2ani DR without periodic
// boundaries under pcf_set_ref_bugs, with some weight other than 1.
bool ref_bugs_alias_needed(pcf_ctx* c, const pcf_job* job) {
  if (!ref_bugs(c) || job->type != PCF_2ANI || job->hist != PCF_XY || job->bpc) return false;
  Catalog* X = get_cat(c, job->cat_x);
  Catalog* Y = get_cat(c, job->cat_y);
  return X && Y && !(X->all_w1 && Y->all_w1);
}

// One histogram on this context's stream (also the engine lanes' entry).
int dispatch_histogram(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights) {
  if (!job) return fail(c, PCF_ERR_ARG, "null argument");
  switch (job->type) {
    case PCF_2ISO:
    case PCF_2ANI:
      if (job->hist != PCF_XX && job->hist != PCF_XY)
        return fail(c, PCF_ERR_ARG, "pair types take PCF_XX or PCF_XY");
      {
        pcf_job tuned = *job;
        if (get_cat(c, job->cat_x) && (job->hist == PCF_XX || get_cat(c, job->cat_y)))
          apply_tuning(c, &tuned, 1);
        if (ref_bugs_alias_needed(c, job)) {
          // The reference's XY2ani weight (cuda/kernels/2ani_k.cuh:128): the
          // random points take the data weights of their ordered slots.
          pcf_catalog alias = -1;
          int rc = ref_alias_catalog(c, job->cat_x, job->cat_y, job->box, job->partitions, &alias);
          if (rc != PCF_OK) return rc;
          tuned.cat_y = alias;
          rc = run_pairs(c, &tuned, counts, weights, nullptr, 0);
          std::string keep = c->err;
          (void)pcf_catalog_release(c, alias);
          c->err = keep;
          return rc;
        }
        return run_pairs(c, &tuned, counts, weights, nullptr, 0);
      }
    case PCF_3ISO:
    case PCF_3ANI:
      if (job->hist != PCF_XXX && job->hist != PCF_XXY)
        return fail(c, PCF_ERR_ARG, "triplet types take PCF_XXX or PCF_XXY");
      return run_triplets(c, job, counts, weights);
  }
  return fail(c, PCF_ERR_ARG, "unknown histogram type");
}

// A serving lane: same device and resident catalogs as the parent, its own
// stream, events and scratch buffers.
int create_lane(pcf_ctx* parent, pcf_ctx** out) {
  auto c = new pcf_ctx();
  c->device = parent->device;
  c->ncu = parent->ncu;
  c->store = parent->store;
  c->lane = true;
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->grid_stream, hipStreamNonBlocking);
  for (auto& ev : c->ev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  if (e != hipSuccess) {
    pcf_destroy(c);
    return hipfail(parent, e, "engine lane");
  }
  *out = c;
  return PCF_OK;
}

}  // namespace pcf

extern "C" {

int pcf_last_stats(const pcf_ctx* c, pcf_stats* out) {
  if (!c || !out) return PCF_ERR_ARG;
  *out = c->last;
  return PCF_OK;
}

int pcf_analytic_rr(pcf_ctx* c, int32_t bins, int64_t np, float dmax, float box, double* rr) {
  if (!c || !rr || bins <= 0 || np < 0 || !(dmax > 0) || !(box > 0))
    return fail(c, PCF_ERR_ARG, "bad analytic RR arguments");
  HIPCHK(c, hipSetDevice(c->device));
  // Host constants of cuda/pcf2isoBPC.cuh:199-202 with FP32 dr and beta as the
  // reference forms them, np*np in 64-bit (its 32-bit wrap under
  // pcf_set_ref_bugs).
  const bool wrap = ref_bugs(c);
  float drf = dmax / (float)bins;
  double dr = drf;
  float np2 = wrap ? (float)(int32_t)(uint32_t)((uint64_t)np * (uint64_t)np)
                   : (float)((double)np * (double)np);
  float L3 = box * box * box;
  float betaf = np2 / L3;
  double beta = betaf;
  double alpha = 8 * dr * dr * dr * std::acos(0.0) * beta / 3;
  HIPCHK(c, c->misc_buf.ensure(sizeof(double) * bins));
  hipLaunchKernelGGL(k_analytic_rr, dim3((bins + 255) / 256), dim3(256), 0, c->stream, alpha,
                     (int)bins, c->misc_buf.as<double>(), wrap ? 1 : 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(rr, c->misc_buf.p, sizeof(double) * bins, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PCF_OK;
}

}  // extern "C"

extern "C" int pcf_selftest_scan(pcf_ctx* c, const int32_t* counts, int64_t n, int32_t* total) {
  using namespace pcf;
  if (!c || !total || n < 0 || (n > 0 && !counts)) return fail(c, PCF_ERR_ARG, "bad scan arguments");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t ibytes = (sizeof(int) * (size_t)(2 * n + 2) + 255) & ~(size_t)255;
  HIPCHK(c, c->misc_buf.ensure(ibytes + sizeof(long long) * (size_t)scan_tiles(n)));
  int* in = c->misc_buf.as<int>();
  int* out = in + n + 1;
  long long* tsum = reinterpret_cast<long long*>(c->misc_buf.as<char>() + ibytes);
  if (n > 0)
    HIPCHK(c, hipMemcpyAsync(in, counts, sizeof(int) * (size_t)n, hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, launch_scan(in, n, out, tsum, c->stream));
  HIPCHK(c, hipMemcpyAsync(total, out + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PCF_OK;
}
