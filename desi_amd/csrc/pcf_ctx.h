// pcf_ctx.h -- the context object behind the C ABI and the host helpers
// shared by the pair, triplet and serving translation units.
#pragma once

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pcf_abi.h"
#include "pcf_internal.h"

namespace pcf {
struct JobEngine;
struct BatchPlan;
void destroy_engine(JobEngine* e);
void destroy_batch_plan(BatchPlan* p);
void engine_drain(JobEngine* e);
}  // namespace pcf

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
namespace pcf {
// Resident catalogs (and their grids) and the tuning cache of a context,
// shared with the serving engine's lanes, which run jobs concurrently on their
// own streams.  `mu` guards the catalog table, every grid cache and the tuning
// map; the device work itself is ordered by events (DevGrid::ready).
struct CatalogStore {
  std::vector<std::unique_ptr<Catalog>> cats;
  std::map<std::string, pcf_tuning> tuned;
  uint64_t next_serial = 1;
  std::mutex mu;
  // pcf_set_ref_bugs: reproduce the reference defects that depend on its node
  // order and 32-bit arithmetic (read by the context and its serving lanes).
  std::atomic<bool> ref_bugs{false};
};

// The triplet pair list left in trip_idx (offsets) and trip_pairs (records)
// by the last triplet job, and what it was built from.  DDD and DDR (DRR and
// RRR) list the same pairs of one catalog, so the second job reuses it.
struct TripPairKey {
  uint64_t grid_build = 0;  // DevGrid::build_id of the catalog's grid
  int bins = 0, bpc = 0, ani = 0, key_t = 0, nt = 0;
  int K[3] = {0, 0, 0};
  float dmax = 0.f, L = 0.f, sigma = 0.f, reach = 0.f;
  double R = 0.0;
  bool same(const TripPairKey& o) const {
    return grid_build == o.grid_build && bins == o.bins && bpc == o.bpc && ani == o.ani &&
           key_t == o.key_t && nt == o.nt && K[0] == o.K[0] && K[1] == o.K[1] &&
           K[2] == o.K[2] && dmax == o.dmax && L == o.L && sigma == o.sigma &&
           reach == o.reach && R == o.R;
  }
};
struct TripPairCache {
  bool valid = false;
  TripPairKey key;
  std::vector<uint8_t> ptab;  // the pair table the list was classified with
  int npairs = 0;
};
}  // namespace pcf

struct pcf_ctx {
  int device = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;
  hipStream_t grid_stream = nullptr;  // grid builds (never captured into a graph)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::string err;
  std::shared_ptr<pcf::CatalogStore> store;
  bool lane = false;  // an engine lane: shares the store of the context that made it
  pcf::DevBuf scratch_cells, scratch_perm, wl_counts, wl_items, table_buf, rules_buf, hist_buf,
      misc_buf, stats_partials, scaled_x, scaled_y, trip_idx, trip_pairs, trip_hist, wl_items2,
      trip_qidx, trip_q;
  // Pair path: BinRules + d2 -> bin table on the device, and the host bytes
  // last uploaded there (an unchanged job skips the upload).
  pcf::DevBuf pair_rt;
  std::vector<uint8_t> pair_rt_host;
  pcf::PinnedBuf pin;  // staging for uploads and result read-backs
  pcf_stats last{};
  pcf::JobEngine* engine = nullptr;  // asynchronous serving engine (jobs.hip)
  pcf::BatchPlan* batch = nullptr;   // last batched launch plan + graph (batch.hip)
  pcf::TripPairCache trip_cache;     // what trip_idx / trip_pairs hold (triplets.hip)
};


namespace pcf {

inline int fail(pcf_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

inline int hipfail(pcf_ctx* c, hipError_t e, const char* what) {
  return fail(c, PCF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, call)                                        \
  do {                                                           \
    hipError_t e_ = (call);                                      \
    if (e_ != hipSuccess) {                                      \
      (void)hipGetLastError();                                   \
      if (e_ == hipErrorOutOfMemory) return fail(ctx, PCF_ERR_NOMEM, #call); \
      return hipfail(ctx, e_, #call);                            \
    }                                                            \
  } while (0)

inline Catalog* get_cat(pcf_ctx* c, pcf_catalog id) {
  std::lock_guard<std::mutex> lk(c->store->mu);
  if (id < 0 || id >= (pcf_catalog)c->store->cats.size()) return nullptr;
  return c->store->cats[id].get();
}

// Grids handed out by get_grid are pinned (not evicted or rebuilt by another
// lane) until released; GridPins releases them when it goes out of scope,
// after the context's stream has drained.
void release_grid(pcf_ctx* c, DevGrid* g);
struct GridPins {
  pcf_ctx* c;
  std::vector<DevGrid*> g;
  explicit GridPins(pcf_ctx* ctx) : c(ctx) {}
  DevGrid* add(DevGrid* d) {
    if (d) g.push_back(d);
    return d;
  }
  ~GridPins() {
    if (!g.empty()) (void)hipStreamSynchronize(c->stream);
    for (DevGrid* d : g) release_grid(c, d);
  }
};
// FP64 copy of a pinned grid (built once, under the store lock).
hipError_t ensure_grid64(pcf_ctx* c, DevGrid* g);


// Defined in abi.hip.
double ds_floored(int bins, float dmax);
void bin_thresholds(int bins, double ds, double* T);
// rho >= 0: the table also covers the Gram-form pair path for tiles of box
// half-diagonal <= rho (unscaled units).
// flip: the table is indexed by cvt_rz(*top_out - d2) (d2 in table units)
// instead of min(cvt_rz(d2), nt - 1); used by the pair kernels (2iso, and each
// component of 2ani).
int build_table(const BinRules& r, double A, int nt_max, std::vector<uint8_t>* table, float* inv_h,
                int* nt_out, double rho = -1.0, bool flip = false, float* top_out = nullptr);
// zsub > 1 cuts every cell into zsub thinner cells along z (the line of the
// grid's columns), so the j range of a tile (whole z cells) overshoots what
// the tile can reach by at most a thin cell at either end.
// Largest tile half-diagonal (unscaled units) the Gram pair path's error
// bound is built for on geometry g with n points per catalog.
double gram_rho(const GridGeom& g, int64_t n);
double pair_reach(const double lo[3], const double hi[3], float dmax, int bpc, float box);
int choose_geometry(const Catalog* X, const Catalog* Y, float dmax, int bpc, float box,
                    int cells_per_rmax, GridGeom* g, int K[3], float* reach_out, int* m_out,
                    int zsub = 1);
DevGrid* get_grid(pcf_ctx* c, Catalog* cat, const GridGeom& g, bool force_rebuild, int* rc);
void free_catalog(Catalog* cat);

// Defined in abi.hip: one pair histogram; with arena != nullptr every device
// buffer is carved from those arena_bytes (nothing allocated).
int run_pairs(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights, void* arena,
              size_t arena_bytes);

// Defined in noalloc.hip: device address of this GPU's scratch array of
// pcf_histogram_noalloc (once per device, with the device current).
hipError_t resolve_noalloc_scratch(int device);

// Defined in triplets.hip.
int run_triplets(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights);

// Defined in batch.hip.
int run_pairs_batch(pcf_ctx* c, const pcf_job* jobs, int njobs, uint64_t* counts,
                    double* weights);

// Defined in abi.hip: one histogram on c's stream (pcf_histogram's body, also
// run by the serving lanes), and a serving lane of a context.
int dispatch_histogram(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights);
bool ref_bugs_alias_needed(pcf_ctx* c, const pcf_job* job);
int create_lane(pcf_ctx* parent, pcf_ctx** out);

// Defined in tune.hip.
void apply_tuning(pcf_ctx* c, pcf_job* jobs, int njobs);

// Defined in refbugs.hip: the reference's node order of a catalog (stable sort
// of its node ids on partitions^3 nodes of side box / partitions): order[q]
// is the catalog index in slot q, node_start[n] the first slot of node n
// (nodes + 1 entries), sorted_pts (may be null) the points in slot order;
// and a catalog whose weights are the ones XY2ani reads for `rand`.
int ref_partitions(int partitions);
int build_ref_order(pcf_ctx* c, const Catalog* X, float box, int partitions, DevBuf* order,
                    DevBuf* node_start, DevBuf* sorted_pts);
int ref_alias_catalog(pcf_ctx* c, pcf_catalog data, pcf_catalog rand, float box, int partitions,
                      pcf_catalog* out);
inline bool ref_bugs(const pcf_ctx* c) { return c->store && c->store->ref_bugs.load(); }

// Defined in analytic.hip.
int run_analytic_3iso(pcf_ctx* c, pcf_catalog data, int bins, float dmax, float box, int64_t np,
                      int ptt, int bn_ref, double* ddr, double* rrr, double* ff_out,
                      double* ffr_out);

}  // namespace pcf
