// tune.hip -- autotuner of the pair kernels (pcf_autotune).
//
// The knobs are the pair kernel (unweighted jobs: the direct per-lane form or
// the Gram form on the matrix cores), the cell edge (rmax / m:
// candidates per useful pair against work items per point), the number of LDS
// copies of each histogram row
// (fewer bank conflicts on the deposit against LDS per workgroup) and the
// resident workgroups per CU.  The workgroup size stays 512 (8 waves share one
// d2 -> bin table in LDS).  Tuning runs the real histogram(s) on the caller's
// catalogs for each candidate, coordinate-wise (m first, then LDS copies, then
// occupancy), keeps the fastest, and checks that every candidate produced the
// same counts as the default configuration (a failure is an error, not a
// silent choice).  Results are cached on the context per workload signature
// and applied to later calls whose knobs are left at 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pcf_ctx.h"

namespace pcf {

std::string tune_key(pcf_ctx* c, const pcf_job* jobs, int njobs) {
  std::string k;
  char buf[160];
  const pcf_job& j = jobs[0];
  std::snprintf(buf, sizeof(buf), "t%d b%d n%d d%a L%a p%d|", j.type, j.bpc, j.bins,
                (double)j.dmax, (double)j.box, j.precision);
  k += buf;
  for (int i = 0; i < njobs; ++i) {
    Catalog* X = get_cat(c, jobs[i].cat_x);
    Catalog* Y = jobs[i].hist == PCF_XY || jobs[i].hist == PCF_XXY ? get_cat(c, jobs[i].cat_y)
                                                                   : nullptr;
    std::snprintf(buf, sizeof(buf), "%d:%lld:%lld;", jobs[i].hist, X ? (long long)X->n : -1LL,
                  Y ? (long long)Y->n : -1LL);
    k += buf;
  }
  return k;
}

// Apply a cached tuning to jobs whose knobs are all at their defaults.
void apply_tuning(pcf_ctx* c, pcf_job* jobs, int njobs) {
  for (int i = 0; i < njobs; ++i)
    if (jobs[i].cells_per_rmax || jobs[i].lds_copies_log2 || jobs[i].blocks_per_cu ||
        jobs[i].pair_kernel)
      return;
  const std::string key = tune_key(c, jobs, njobs);
  std::lock_guard<std::mutex> lk(c->store->mu);
  auto it = c->store->tuned.find(key);
  if (it == c->store->tuned.end()) return;
  for (int i = 0; i < njobs; ++i) {
    jobs[i].cells_per_rmax = it->second.cells_per_rmax;
    jobs[i].lds_copies_log2 = it->second.lds_copies_log2;
    jobs[i].blocks_per_cu = it->second.blocks_per_cu;
    jobs[i].pair_kernel = it->second.pair_kernel;
  }
}

namespace {

struct Trial {
  int m, lds, per_cu, kernel;
};

// Runs the jobs once with the given knobs; returns PCF_OK and the device time
// of the counting kernel(s), or an error (UNSUPPORTED: the geometry does not
// exist for this m).
int run_trial(pcf_ctx* c, const std::vector<pcf_job>& base, const Trial& t,
              std::vector<uint64_t>* counts, double* ms) {
  std::vector<pcf_job> jobs = base;
  for (auto& j : jobs) {
    j.cells_per_rmax = t.m;
    j.lds_copies_log2 = t.lds;
    j.blocks_per_cu = t.per_cu;
    j.pair_kernel = t.kernel;
  }
  int64_t total = 0;
  for (const auto& j : jobs) total += pcf_hist_size(j.type, j.bins);
  counts->assign((size_t)total, 0);
  int rc;
  if (jobs.size() == 1) rc = pcf_histogram(c, jobs.data(), counts->data(), nullptr);
  else rc = run_pairs_batch(c, jobs.data(), (int)jobs.size(), counts->data(), nullptr);
  if (rc != PCF_OK) return rc;
  *ms = c->last.kernel_ms;
  return PCF_OK;
}

}  // namespace

}  // namespace pcf

extern "C" int pcf_autotune(pcf_ctx* c, const pcf_job* jobs, int32_t njobs, pcf_tuning* out) {
  using namespace pcf;
  if (!c || !jobs || njobs <= 0) return fail(c, PCF_ERR_ARG, "bad autotune arguments");
  HIPCHK(c, hipSetDevice(c->device));
  c->err.clear();
  const int type = jobs[0].type;
  if (type != PCF_2ISO && type != PCF_2ANI)
    return fail(c, PCF_ERR_UNSUPPORTED, "autotuning covers the pair kernels");
  if (jobs[0].precision != PCF_PREC_MIXED)
    return fail(c, PCF_ERR_UNSUPPORTED, "autotuning covers the mixed-precision path");
  bool weighted = false;
  int bins_max = 0;  // the batch's largest bins sizes the LDS layout
  for (int i = 0; i < njobs; ++i) {
    bins_max = std::max(bins_max, jobs[i].bins);
    Catalog* X = get_cat(c, jobs[i].cat_x);
    Catalog* Y = jobs[i].hist == PCF_XY ? get_cat(c, jobs[i].cat_y) : X;
    if (!X || !Y) return fail(c, PCF_ERR_ARG, "unknown catalog id");
    weighted = weighted || !(X->all_w1 && Y->all_w1);
  }
  std::vector<pcf_job> base(jobs, jobs + njobs);
  for (auto& j : base) {
    j.cells_per_rmax = 0;
    j.lds_copies_log2 = 0;
    j.blocks_per_cu = 0;
    j.pair_kernel = PCF_PAIR_AUTO;
    j.rebuild_grid = 0;
  }
  const std::string key = tune_key(c, base.data(), njobs);
  {
    std::lock_guard<std::mutex> lk(c->store->mu);
    c->store->tuned.erase(key);
  }

  // Default configuration: the reference result and time (two runs, the first
  // also builds the grids).
  std::vector<uint64_t> ref, got;
  double t_def = 0, t = 0;
  const Trial def{0, 0, 0, PCF_PAIR_AUTO};
  int rc = run_trial(c, base, def, &ref, &t);
  if (rc != PCF_OK) return rc;
  if ((rc = run_trial(c, base, def, &ref, &t_def)) != PCF_OK) return rc;
  const int m_def = c->last.cells_per_rmax;
  int trials = 2;
  auto timed = [&](const Trial& tr, double* best) -> int {
    // Two runs (grids of a new geometry are built by the first); the faster
    // kernel time counts.  Every run must reproduce the default counts.
    double a = 0, b = 0;
    int r = run_trial(c, base, tr, &got, &a);
    if (r != PCF_OK) return r;
    if (got != ref) return fail(c, PCF_ERR_HIP, "a tuning candidate changed the counts");
    r = run_trial(c, base, tr, &got, &b);
    if (r != PCF_OK) return r;
    if (got != ref) return fail(c, PCF_ERR_HIP, "a tuning candidate changed the counts");
    trials += 2;
    *best = std::min(a, b);
    return PCF_OK;
  };

  // 0. pair kernel: unweighted jobs (and batches) run the Gram form by
  // default; the direct form competes for them.
  const bool gram_ok = !weighted;
  Trial best{m_def, 0, 0, gram_ok ? PCF_PAIR_GRAM : PCF_PAIR_AUTO};
  double t_best = t_def;
  if (gram_ok) {
    double tm;
    const Trial tr{m_def, 0, 0, PCF_PAIR_DIRECT};
    if ((rc = timed(tr, &tm)) != PCF_OK) return rc;
    if (tm < t_best) {
      t_best = tm;
      best = tr;
    }
  }
  // 1. cell edge (for each kernel candidate)
  const int nk = gram_ok ? 2 : 1;
  for (int kk = 0; kk < nk; ++kk) {
    const int kern = gram_ok ? (kk == 0 ? PCF_PAIR_GRAM : PCF_PAIR_DIRECT) : PCF_PAIR_AUTO;
    for (int m = 1; m <= 6; ++m) {
      if (m == m_def) continue;
      double tm;
      rc = timed(Trial{m, 0, 0, kern}, &tm);
      if (rc == PCF_ERR_UNSUPPORTED) continue;
      if (rc != PCF_OK) return rc;
      if (tm < t_best) {
        t_best = tm;
        best = Trial{m, 0, 0, kern};
      }
    }
  }
  // 2. LDS copies of the histogram rows (log2 + 1 in the job)
  {
    const int nflat = type == PCF_2ANI ? bins_max * bins_max : bins_max;
    const int cs_def = pair_shape(nflat, bins_max, weighted, false, 0, 0).cshift;
    for (int cs = std::max(0, cs_def - 3); cs <= std::min(6, cs_def + 1); ++cs) {
      if (cs == cs_def) continue;
      double tm;
      Trial tr{best.m, cs + 1, 0, best.kernel};
      if ((rc = timed(tr, &tm)) != PCF_OK) return rc;
      if (tm < t_best) {
        t_best = tm;
        best = tr;
      }
    }
  }
  // 3. resident workgroups per CU
  {
    const int nflat = type == PCF_2ANI ? bins_max * bins_max : bins_max;
    const int cs = best.lds ? best.lds - 1
                            : pair_shape(nflat, bins_max, weighted, false, 0, 0).cshift;
    const int pc_def = pair_shape(nflat, bins_max, weighted, false, cs + 1, 0).per_cu;
    for (int pc = 1; pc < pc_def; ++pc) {
      double tm;
      Trial tr{best.m, best.lds, pc, best.kernel};
      if ((rc = timed(tr, &tm)) != PCF_OK) return rc;
      if (tm < t_best) {
        t_best = tm;
        best = tr;
      }
    }
  }
  pcf_tuning res{};
  res.cells_per_rmax = best.m;
  res.lds_copies_log2 = best.lds;
  res.blocks_per_cu = best.per_cu;
  res.pair_kernel = best.kernel;
  res.trials = trials;
  res.best_ms = t_best;
  res.default_ms = t_def;
  {
    std::lock_guard<std::mutex> lk(c->store->mu);
    c->store->tuned[key] = res;
  }
  if (out) *out = res;
  return PCF_OK;
}
