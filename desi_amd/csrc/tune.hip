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
#include <cstdio>
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

// ---------------------------------------------------------------------------
// Tuning cache file.  Text, one tuned workload per line after a two-line
// header naming the format and the device it was measured on:
//   pcf-tuning-cache 1
//   device <gcnArchName> <compute units>
//   <workload signature>\t<cells_per_rmax> <lds_copies_log2> <blocks_per_cu>
//       <pair_kernel> <trials> <best_ms> <default_ms>
// The signature is tune_key's (type, periodicity, bins, dmax, box, precision
// and catalog sizes); it holds no tab or newline.  A file from another device
// model is refused: its timings do not carry over.
// ---------------------------------------------------------------------------
namespace pcf {
namespace {

int device_tag(pcf_ctx* c, std::string* tag) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, c->device) != hipSuccess) return -1;
  char buf[320];
  std::snprintf(buf, sizeof(buf), "%s %d", p.gcnArchName, p.multiProcessorCount);
  for (char* q = buf; *q; ++q)
    if (*q == '\t' || *q == '\n') *q = ' ';
  *tag = buf;
  return 0;
}

}  // namespace
}  // namespace pcf

extern "C" int pcf_tune_lookup(pcf_ctx* c, const pcf_job* jobs, int32_t njobs, pcf_tuning* out) {
  using namespace pcf;
  if (!c || !jobs || njobs <= 0) return fail(c, PCF_ERR_ARG, "bad tuning lookup arguments");
  c->err.clear();
  std::vector<pcf_job> base(jobs, jobs + njobs);
  for (auto& j : base) {
    if (!get_cat(c, j.cat_x)) return fail(c, PCF_ERR_ARG, "unknown catalog id");
    j.cells_per_rmax = j.lds_copies_log2 = j.blocks_per_cu = j.pair_kernel = 0;
  }
  const std::string key = tune_key(c, base.data(), njobs);
  std::lock_guard<std::mutex> lk(c->store->mu);
  auto it = c->store->tuned.find(key);
  if (it == c->store->tuned.end()) return 0;
  if (out) *out = it->second;
  return 1;
}

extern "C" int pcf_tune_save(pcf_ctx* c, const char* path) {
  using namespace pcf;
  if (!c || !path || !*path) return fail(c, PCF_ERR_ARG, "bad tuning cache path");
  c->err.clear();
  std::string tag;
  if (device_tag(c, &tag) != 0) return fail(c, PCF_ERR_HIP, "hipGetDeviceProperties failed");
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return fail(c, PCF_ERR_ARG, std::string("cannot write ") + tmp);
  bool ok = std::fprintf(f, "pcf-tuning-cache 1\ndevice %s\n", tag.c_str()) > 0;
  {
    std::lock_guard<std::mutex> lk(c->store->mu);
    for (const auto& kv : c->store->tuned) {
      const pcf_tuning& t = kv.second;
      ok = ok && std::fprintf(f, "%s\t%d %d %d %d %d %.17g %.17g\n", kv.first.c_str(),
                              t.cells_per_rmax, t.lds_copies_log2, t.blocks_per_cu,
                              t.pair_kernel, t.trials, t.best_ms, t.default_ms) > 0;
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(c, PCF_ERR_ARG, std::string("cannot write ") + path);
  }
  return PCF_OK;
}

extern "C" int pcf_tune_load(pcf_ctx* c, const char* path) {
  using namespace pcf;
  if (!c || !path || !*path) return fail(c, PCF_ERR_ARG, "bad tuning cache path");
  c->err.clear();
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(c, PCF_ERR_ARG, std::string("cannot read ") + path);
  std::string tag;
  if (device_tag(c, &tag) != 0) {
    std::fclose(f);
    return fail(c, PCF_ERR_HIP, "hipGetDeviceProperties failed");
  }
  std::vector<std::pair<std::string, pcf_tuning>> got;
  std::string err;
  char line[2048];
  int ln = 0;
  while (err.empty() && std::fgets(line, sizeof(line), f)) {
    ++ln;
    std::string s(line);
    if (s.empty() || s.back() != '\n') {
      err = "line " + std::to_string(ln) + " is too long or unterminated";
      break;
    }
    s.pop_back();
    if (ln == 1) {
      if (s != "pcf-tuning-cache 1") err = "not a tuning cache file (format line)";
      continue;
    }
    if (ln == 2) {
      if (s.rfind("device ", 0) != 0) err = "missing device line";
      else if (s.substr(7) != tag)
        err = "tuning cache was measured on device '" + s.substr(7) + "', this one is '" + tag +
              "'";
      continue;
    }
    const size_t tab = s.find('\t');
    pcf_tuning t{};
    int used = 0;
    if (tab == std::string::npos || tab == 0 ||
        std::sscanf(s.c_str() + tab + 1, "%d %d %d %d %d %lf %lf%n", &t.cells_per_rmax,
                    &t.lds_copies_log2, &t.blocks_per_cu, &t.pair_kernel, &t.trials, &t.best_ms,
                    &t.default_ms, &used) != 7 ||
        s[tab + 1 + used] != '\0' || t.cells_per_rmax < 0 || t.cells_per_rmax > 64 ||
        t.lds_copies_log2 < 0 || t.lds_copies_log2 > 7 || t.blocks_per_cu < 0 ||
        t.blocks_per_cu > 64 || t.pair_kernel < PCF_PAIR_AUTO || t.pair_kernel > PCF_PAIR_GRAM) {
      err = "malformed entry on line " + std::to_string(ln);
      break;
    }
    got.emplace_back(s.substr(0, tab), t);
  }
  std::fclose(f);
  if (err.empty() && ln < 2) err = "truncated tuning cache (no device line)";
  if (!err.empty()) return fail(c, PCF_ERR_ARG, std::string(path) + ": " + err);
  std::lock_guard<std::mutex> lk(c->store->mu);
  for (auto& kv : got) c->store->tuned[kv.first] = kv.second;
  return PCF_OK;
}
