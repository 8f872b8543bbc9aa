// batch.hip -- many pair histograms in one persistent launch, captured in a
// HIP graph (pcf_histogram_batch).
//
// The reference's many-random drivers launch XX and XY once per random file,
// each over its own node grid (cuda/pcf2ani.cuh:90-98, cuda/pcf2iso.cuh:95-105);
// a 64-catalog run is 129 launches with uneven sizes and a device-wide sync.
// Here every catalog of the batch is gridded on one geometry, the work lists of
// all histograms are concatenated and cut into job-pure chunks, and one
// persistent k_pairs_mixed<..., BATCH = 1> grid walks them (workgroups take
// chunks, waves take items, each chunk ends with a flush of the workgroup's LDS
// histogram into its job's rows).  The plan (grids, lists, chunks, job table,
// device buffers) is kept on the context; a repeated call with the same jobs
// over the same resident catalogs replays the captured graph (scaled copies,
// histogram clear, counting kernel) without re-planning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pcf_ctx.h"

namespace pcf {

void make_rules(int type, int bpc, int auto_, int bins, float dmax, float box, BinRules* r);
hipError_t build_batch_worklists(const std::vector<WorkListArgs>& args, DevBuf& counts,
                                 DevBuf& items, std::vector<int>* job_offsets, hipStream_t s);

struct BatchPlan {
  // signature
  std::vector<pcf_job> jobs;
  std::vector<uint64_t> serials;      // catalog serial per job side (x, y)
  std::vector<const void*> grid_pts;  // grid point arrays the plan was built on
  std::vector<DevGrid*> grids;        // and their grids (same order as serials)
  // geometry / rules.  Jobs of one (bins, dmax) share a group: its bin
  // table, exact rules and scale; nflat and bins are the largest of the batch
  // (they size the kernel's LDS layout).
  GridGeom g{};
  int nflat = 0, bins = 0, ani = 0, weighted = 0, cshift = 0, nblocks = 0, m = 0;
  bool gram = false;  // Gram form of the pair kernel (matrix cores)
  struct Group {
    int bins = 0;
    float dmax = 0.f;
    BinRules rules{};
    float inv_h = 1.f, sigma = 1.f, rho2 = 0.f, tflip = 0.f;
    int nt = 0;
    size_t table_off = 0;  // in table_d
  };
  std::vector<Group> groups;
  std::vector<int> job_group;   // group of each job
  std::vector<int> job_nflat;   // histogram size of each job
  std::vector<int64_t> job_row;  // first row of each job in the output
  int64_t total_rows = 0;
  int total_items = 0;
  std::vector<int> job_offsets;
  std::vector<BatchChunk> chunks;
  std::vector<int> mult;  // 2 for auto histograms (unordered pairs counted once)
  // jobs with rebuild_grid: the launch re-runs the device counting sort of
  // every catalog into its resident grid first (same data, same cell counts,
  // so the work lists stay valid; only the order inside a cell may change)
  bool regrid = false;
  std::vector<const float4*> regrid_raw;
  std::vector<int64_t> regrid_n;
  DevBuf regrid_cells, regrid_perm;
  // Regrid of up to kMultiGridMax catalogs of one z sub-slab count in one
  // counting sort (fewer dependent launches in the step).
  bool regrid_joint = false;
  MultiGridArgs regrid_mg{};
  // scaled copies: one per distinct (catalog, group scale)
  std::vector<const float4*> scale_src;
  std::vector<int64_t> scale_n;
  std::vector<int64_t> scale_off;
  std::vector<float> scale_sigma;
  // device memory owned by the plan
  DevBuf counts, items, jobs_d, chunks_d, table_d, rules_d, hist_d, scaled;
  // Two captured graphs: the preparation (re-gridding, scaled copies,
  // histogram clear) and the counting kernel, launched back to back with an
  // event between them, so the kernel's time is measured on its own.
  hipGraph_t graph_prep = nullptr, graph = nullptr;
  hipGraphExec_t exec_prep = nullptr, exec = nullptr;
  hipEvent_t kev = nullptr;
  uint64_t evals_per_run = 0;
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (exec_prep) (void)hipGraphExecDestroy(exec_prep);
    if (graph_prep) (void)hipGraphDestroy(graph_prep);
    if (kev) (void)hipEventDestroy(kev);
    for (DevBuf* b : {&counts, &items, &jobs_d, &chunks_d, &table_d, &rules_d, &hist_d, &scaled,
                      &regrid_cells, &regrid_perm})
      b->release();
  }
};

void destroy_batch_plan(BatchPlan* p) { delete p; }

namespace {

bool same_job(const pcf_job& a, const pcf_job& b) {
  return a.type == b.type && a.hist == b.hist && a.bpc == b.bpc && a.bins == b.bins &&
         a.dmax == b.dmax && a.box == b.box && a.precision == b.precision &&
         a.cat_x == b.cat_x && a.cat_y == b.cat_y && a.cells_per_rmax == b.cells_per_rmax &&
         a.lds_copies_log2 == b.lds_copies_log2 && a.blocks_per_cu == b.blocks_per_cu &&
         a.pair_kernel == b.pair_kernel && (a.rebuild_grid != 0) == (b.rebuild_grid != 0);
}

// Graph bodies.  Part 0 (preparation): with regrid the counting sort of every
// catalog, then the scaled copies and the histogram + counter clear.  Part 1:
// the batched kernel.  prepare: only set the kernel attributes (not allowed
// during capture).
hipError_t enqueue_plan(pcf_ctx* c, BatchPlan* p, size_t njobs, bool prepare, bool regrid,
                        int part) {
  hipError_t e;
  (void)njobs;
  float4* sc = p->scaled.as<float4>();
  const size_t hist_bytes = sizeof(unsigned long long) * (size_t)p->total_rows;
  const size_t w_bytes = sizeof(double) * (size_t)p->total_rows;
  if (part == 0 && !prepare && regrid && p->regrid_joint) {
    if ((e = build_grids_into(p->regrid_mg, p->g, p->grids[0]->zq, p->regrid_cells.p,
                              p->regrid_perm.as<int>(), c->stream)) != hipSuccess)
      return e;
  } else if (part == 0 && !prepare && regrid) {
    for (size_t k = 0; k < p->grids.size(); ++k)
      if ((e = build_grid_into(p->regrid_raw[k], p->regrid_n[k], p->g, p->grids[k]->pts,
                               p->grids[k]->cell_start, p->grids[k]->fine_start,
                               p->grids[k]->zq, p->regrid_cells.p, p->regrid_perm.as<int>(),
                               c->stream)) != hipSuccess)
        return e;
  }
  if (part == 0 && !prepare) {
    for (size_t k = 0; k < p->scale_src.size(); ++k)
      if ((e = launch_scale_points(p->scale_src[k], p->scale_n[k], p->scale_sigma[k],
                                   sc + p->scale_off[k], c->stream)) != hipSuccess)
        return e;
    if ((e = launch_zero(p->hist_d.p, hist_bytes + w_bytes + kBatchTailBytes, c->stream)) !=
        hipSuccess)
      return e;
  }
  unsigned long long* d_stats =
      reinterpret_cast<unsigned long long*>(p->hist_d.as<char>() + hist_bytes + w_bytes);
  // Launch parameters of the first job's group; the kernel switches to each
  // chunk's job parameters from the job table.
  const BatchPlan::Group& g0 = p->groups[p->job_group[0]];
  PairLaunch pl{};
  pl.items = p->items.as<PairItem>();
  pl.nitems = p->total_items;
  // Chunk cursors: one per XCD range, each on a 64-byte line of its own.
  pl.counter = reinterpret_cast<int*>(d_stats + 8);
  pl.table = p->table_d.as<uint8_t>() + g0.table_off;
  pl.nt = g0.nt;
  pl.inv_h = g0.inv_h;
  pl.L = (float)g0.rules.L * g0.sigma;
  pl.bins = p->bins;
  pl.nflat = p->nflat;
  pl.cshift = p->cshift;
  pl.hist = p->hist_d.as<unsigned long long>();
  pl.whist = reinterpret_cast<double*>(p->hist_d.as<char>() + hist_bytes);
  pl.stats = d_stats;
  pl.rules = p->rules_d.as<BinRules>() + p->job_group[0];
  pl.ani = p->ani;
  pl.weighted = p->weighted;
  pl.plain64 = 0;
  pl.nblocks = p->nblocks;
  pl.R = g0.rules.R;
  pl.Lx = g0.rules.L;
  pl.le = g0.rules.le;
  pl.zero = g0.rules.zero;
  pl.batch = 1;
  pl.mfma = p->gram ? 1 : 0;
  pl.rho2 = g0.rho2;
  pl.tflip = g0.tflip;
  pl.geo = p->g;
  pl.sigma = g0.sigma;
  pl.jobs = p->jobs_d.as<BatchJob>();
  pl.chunks = p->chunks_d.as<BatchChunk>();
  pl.nchunks = (int)p->chunks.size();
  pl.prepare_only = prepare ? 1 : 0;
  if (part == 0 || pl.nchunks == 0) return hipSuccess;
  return launch_pairs(pl, c->stream);
}

// Build (or rebuild) the plan of a batch.  Returns PCF_OK or an error code.
int make_plan(pcf_ctx* c, const pcf_job* jobs, int njobs, BatchPlan* p, GridPins& pins) {
  const pcf_job& j0 = jobs[0];
  const bool ani = j0.type == PCF_2ANI;
  const int bpc = j0.bpc ? 1 : 0;
  const float L = j0.box;
  // Distinct catalogs and the joint bounds / occupancy used for the geometry.
  std::vector<pcf_catalog> ids;
  auto add_id = [&](pcf_catalog id) {
    if (std::find(ids.begin(), ids.end(), id) == ids.end()) ids.push_back(id);
  };
  float dmax_max = 0.f;
  int bins_max = 0;
  for (int k = 0; k < njobs; ++k) {
    add_id(jobs[k].cat_x);
    if (jobs[k].hist == PCF_XY) add_id(jobs[k].cat_y);
    dmax_max = std::max(dmax_max, jobs[k].dmax);
    bins_max = std::max(bins_max, jobs[k].bins);
  }
  Catalog U;
  bool any = false;
  double nsum = 0;
  float maxabs = 0;
  bool all_w1 = true;
  for (pcf_catalog id : ids) {
    Catalog* X = get_cat(c, id);
    nsum += (double)X->n;
    maxabs = std::max(maxabs, X->maxabs);
    all_w1 = all_w1 && X->all_w1;
    if (!X->n) continue;
    for (int a = 0; a < 3; ++a) {
      U.lo[a] = any ? std::min(U.lo[a], X->lo[a]) : X->lo[a];
      U.hi[a] = any ? std::max(U.hi[a], X->hi[a]) : X->hi[a];
    }
    any = true;
  }
  U.n = any ? (int64_t)std::llround(nsum / (double)ids.size()) : 0;
  // Gram form (matrix cores) for unweighted mixed-precision batches, on thin
  // z cells; PCF_PAIR_DIRECT in the first job selects the direct form.
  p->gram = all_w1 && j0.pair_kernel != PCF_PAIR_DIRECT;
  // One grid for the batch, cut for its largest range; each job reaches its
  // own number of cells.
  int Kmax[3];
  float reach_max;
  if (choose_geometry(&U, nullptr, dmax_max, bpc, L, j0.cells_per_rmax, &p->g, Kmax, &reach_max,
                      &p->m, p->gram ? kGramZsub : 1) != 0)
    return fail(c, PCF_ERR_UNSUPPORTED, "geometry");
  // Grids (resident, cached per catalog and geometry).
  int rc = PCF_OK;
  size_t ncat = 0;
  {
    std::lock_guard<std::mutex> lk(c->store->mu);
    ncat = c->store->cats.size();
  }
  std::vector<DevGrid*> grid_of(ncat, nullptr);
  for (pcf_catalog id : ids) {
    bool rebuild = false;
    for (int k = 0; k < njobs; ++k)
      if ((jobs[k].cat_x == id || (jobs[k].hist == PCF_XY && jobs[k].cat_y == id)) &&
          jobs[k].rebuild_grid)
        rebuild = true;
    DevGrid* gr = pins.add(get_grid(c, get_cat(c, id), p->g, rebuild, &rc));
    if (!gr) return rc;
    grid_of[id] = gr;
  }
  // Groups: one bin table, rule set and scale per (bins, dmax).  Near-zero
  // table cells are marked ambiguous (the rules of an auto histogram) so one
  // table serves the XX and XY jobs of a group; each job's exact path
  // applies its own zero rule.
  const double A = bpc ? (double)maxabs + (double)L : 0.0;
  const double rho = p->gram ? gram_rho(p->g, U.n) : -1.0;
  p->groups.clear();
  p->job_group.assign(njobs, 0);
  std::vector<uint8_t> tables;
  for (int k = 0; k < njobs; ++k) {
    int gi = -1;
    for (size_t q = 0; q < p->groups.size(); ++q)
      if (p->groups[q].bins == jobs[k].bins && p->groups[q].dmax == jobs[k].dmax) gi = (int)q;
    if (gi < 0) {
      BatchPlan::Group gr;
      gr.bins = jobs[k].bins;
      gr.dmax = jobs[k].dmax;
      make_rules(j0.type, bpc, 1, gr.bins, gr.dmax, L, &gr.rules);
      std::vector<uint8_t> table;
      if (build_table(gr.rules, A, kMixedTableMax, &table, &gr.inv_h, &gr.nt, rho, true,
                      &gr.tflip) != 0)
        return fail(c, PCF_ERR_UNSUPPORTED, "bin table");
      gr.rho2 = p->gram ? (float)(rho * rho / (1.02 * 1.02) * (double)gr.inv_h) : 0.f;
      gr.sigma = std::sqrt(gr.inv_h);
      gr.table_off = tables.size();
      tables.insert(tables.end(), table.begin(), table.end());
      tables.resize((tables.size() + 255) & ~(size_t)255, 0);
      p->groups.push_back(gr);
      gi = (int)p->groups.size() - 1;
    }
    p->job_group[k] = gi;
  }
  p->bins = bins_max;
  p->ani = ani ? 1 : 0;
  p->nflat = ani ? bins_max * bins_max : bins_max;
  p->job_nflat.assign(njobs, 0);
  p->job_row.assign(njobs, 0);
  p->total_rows = 0;
  for (int k = 0; k < njobs; ++k) {
    p->job_nflat[k] = ani ? jobs[k].bins * jobs[k].bins : jobs[k].bins;
    p->job_row[k] = p->total_rows;
    p->total_rows += p->job_nflat[k];
  }
  p->weighted = all_w1 ? 0 : 1;
  const PairShape sh = pair_shape(p->nflat, p->bins, p->weighted != 0, false, j0.lds_copies_log2,
                                  j0.blocks_per_cu);
  p->cshift = sh.cshift;
  p->nblocks = std::max(1, c->ncu * sh.per_cu);

  // Scaled copies of every distinct (catalog, group scale), one buffer.
  p->scale_src.clear();
  p->scale_n.clear();
  p->scale_off.clear();
  p->scale_sigma.clear();
  auto scaled_of = [&](pcf_catalog id, float sigma) -> int64_t {
    for (size_t q = 0; q < p->scale_src.size(); ++q)
      if (p->scale_src[q] == grid_of[id]->pts && p->scale_sigma[q] == sigma)
        return p->scale_off[q];
    const int64_t off = p->scale_src.empty() ? 0 : p->scale_off.back() + p->scale_n.back();
    p->scale_src.push_back(grid_of[id]->pts);
    p->scale_n.push_back(grid_of[id]->n);
    p->scale_off.push_back(off);
    p->scale_sigma.push_back(sigma);
    return off;
  };
  std::vector<int64_t> xs_off(njobs), ys_off(njobs);
  for (int k = 0; k < njobs; ++k) {
    const float sg = p->groups[p->job_group[k]].sigma;
    xs_off[k] = scaled_of(jobs[k].cat_x, sg);
    ys_off[k] = scaled_of(jobs[k].hist == PCF_XX ? jobs[k].cat_x : jobs[k].cat_y, sg);
  }
  const int64_t tot = p->scale_src.empty() ? 0 : p->scale_off.back() + p->scale_n.back();
  HIPCHK(c, p->scaled.ensure(sizeof(float4) * (size_t)std::max<int64_t>(tot, 1)));

  // Work lists of all jobs, concatenated; each job reaches the cells of its
  // own range on the batch's grid.
  double ulo[3], uhi[3];
  for (int a = 0; a < 3; ++a) {
    ulo[a] = any ? U.lo[a] : 0.0;
    uhi[a] = any ? U.hi[a] : 1.0;
  }
  std::vector<WorkListArgs> wa(njobs);
  for (int k = 0; k < njobs; ++k) {
    const bool auto_ = jobs[k].hist == PCF_XX;
    DevGrid* gx = grid_of[jobs[k].cat_x];
    DevGrid* gy = auto_ ? gx : grid_of[jobs[k].cat_y];
    wa[k].x_cell_start = gx->cell_start;
    wa[k].y_cell_start = gy->cell_start;
    wa[k].g = p->g;
    if (jobs[k].dmax == dmax_max) {
      for (int a = 0; a < 3; ++a) wa[k].K[a] = Kmax[a];
      wa[k].reach = reach_max;
    } else {
      const double reach = pair_reach(ulo, uhi, jobs[k].dmax, bpc, L);
      for (int a = 0; a < 3; ++a) {
        wa[k].K[a] = (int)std::ceil(reach / p->g.cs[a]);
        if (bpc && wa[k].K[a] > p->g.P[a]) wa[k].K[a] = p->g.P[a];
      }
      wa[k].reach = (float)(reach * (1.0 + 1e-6));
    }
    wa[k].auto_ = auto_ ? 1 : 0;
    wa[k].cylinder = ani ? 1 : 0;
    wa[k].map = nullptr;
    const bool fine = gx->zq > 1 && gy->zq == gx->zq;
    wa[k].y_fine_start = fine ? gy->fine_start : nullptr;
    wa[k].x_pts = fine ? gx->pts : nullptr;
    wa[k].zq = fine ? gx->zq : 1;
  }
  HIPCHK(c, build_batch_worklists(wa, p->counts, p->items, &p->job_offsets, c->stream));
  p->total_items = p->job_offsets[njobs];

  // Chunks: job-pure runs of items, about 8 per workgroup of the grid.
  const int target = std::max(64, std::min(4096, p->total_items / std::max(1, 8 * p->nblocks)));
  p->chunks.clear();
  for (int k = 0; k < njobs; ++k)
    for (int b = p->job_offsets[k]; b < p->job_offsets[k + 1]; b += target)
      p->chunks.push_back({b, std::min(b + target, p->job_offsets[k + 1]), k});

  // Device tables and rules of the groups, then the job table.
  std::vector<BinRules> grules(p->groups.size());
  for (size_t q = 0; q < p->groups.size(); ++q) grules[q] = p->groups[q].rules;
  HIPCHK(c, p->table_d.ensure(std::max<size_t>(tables.size(), 1)));
  HIPCHK(c, p->rules_d.ensure(sizeof(BinRules) * grules.size()));
  std::vector<BatchJob> jd(njobs);
  p->mult.assign(njobs, 1);
  float4* sc = p->scaled.as<float4>();
  for (int k = 0; k < njobs; ++k) {
    const bool auto_ = jobs[k].hist == PCF_XX;
    const pcf_catalog yid = auto_ ? jobs[k].cat_x : jobs[k].cat_y;
    const BatchPlan::Group& gr = p->groups[p->job_group[k]];
    BinRules r;
    make_rules(j0.type, bpc, auto_ ? 1 : 0, jobs[k].bins, jobs[k].dmax, L, &r);
    jd[k].xs = sc + xs_off[k];
    jd[k].ys = sc + ys_off[k];
    jd[k].xo = grid_of[jobs[k].cat_x]->pts;
    jd[k].yo = grid_of[yid]->pts;
    jd[k].hoff = (int)p->job_row[k];
    jd[k].zero = r.zero;
    jd[k].table = p->table_d.as<uint8_t>() + gr.table_off;
    jd[k].rules = p->rules_d.as<BinRules>() + p->job_group[k];
    jd[k].nt = gr.nt;
    jd[k].bins = gr.bins;
    jd[k].nflat = p->job_nflat[k];
    jd[k].Ls = (float)L * gr.sigma;
    jd[k].tflip = gr.tflip;
    jd[k].rho2 = gr.rho2;
    for (int a = 0; a < 3; ++a) {
      // Scaling by sigma (a power of two) is exact.
      jd[k].glo[a] = p->g.lo[a] * gr.sigma;
      jd[k].gcs[a] = p->g.cs[a] * gr.sigma;
      jd[k].ginv[a] = p->g.inv_cs[a] / gr.sigma;
    }
    p->mult[k] = auto_ ? 2 : 1;
  }
  HIPCHK(c, p->jobs_d.ensure(sizeof(BatchJob) * njobs));
  HIPCHK(c, p->chunks_d.ensure(sizeof(BatchChunk) * std::max<size_t>(1, p->chunks.size())));
  HIPCHK(c, p->hist_d.ensure((sizeof(unsigned long long) + sizeof(double)) *
                                 (size_t)p->total_rows + kBatchTailBytes));
  HIPCHK(c, hipMemcpyAsync(p->jobs_d.p, jd.data(), sizeof(BatchJob) * njobs,
                           hipMemcpyHostToDevice, c->stream));
  if (!p->chunks.empty())
    HIPCHK(c, hipMemcpyAsync(p->chunks_d.p, p->chunks.data(),
                             sizeof(BatchChunk) * p->chunks.size(), hipMemcpyHostToDevice,
                             c->stream));
  if (!tables.empty())
    HIPCHK(c, hipMemcpyAsync(p->table_d.p, tables.data(), tables.size(), hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(p->rules_d.p, grules.data(), sizeof(BinRules) * grules.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors above go out of scope

  // Signature.
  p->jobs.assign(jobs, jobs + njobs);
  p->serials.clear();
  p->grid_pts.clear();
  p->grids.clear();
  p->regrid_raw.clear();
  p->regrid_n.clear();
  int64_t nmax = 1;
  for (pcf_catalog id : ids) {
    Catalog* X = get_cat(c, id);
    p->serials.push_back(X->serial);
    p->grid_pts.push_back(grid_of[id]->pts);
    p->grids.push_back(grid_of[id]);
    p->regrid_raw.push_back(X->raw);
    p->regrid_n.push_back(X->n);
    nmax = std::max(nmax, X->n);
  }
  p->regrid = false;
  for (int k = 0; k < njobs; ++k) p->regrid = p->regrid || jobs[k].rebuild_grid != 0;
  if (p->regrid) {
    // Each grid is rebuilt in its own point order (z sub-slabs): the plan's
    // step windows assume it.
    int zq = 1;
    bool same_zq = true;
    int64_t ntot = 0;
    for (DevGrid* g : p->grids) {
      zq = std::max(zq, g->zq);
      same_zq = same_zq && g->zq == p->grids[0]->zq;
    }
    for (int64_t nk : p->regrid_n) ntot += nk;
    const int m = (int)p->grids.size();
    p->regrid_joint = m >= 2 && m <= kMultiGridMax && same_zq &&
                      (int64_t)m * p->g.ncell() * zq + 1 <= (int64_t)INT32_MAX &&
                      ntot <= (int64_t)INT32_MAX;
    if (p->regrid_joint) {
      MultiGridArgs& mg = p->regrid_mg;
      mg = MultiGridArgs{};
      mg.m = m;
      mg.off[0] = 0;
      for (int k = 0; k < m; ++k) {
        mg.raw[k] = p->regrid_raw[k];
        mg.pts[k] = p->grids[k]->pts;
        mg.cell_start[k] = p->grids[k]->cell_start;
        mg.fine_start[k] = p->grids[k]->fine_start;
        mg.off[k + 1] = mg.off[k] + p->regrid_n[k];
      }
      HIPCHK(c, p->regrid_cells.ensure(grid_cells_tmp_bytes((int64_t)m * p->g.ncell() * zq)));
      HIPCHK(c, p->regrid_perm.ensure(2 * sizeof(int) * (size_t)std::max<int64_t>(ntot, 1)));
    } else {
      HIPCHK(c, p->regrid_cells.ensure(grid_cells_tmp_bytes(p->g.ncell() * zq)));
      HIPCHK(c, p->regrid_perm.ensure(2 * sizeof(int) * (size_t)nmax));
    }
  }

  // Capture the run into a graph (serving lanes run their batches directly:
  // their job sets rarely repeat).
  for (hipGraphExec_t* x : {&p->exec, &p->exec_prep})
    if (*x) (void)hipGraphExecDestroy(*x);
  for (hipGraph_t* g : {&p->graph, &p->graph_prep})
    if (*g) (void)hipGraphDestroy(*g);
  p->exec = p->exec_prep = nullptr;
  p->graph = p->graph_prep = nullptr;
  if (!p->kev) HIPCHK(c, hipEventCreate(&p->kev));
  HIPCHK(c, enqueue_plan(c, p, (size_t)njobs, true, false, 1));
  if (c->lane) return PCF_OK;
  for (int part = 0; part < 2; ++part) {
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_plan(c, p, (size_t)njobs, false, p->regrid, part);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      return hipfail(c, e != hipSuccess ? e : e2, "batch graph capture");
    }
    (part == 0 ? p->graph_prep : p->graph) = graph;
    HIPCHK(c, hipGraphInstantiate(part == 0 ? &p->exec_prep : &p->exec, graph, nullptr, nullptr,
                                  0));
  }
  return PCF_OK;
}

// The plan still describes these jobs over the same resident data.
// The plan still describes these jobs over the same resident data; if so its
// grids are pinned (into `pins`) for this launch.
bool plan_valid(pcf_ctx* c, const BatchPlan* p, const pcf_job* jobs, int njobs, GridPins& pins) {
  if (!p || !p->exec || (int)p->jobs.size() != njobs) return false;
  for (int k = 0; k < njobs; ++k) {
    if (!same_job(p->jobs[k], jobs[k])) return false;
  }
  std::vector<pcf_catalog> ids;
  for (int k = 0; k < njobs; ++k) {
    for (pcf_catalog id : {jobs[k].cat_x, jobs[k].hist == PCF_XY ? jobs[k].cat_y : jobs[k].cat_x})
      if (std::find(ids.begin(), ids.end(), id) == ids.end()) ids.push_back(id);
  }
  if (ids.size() != p->serials.size()) return false;
  std::lock_guard<std::mutex> lk(c->store->mu);
  auto& cats = c->store->cats;
  for (size_t i = 0; i < ids.size(); ++i) {
    if (ids[i] < 0 || ids[i] >= (pcf_catalog)cats.size() || !cats[ids[i]]) return false;
    Catalog* X = cats[ids[i]].get();
    if (X->serial != p->serials[i]) return false;
    bool found = false;
    for (auto& gr : X->grids)
      if (gr.get() == p->grids[i] && gr->valid && gr->g.same(p->g) && gr->pts == p->grid_pts[i])
        found = true;
    if (!found) return false;
  }
  for (DevGrid* g : p->grids) {
    g->users++;
    pins.g.push_back(g);
    (void)hipStreamWaitEvent(c->stream, g->ready, 0);
  }
  return true;
}

}  // namespace

int run_pairs_batch(pcf_ctx* c, const pcf_job* jobs_in, int njobs, uint64_t* counts,
                    double* weights) {
  if (njobs <= 0 || !jobs_in) return fail(c, PCF_ERR_ARG, "empty batch");
  std::vector<pcf_job> tuned(jobs_in, jobs_in + njobs);
  const pcf_job* jobs = tuned.data();
  const pcf_job& j0 = jobs[0];
  if (j0.type != PCF_2ISO && j0.type != PCF_2ANI)
    return fail(c, PCF_ERR_UNSUPPORTED, "batched launches are for pair histograms");
  for (int k = 0; k < njobs; ++k) {
    const pcf_job& j = jobs[k];
    if (ref_bugs_alias_needed(c, &j))
      return fail(c, PCF_ERR_UNSUPPORTED,
                  "reference 2ani DR weights (pcf_set_ref_bugs) take pcf_histogram calls");
    if (j.type != j0.type || j.bpc != j0.bpc || (j0.bpc && j.box != j0.box) ||
        j.precision != j0.precision || j.cells_per_rmax != j0.cells_per_rmax ||
        j.lds_copies_log2 != j0.lds_copies_log2 || j.blocks_per_cu != j0.blocks_per_cu)
      return fail(c, PCF_ERR_ARG, "jobs of a batch must share type, box and precision");
    if (j.hist != PCF_XX && j.hist != PCF_XY) return fail(c, PCF_ERR_ARG, "pair histogram kind");
    if (!get_cat(c, j.cat_x) || (j.hist == PCF_XY && !get_cat(c, j.cat_y)))
      return fail(c, PCF_ERR_ARG, "unknown catalog id");
    if (j.bins <= 0 || pcf_hist_size(j.type, j.bins) > kMaxFlatBins)
      return fail(c, PCF_ERR_ARG, "bins out of range");
    if (!(j.dmax > 0) || !std::isfinite(j.dmax)) return fail(c, PCF_ERR_ARG, "dmax must be > 0");
  }
  apply_tuning(c, tuned.data(), njobs);
  // Each job's output: pcf_hist_size values, one job after the other.
  std::vector<int64_t> out_off(njobs + 1, 0);
  for (int k = 0; k < njobs; ++k) out_off[k + 1] = out_off[k] + pcf_hist_size(j0.type, jobs[k].bins);
  // The plain FP64 path and out-of-range arguments go through the single-job
  // driver (same kernels, one launch per job).
  // Histograms the mixed kernel does not serve (bin counts past its table or
  // LDS) likewise.
  bool single = j0.precision == PCF_PREC_FP64;
  bool any_weighted = false;
  for (int k = 0; k < njobs && !single; ++k)
    for (pcf_catalog id : {jobs[k].cat_x, jobs[k].hist == PCF_XY ? jobs[k].cat_y : jobs[k].cat_x}) {
      const Catalog* X = get_cat(c, id);
      if (X->n >= kRecheckMaxPoints) single = true;
      if (!X->all_w1) any_weighted = true;
    }
  for (int k = 0; k < njobs && !single; ++k)
    if (!pair_mixed_fits(jobs[k].bins, j0.type == PCF_2ANI, any_weighted)) single = true;
  if (!single) {
    // Argument checks of the single-job driver (range, box) on each job.
    if (j0.bpc) {
      if (!(j0.box > 0)) return fail(c, PCF_ERR_ARG, "box side required for periodic boundaries");
      for (int k = 0; k < njobs; ++k)
        if (!(jobs[k].dmax < j0.box))
          return fail(c, PCF_ERR_UNSUPPORTED, "periodic boundaries need dmax < box side");
      for (int k = 0; k < njobs; ++k)
        for (pcf_catalog id : {jobs[k].cat_x, jobs[k].hist == PCF_XY ? jobs[k].cat_y : jobs[k].cat_x}) {
          Catalog* X = get_cat(c, id);
          for (int a = 0; a < 3; ++a)
            if (X->n && (X->lo[a] < 0.f || !(X->hi[a] < j0.box)))
              return fail(c, PCF_ERR_RANGE, "coordinates outside [0, box) with periodic boundaries");
        }
    }
  }
  if (single) {
    pcf_stats agg{};
    for (int k = 0; k < njobs; ++k) {
      int rc = pcf_histogram(c, &jobs[k], counts ? counts + out_off[k] : nullptr,
                             weights ? weights + out_off[k] : nullptr);
      if (rc != PCF_OK) return rc;
      agg.evaluations += c->last.evaluations;
      agg.rechecks += c->last.rechecks;
      agg.work_items += c->last.work_items;
      agg.kernel_ms += c->last.kernel_ms;
      agg.total_ms += c->last.total_ms;
      agg.cells_per_rmax = c->last.cells_per_rmax;
    }
    c->last = agg;
    return PCF_OK;
  }

  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  GridPins pins(c);
  const bool replay = plan_valid(c, c->batch, jobs, njobs, pins);
  if (!replay) {
    if (!c->batch) c->batch = new BatchPlan();
    int rc = make_plan(c, jobs, njobs, c->batch, pins);
    if (rc != PCF_OK) {
      delete c->batch;
      c->batch = nullptr;
      return rc;
    }
  }
  BatchPlan* p = c->batch;
  // Re-gridding rewrites resident grids in place: only while no other job
  // (a serving lane) holds them; otherwise the grids, whose contents would be
  // the same, are used as they are.
  bool shared = false;
  if (p->regrid) {
    std::lock_guard<std::mutex> lk(c->store->mu);
    for (DevGrid* g : p->grids) shared = shared || g->users > 1;
    // The launch rewrites the grids' point order.
    if (!shared)
      for (DevGrid* g : p->grids) g->build_id = next_grid_build_id();
  }
  // ev[1] .. kev: preparation; kev .. ev[2]: the counting kernel.
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  if (p->exec && !(p->regrid && shared)) {
    HIPCHK(c, hipGraphLaunch(p->exec_prep, c->stream));
    HIPCHK(c, hipEventRecord(p->kev, c->stream));
    HIPCHK(c, hipGraphLaunch(p->exec, c->stream));
  } else {
    HIPCHK(c, enqueue_plan(c, p, (size_t)njobs, false, p->regrid && !shared, 0));
    HIPCHK(c, hipEventRecord(p->kev, c->stream));
    HIPCHK(c, enqueue_plan(c, p, (size_t)njobs, false, false, 1));
  }
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  const size_t nf = (size_t)p->total_rows;
  std::vector<unsigned long long> hc(nf + 8);
  std::vector<double> hw(nf);
  const size_t hist_bytes = sizeof(unsigned long long) * nf;
  HIPCHK(c, hipMemcpyAsync(hc.data(), p->hist_d.p, hist_bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hw.data(), p->hist_d.as<char>() + hist_bytes, sizeof(double) * nf,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hc.data() + nf, p->hist_d.as<char>() + hist_bytes + sizeof(double) * nf,
                           4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < njobs; ++k) {
    const uint64_t mult = (uint64_t)p->mult[k];
    for (int b = 0; b < p->job_nflat[k]; ++b) {
      const size_t q = (size_t)p->job_row[k] + b;
      if (counts) counts[out_off[k] + b] = (uint64_t)hc[q] * mult;
      if (weights)
        weights[out_off[k] + b] =
            p->weighted ? hw[q] * (double)mult : (double)((uint64_t)hc[q] * mult);
    }
  }
  pcf_stats st{};
  st.evaluations = hc[nf];
  st.rechecks = hc[nf + 1];
  st.work_items = (uint64_t)p->total_items;
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, p->kev, c->ev[2]));
  st.kernel_ms = ms;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
  st.total_ms = ms;
  st.cells_per_rmax = p->m;
  st.graph_replay = replay ? 1 : 0;
  st.chunks = (int32_t)p->chunks.size();
  c->last = st;
  return PCF_OK;
}

}  // namespace pcf

extern "C" int pcf_histogram_batch(pcf_ctx* c, const pcf_job* jobs, int32_t njobs,
                                   uint64_t* counts, double* weights) {
  using namespace pcf;
  if (!c) return PCF_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  c->err.clear();
  return run_pairs_batch(c, jobs, njobs, counts, weights);
}
