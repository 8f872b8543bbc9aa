// jobs.hip -- the serving engine behind pcf_submit / pcf_poll / pcf_wait.
//
// The reference serves one request per process run: main() reads the
// catalogs, grids them, runs one driver and exits (cuda/main.cu:56-417).  A
// service answering many requests over resident catalogs needs what this file
// adds:
//   * a queue of submitted histogram jobs (tickets), FIFO;
//   * admission against an HBM budget: a job starts when its estimated device
//     memory fits beside the jobs in flight (a lone job always starts);
//   * max_inflight lanes, each a worker thread with its own HIP stream and
//     scratch buffers, sharing the context's resident catalogs and grids (the
//     grid cache pins grids in use and orders cross-stream builds by events);
//   * continuous batching: a lane that takes a pair job also takes the queued
//     pair jobs compatible with it (same type, periodicity, box and
//     precision; bins and dmax may differ, each job keeping its own bin
//     table, range and scale) and runs them as one persistent launch
//     (pcf_histogram_batch's path);
//   * per-ticket host timestamps, so callers can report queueing and latency
//     percentiles.
// Results are the same as pcf_histogram's for the job alone (counts bit-exact,
// weighted sums to 1e-12 relative).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pcf_ctx.h"

namespace pcf {

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Ticket {
  pcf_ticket id = 0;
  pcf_job job{};
  int rc = PCF_OK;
  std::string err;
  bool done = false;
  std::vector<uint64_t> counts;
  std::vector<double> weights;
  pcf_stats st{};
  double t_submit = 0, t_start = 0, t_done = 0;
  int64_t bytes = 0;
};

bool batchable(const pcf_job& a) {
  return (a.type == PCF_2ISO || a.type == PCF_2ANI) && a.precision == PCF_PREC_MIXED &&
         (a.hist == PCF_XX || a.hist == PCF_XY) && !a.rebuild_grid;
}

// (Batching is also off under pcf_set_ref_bugs, whose 2ani DR weights are
// formed per call.)
bool compatible(const pcf_job& a, const pcf_job& b) {
  return batchable(a) && batchable(b) && a.type == b.type && a.bpc == b.bpc &&
         (!a.bpc || a.box == b.box) && a.cells_per_rmax == b.cells_per_rmax &&
         a.lds_copies_log2 == b.lds_copies_log2 && a.blocks_per_cu == b.blocks_per_cu &&
         a.pair_kernel == b.pair_kernel;
}

}  // namespace

struct JobEngine {
  pcf_ctx* parent = nullptr;
  std::vector<pcf_ctx*> lanes;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Ticket*> queue;
  std::unordered_map<pcf_ticket, std::unique_ptr<Ticket>> tickets;
  int64_t budget = 0;
  int64_t inflight_bytes = 0;
  int inflight = 0;
  int max_batch = 1;
  size_t max_queue = 1 << 16;
  pcf_ticket next_id = 1;
  bool stop = false;

  bool head_admissible() const {
    if (queue.empty()) return false;
    return inflight == 0 || inflight_bytes + queue.front()->bytes <= budget;
  }

  void worker(int k) {
    pcf_ctx* lane = lanes[k];
    (void)hipSetDevice(lane->device);
    for (;;) {
      std::vector<Ticket*> take;
      int64_t bytes = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || head_admissible(); });
        if (stop) return;
        Ticket* t = queue.front();
        queue.pop_front();
        take.push_back(t);
        bytes += t->bytes;
        // Continuous batching: compatible pair jobs queued behind it.
        if (max_batch > 1 && batchable(t->job) && !ref_bugs(parent)) {
          for (auto it = queue.begin(); it != queue.end() && (int)take.size() < max_batch;) {
            if (compatible(t->job, (*it)->job) && inflight_bytes + bytes + (*it)->bytes <= budget) {
              bytes += (*it)->bytes;
              take.push_back(*it);
              it = queue.erase(it);
            } else {
              ++it;
            }
          }
        }
        inflight++;
        inflight_bytes += bytes;
      }
      const double t0 = now_ms();
      for (Ticket* t : take) t->t_start = t0;
      run(lane, take);
      const double t1 = now_ms();
      {
        std::lock_guard<std::mutex> lk(mu);
        for (Ticket* t : take) {
          t->t_done = t1;
          t->done = true;
        }
        inflight--;
        inflight_bytes -= bytes;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  void run(pcf_ctx* lane, const std::vector<Ticket*>& take) {
    lane->err.clear();
    for (Ticket* t : take) {
      const int64_t hs = pcf_hist_size(t->job.type, t->job.bins);
      t->counts.assign((size_t)std::max<int64_t>(hs, 0), 0);
      t->weights.assign((size_t)std::max<int64_t>(hs, 0), 0.0);
    }
    if (take.size() == 1) {
      Ticket* t = take[0];
      t->rc = dispatch_histogram(lane, &t->job, t->counts.data(), t->weights.data());
      t->st = lane->last;
      t->st.batched_jobs = 1;
      if (t->rc != PCF_OK) t->err = lane->err;
      return;
    }
    std::vector<pcf_job> jobs;
    std::vector<int64_t> off(1, 0);
    for (Ticket* t : take) {
      jobs.push_back(t->job);
      off.push_back(off.back() + pcf_hist_size(t->job.type, t->job.bins));
    }
    std::vector<uint64_t> c((size_t)off.back());
    std::vector<double> w((size_t)off.back());
    const int rc = run_pairs_batch(lane, jobs.data(), (int)jobs.size(), c.data(), w.data());
    if (rc != PCF_OK) {
      // One job of the batch is refused (its catalog outside the box, say):
      // every job runs alone, so each gets exactly its own result or error.
      for (Ticket* t : take) {
        t->rc = dispatch_histogram(lane, &t->job, t->counts.data(), t->weights.data());
        t->st = lane->last;
        t->st.batched_jobs = 1;
        if (t->rc != PCF_OK) t->err = lane->err;
      }
      return;
    }
    for (size_t k = 0; k < take.size(); ++k) {
      Ticket* t = take[k];
      t->rc = PCF_OK;
      t->st = lane->last;
      t->st.batched_jobs = (int32_t)take.size();
      std::copy(c.begin() + off[k], c.begin() + off[k + 1], t->counts.begin());
      std::copy(w.begin() + off[k], w.begin() + off[k + 1], t->weights.begin());
    }
  }

  void drain() {
    std::unique_lock<std::mutex> lk(mu);
    cv_done.wait(lk, [&] { return queue.empty() && inflight == 0; });
  }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    for (auto& w : workers)
      if (w.joinable()) w.join();
    workers.clear();
    for (pcf_ctx* l : lanes) pcf_destroy(l);
    lanes.clear();
  }
};

void destroy_engine(JobEngine* e) {
  if (!e) return;
  e->drain();
  e->shutdown();
  delete e;
}

void engine_drain(JobEngine* e) {
  if (e) e->drain();
}

namespace {

int start_engine(pcf_ctx* c, int64_t budget, int max_inflight, int max_batch) {
  if (c->lane) return fail(c, PCF_ERR_ARG, "engine lanes cannot serve");
  if (max_inflight < 1) max_inflight = 1;
  if (max_inflight > 16) max_inflight = 16;
  if (budget <= 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    budget = (int64_t)((double)fr * 0.8);
  }
  auto e = new JobEngine();
  e->parent = c;
  e->budget = budget;
  e->max_batch = std::max(1, max_batch);
  for (int k = 0; k < max_inflight; ++k) {
    pcf_ctx* lane = nullptr;
    int rc = create_lane(c, &lane);
    if (rc != PCF_OK) {
      e->shutdown();
      delete e;
      return rc;
    }
    e->lanes.push_back(lane);
  }
  for (int k = 0; k < max_inflight; ++k) e->workers.emplace_back([e, k] { e->worker(k); });
  c->engine = e;
  return PCF_OK;
}

// Device memory a job may need beyond what is already resident: grids of its
// catalogs, the scaled copies and work list of the pair path, the (1,2) pair
// list and histograms of the triplet path.  An estimate for admission only.
int64_t estimate_bytes(pcf_ctx* c, const pcf_job& j) {
  Catalog* X = get_cat(c, j.cat_x);
  const bool cross = j.hist == PCF_XY || j.hist == PCF_XXY;
  Catalog* Y = cross ? get_cat(c, j.cat_y) : nullptr;
  const double nx = X ? (double)X->n : 0, ny = Y ? (double)Y->n : 0;
  double b = 16.0 * (nx + ny) * 2.0;          // grids + scaled copies
  if (j.precision == PCF_PREC_FP64) b += 32.0 * (nx + ny);
  b += 8.0 * (nx + ny) + (1 << 20);           // work list, cell tables, histograms
  if (j.type == PCF_3ISO || j.type == PCF_3ANI) {
    double vol = 1.0;
    for (int a = 0; a < 3; ++a)
      vol *= j.bpc ? (double)j.box
                   : std::max((double)(X ? X->hi[a] - X->lo[a] : 1.0f), (double)j.dmax);
    const double nb = nx * 4.18879 * (double)j.dmax * j.dmax * j.dmax / vol;  // neighbours
    b += 24.0 * nx * nb * (j.type == PCF_3ANI ? 2.0 : 1.0);
    const double bins = j.bins;
    b += 16.0 * (j.type == PCF_3ANI ? bins * bins * bins * bins * bins : bins * bins * bins);
  }
  return (int64_t)b;
}

}  // namespace
}  // namespace pcf

extern "C" {

int pcf_set_hbm_budget(pcf_ctx* c, int64_t bytes, int32_t max_inflight) {
  using namespace pcf;
  if (!c) return PCF_ERR_ARG;
  int mb = 1;
  // Finished tickets not yet collected move to the restarted engine (the old
  // one drains first, so every ticket is finished by then).
  std::unordered_map<pcf_ticket, std::unique_ptr<Ticket>> kept;
  pcf_ticket next_id = 1;
  if (c->engine) {
    mb = c->engine->max_batch;
    c->engine->drain();
    {
      std::lock_guard<std::mutex> lk(c->engine->mu);
      kept.swap(c->engine->tickets);
      next_id = c->engine->next_id;
    }
    destroy_engine(c->engine);
    c->engine = nullptr;
  }
  const int rc = start_engine(c, bytes, max_inflight, mb);
  if (rc == PCF_OK) {
    std::lock_guard<std::mutex> lk(c->engine->mu);
    c->engine->tickets.swap(kept);
    c->engine->next_id = next_id;
  }
  return rc;
}

int pcf_set_max_batch(pcf_ctx* c, int32_t max_batch) {
  using namespace pcf;
  if (!c || max_batch < 1) return fail(c, PCF_ERR_ARG, "max_batch must be >= 1");
  if (!c->engine) {
    int rc = start_engine(c, 0, 4, max_batch);
    if (rc != PCF_OK) return rc;
  }
  std::lock_guard<std::mutex> lk(c->engine->mu);
  c->engine->max_batch = max_batch;
  return PCF_OK;
}

int pcf_submit(pcf_ctx* c, const pcf_job* job, pcf_ticket* ticket) {
  using namespace pcf;
  if (!c || !job || !ticket) return fail(c, PCF_ERR_ARG, "null argument");
  if (c->lane) return fail(c, PCF_ERR_ARG, "engine lanes cannot serve");
  if (job->type < PCF_2ISO || job->type > PCF_3ANI || job->bins <= 0)
    return fail(c, PCF_ERR_ARG, "bad job");
  if (!get_cat(c, job->cat_x) ||
      ((job->hist == PCF_XY || job->hist == PCF_XXY) && !get_cat(c, job->cat_y)))
    return fail(c, PCF_ERR_ARG, "unknown catalog id");
  if (!c->engine) {
    int rc = start_engine(c, 0, 4, 1);
    if (rc != PCF_OK) return rc;
  }
  JobEngine* e = c->engine;
  auto t = std::make_unique<Ticket>();
  t->job = *job;
  t->bytes = estimate_bytes(c, *job);
  t->t_submit = now_ms();
  {
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->queue.size() >= e->max_queue) return fail(c, PCF_ERR_BUSY, "serving queue full");
    t->id = e->next_id++;
    *ticket = t->id;
    e->queue.push_back(t.get());
    e->tickets[t->id] = std::move(t);
  }
  e->cv_work.notify_one();
  return PCF_OK;
}

int pcf_poll(pcf_ctx* c, pcf_ticket ticket, int32_t* done) {
  using namespace pcf;
  if (!c || !done) return PCF_ERR_ARG;
  if (!c->engine) return fail(c, PCF_ERR_ARG, "unknown ticket");
  std::lock_guard<std::mutex> lk(c->engine->mu);
  auto it = c->engine->tickets.find(ticket);
  if (it == c->engine->tickets.end()) return fail(c, PCF_ERR_ARG, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  return PCF_OK;
}

int pcf_wait(pcf_ctx* c, pcf_ticket ticket, uint64_t* counts, double* weights, pcf_stats* st) {
  using namespace pcf;
  if (!c) return PCF_ERR_ARG;
  if (!c->engine) return fail(c, PCF_ERR_ARG, "unknown ticket");
  JobEngine* e = c->engine;
  std::unique_ptr<Ticket> t;
  {
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->tickets.find(ticket);
    if (it == e->tickets.end()) return fail(c, PCF_ERR_ARG, "unknown ticket");
    Ticket* tp = it->second.get();
    e->cv_done.wait(lk, [&] { return tp->done; });
    t = std::move(it->second);
    e->tickets.erase(it);
  }
  if (t->rc != PCF_OK) return fail(c, t->rc, t->err);
  if (counts) std::copy(t->counts.begin(), t->counts.end(), counts);
  if (weights) std::copy(t->weights.begin(), t->weights.end(), weights);
  if (st) {
    *st = t->st;
    st->queue_ms = t->t_start - t->t_submit;
    st->latency_ms = t->t_done - t->t_submit;
  }
  return PCF_OK;
}

}  // extern "C"
