/* pcf_abi.h -- C ABI of libpcf_hip.so, the MI355X (gfx950) pair/triplet
 * histogram engine behind the PCF command line.
 *
 * The reference (jaarpa/DESI, "3PCFGPU") has no library API: its boundary is
 * the PCF.out command line plus the host functions that main() calls after it
 * has built the node grid,
 *
 *   pcf_2iso, pcf_2iso_BPC, pcf_2iso_BPCanalytic   cuda/pcf2iso.cuh:11, cuda/pcf2isoBPC.cuh:11,166
 *   pcf_2ani, pcf_2aniBPC                          cuda/pcf2ani.cuh:5,  cuda/pcf2aniBPC.cuh:5
 *   pcf_3iso, pcf_3isoBPC, pcf_3isoBPC_analytic    cuda/pcf3iso.cuh:10, cuda/pcf3isoBPC.cuh:10,177
 *   pcf_3ani, pcf_3aniBPC                          cuda/pcf3ani.cuh:10, cuda/pcf3aniBPC.cuh:10
 *
 * each of which allocates device histograms, launches the XX/XY (XXX/XXY)
 * kernels, copies the histograms back and writes the files.  The grid they
 * receive is built on the host by make_nodos (cuda/create_grid.cuh:104-138)
 * and copied to the device in cuda/main.cu:320-329.
 *
 * This ABI cuts at the same place but splits "compute" from "write files":
 *   - pcf_catalog_upload replaces open_files' output hand-over plus the
 *     host grid build and H2D copy (cuda/main.cu:148-329); the grid is built on
 *     the device, per job geometry, and stays resident;
 *   - pcf_histogram computes ONE histogram (XX, XY, XXX or XXY) of one of the
 *     reference's types into caller-owned host arrays; the drivers in the PCF
 *     command line compose these into the reference's DD/RR/DR and
 *     DDD/DDR/DRR/RRR outputs and write the same files;
 *   - pcf_analytic_rr is RR2iso_BPC_analytic (cuda/kernels/2isoBPC_k.cuh:559)
 *     with the host constants of cuda/pcf2isoBPC.cuh:199-202;
 *   - pcf_submit / pcf_wait are the asynchronous form used by the serving loop.
 *
 * Conventions: plain C types, host pointers for results, int status codes
 * (PCF_OK == 0), never exit().  A context is not thread-safe; use one per
 * host thread or lock externally.  Results are deterministic for counts
 * (uint64, bit-exact to the FP64 oracle); weighted sums are FP64 with
 * a summation order that may vary between runs (relative error ~1e-15).
 */
#ifndef PCF_ABI_H
#define PCF_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PCF_OK 0
#define PCF_ERR_ARG -1      /* invalid argument (sizes, ids, dmax, bins ...)   */
#define PCF_ERR_HIP -2      /* HIP runtime error (message in pcf_last_error)    */
#define PCF_ERR_NOMEM -3    /* device allocation failed or over the HBM budget  */
#define PCF_ERR_RANGE -4    /* coordinates outside the periodic box, non-finite */
#define PCF_ERR_UNSUPPORTED -5 /* combination the engine does not implement    */
#define PCF_ERR_BUSY -6     /* serving queue full (retry after pcf_wait)        */

/* ---- histogram types (argv[1] of the reference, cuda/main.cu:59) ------ */
#define PCF_2ISO 0
#define PCF_2ANI 1
#define PCF_3ISO 2
#define PCF_3ANI 3

/* ---- which histogram ------------------------------------------------- */
#define PCF_XX 0   /* DD or RR      (XX2iso, cuda/kernels/2iso_k.cuh:8)   */
#define PCF_XY 1   /* DR            (XY2iso, cuda/kernels/2iso_k.cuh:63)  */
#define PCF_XXX 2  /* DDD or RRR    (XXX3iso, cuda/kernels/3iso_k.cuh:8)  */
#define PCF_XXY 3  /* DDR or DRR    (XXY3iso, cuda/kernels/3iso_k.cuh:94) */

/* ---- arithmetic ------------------------------------------------------ */
#define PCF_PREC_MIXED 0 /* FP32 distances, FP64 re-evaluation near bin edges */
#define PCF_PREC_FP64 1  /* plain FP64 distances and sqrt for every pair      */

typedef struct pcf_ctx pcf_ctx;
typedef int32_t pcf_catalog;

/* One histogram request.  Field meanings follow the reference flags:
 * bins = -b, dmax = -d, bpc = -bpc, box = -s (or the inferred box side),
 * partitions = -p (accepted for drop-in compatibility; the engine picks its
 * own cell size, which does not change any result). */
typedef struct pcf_job {
  int32_t type;        /* PCF_2ISO .. PCF_3ANI                            */
  int32_t hist;        /* PCF_XX .. PCF_XXY                               */
  int32_t bpc;         /* periodic box                                    */
  int32_t bins;        /* bins per histogram dimension                    */
  float dmax;          /* maximum separation                              */
  float box;           /* box side (required when bpc)                    */
  int32_t precision;   /* PCF_PREC_MIXED or PCF_PREC_FP64                 */
  pcf_catalog cat_x;   /* X: the auto catalog, or D of a cross histogram  */
  pcf_catalog cat_y;   /* Y: R of a cross histogram; -1 for XX / XXX      */
  int32_t partitions;  /* -p; informational, except under pcf_set_ref_bugs */
                       /* (the reference's node grid, partitions per side) */
  int32_t cells_per_rmax; /* cell edge = rmax / cells_per_rmax; 0 = the    */
                          /* autotuned value (pcf_autotune) or a default  */
  int32_t rebuild_grid;   /* 1: re-grid the catalogs even if resident      */
  int32_t lds_copies_log2; /* LDS histogram copies per row, log2 + 1; 0 = auto */
  int32_t blocks_per_cu;  /* resident workgroups per CU; 0 = auto          */
  int32_t pair_kernel;    /* mixed-precision pair path: PCF_PAIR_AUTO,      */
                          /* PCF_PAIR_DIRECT or PCF_PAIR_GRAM               */
  int32_t reserved;
} pcf_job;

/* ---- pair kernel variants (pcf_job.pair_kernel) ---------------------- *
 * Both give the same counts; they differ in where d2 is formed.  DIRECT:
 * per lane, (xi - xj)^2 summed in FP32 on the vector units.  GRAM: 16 x 16
 * pairs at a time on the matrix cores as |xi|^2 + |xj|^2 - 2 xi.xj in
 * tile-local coordinates (2ani: perpendicular and line-of-sight parts as two
 * products; unweighted jobs only, weighted ones run DIRECT).  AUTO is GRAM
 * where it applies.  In pcf_histogram_batch the first job's value applies. */
#define PCF_PAIR_AUTO 0
#define PCF_PAIR_DIRECT 1
#define PCF_PAIR_GRAM 2

/* Per-call statistics of the last completed histogram. */
typedef struct pcf_stats {
  uint64_t evaluations;  /* candidate pairs (or triplets) evaluated       */
  uint64_t rechecks;     /* evaluations re-done in FP64 near a bin edge   */
  uint64_t work_items;   /* entries of the flattened work list            */
  double kernel_ms;      /* device time of the counting kernel(s)         */
  double total_ms;       /* device time of grid build + counting + merge  */
  int32_t cells_per_rmax;
  int32_t graph_replay;  /* pcf_histogram_batch: 1 if the captured graph was replayed */
  int32_t chunks;        /* pcf_histogram_batch: work chunks of the launch  */
  int32_t pairs_listed;  /* triplets: entries of the (1,2) pair list        */
  int32_t batched_jobs;  /* pcf_wait: jobs that shared this job's launch    */
  int32_t direct_items;  /* Gram pair path: work items whose tile was too  */
                         /* large for its error bound (run on the direct path) */
  double queue_ms;       /* pcf_wait: submit -> start (host clock)          */
  double latency_ms;     /* pcf_wait: submit -> done (host clock)           */
} pcf_stats;

/* Sizes of a histogram: bins (2iso), bins^2 (2ani), bins^3 (3iso),
 * bins^5 (3ani). */
int64_t pcf_hist_size(int32_t type, int32_t bins);

/* Context on HIP device `device`.  Replaces the CUDA context set-up implied by
 * cuda/main.cu (no explicit call there). */
int pcf_create(pcf_ctx** out, int32_t device);
void pcf_destroy(pcf_ctx* ctx);
/* Message of the last failing call on this context ("" if none). */
const char* pcf_last_error(const pcf_ctx* ctx);
/* Thread-independent message for a failed pcf_create (ctx unavailable). */
const char* pcf_create_error(void);

/* Upload n points, 4 floats each (x, y, z, w; the 4-column catalog format,
 * readme.md:36-40), from host memory into a device-resident catalog.
 * Replaces the hand-over of open_files (cuda/create_grid.cuh:35) to the grid
 * build and the copy to the device (cuda/main.cu:320-329). */
int pcf_catalog_upload(pcf_ctx* ctx, const float* xyzw, int64_t n, pcf_catalog* id);
/* Same, from device memory already in HBM (copied device-to-device). */
int pcf_catalog_upload_device(pcf_ctx* ctx, const void* d_xyzw, int64_t n, pcf_catalog* id);
int pcf_catalog_release(pcf_ctx* ctx, pcf_catalog id);
/* Number of points of a catalog, or -1. */
int64_t pcf_catalog_size(pcf_ctx* ctx, pcf_catalog id);

/* Compute one histogram.  counts[pcf_hist_size] receives the number of
 * ordered pairs (triplets) per bin, weights[...] the sum of the products of
 * the FP32 weights formed in FP64.  Either output may be NULL.  The layout is
 * the reference GPU layout (cuda/kernels/ *_k.cuh files).  Blocks until done.
 * Replaces the kernel launches + D2H copies of one histogram in the
 * reference drivers (e.g. cuda/pcf2iso.cuh:95-124). */
int pcf_histogram(pcf_ctx* ctx, const pcf_job* job, uint64_t* counts, double* weights);

/* Several pair histograms in one persistent launch.  Replaces the per-random
 * loops of the reference's many-random drivers, which launch XX and XY once
 * per random file (cuda/pcf2iso.cuh:96-104, cuda/pcf2ani.cuh:90-98).  All jobs
 * share type (PCF_2ISO or PCF_2ANI), bpc, box, precision and the tuning
 * knobs; hist is PCF_XX or PCF_XY; the catalogs may differ in size, and bins
 * and dmax may differ between jobs (each keeps its own bin table and range;
 * the batch's grid is cut for its largest dmax).  counts / weights receive
 * the njobs histograms one after the other, pcf_hist_size(type, bins of that
 * job) values each (either may be NULL), each equal to what pcf_histogram
 * returns for that job (counts bit-exact, weights to 1e-12 relative).  The plan and its
 * HIP graph stay on the context: repeating the same batch over the same
 * resident catalogs replays the graph (pcf_last_stats: graph_replay, chunks).
 * A batch whose jobs set rebuild_grid re-runs the device grid build (counting
 * sort) of its catalogs at the start of every launch, replays included.
 * kernel_ms is the counting kernel alone, total_ms the whole call. */
int pcf_histogram_batch(pcf_ctx* ctx, const pcf_job* jobs, int32_t njobs, uint64_t* counts,
                        double* weights);

/* pcf_histogram for pair jobs (PCF_2ISO / PCF_2ANI, PCF_XX / PCF_XY) that
 * allocates no device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * hipMalloc).  Inputs are resident catalogs and outputs the caller's host
 * arrays, as for pcf_histogram, with the same results; every device buffer of
 * the call (grids, scaled copies, work list, bin table, device histogram) is
 * carved from one static device array of PCF_NOALLOC_SCRATCH_BYTES per GPU,
 * held by each GPU the library runs on until the process ends.  A job whose
 * buffers exceed it fails with PCF_ERR_NOMEM and a message giving both sizes
 * (about 40 bytes per catalog point plus the work list: two catalogs of about
 * 1.5 million points).  Calls on one GPU share its array and would race, so
 * the library serialises them per device (a lock held for the whole call).
 * Grids are not kept between calls. */
#define PCF_NOALLOC_SCRATCH_BYTES (128LL * 1024 * 1024)
int pcf_histogram_noalloc(pcf_ctx* ctx, const pcf_job* job, uint64_t* counts, double* weights);

/* Autotuner of the pair kernels (no counterpart in the reference, whose
 * launch shape is fixed: 8x8 threads per block, one thread per node pair,
 * cuda/pcf2iso.cuh:74-80, and whose -p partitions flag sets the node size,
 * cuda/main.cu:86).  Runs the given job, or batch of pair jobs (njobs > 1, as
 * for pcf_histogram_batch), on its own catalogs under candidate pair kernels
 * (unweighted jobs: direct or Gram), cell sizes, LDS histogram layouts and
 * occupancies, verifies that every candidate gives
 * the default configuration's counts, and caches the fastest on the context:
 * later pcf_histogram / pcf_histogram_batch calls on the same workload (same
 * type, bins, dmax, box, precision and catalog sizes) whose knobs are 0 use it.
 * out (may be NULL) receives the choice. */
typedef struct pcf_tuning {
  int32_t cells_per_rmax;
  int32_t lds_copies_log2; /* log2 + 1 of the LDS copies; 0 = default        */
  int32_t blocks_per_cu;   /* 0 = default                                    */
  int32_t trials;          /* histogram runs spent                           */
  double best_ms;          /* counting-kernel time of the choice             */
  double default_ms;       /* same, default configuration                    */
  int32_t pair_kernel;     /* PCF_PAIR_DIRECT or PCF_PAIR_GRAM (unweighted   */
                           /* jobs); PCF_PAIR_AUTO otherwise                 */
  int32_t reserved;
} pcf_tuning;
int pcf_autotune(pcf_ctx* ctx, const pcf_job* jobs, int32_t njobs, pcf_tuning* out);
/* The autotuner's cache as a file (no counterpart in the reference), so a
 * tuning measured once is reused by later processes.  pcf_tune_save writes
 * every tuned workload of the context (text: a format line, the device model
 * and compute-unit count, then one line per workload signature with its
 * knobs and timings); pcf_tune_load merges a file's entries into the context
 * and refuses (PCF_ERR_ARG) a malformed file or one measured on another device
 * model.  pcf_tune_lookup returns 1 and the tuning (out may be NULL) when the
 * workload of jobs is cached, 0 when it is not, or a negative status. */
int pcf_tune_save(pcf_ctx* ctx, const char* path);
int pcf_tune_load(pcf_ctx* ctx, const char* path);
int pcf_tune_lookup(pcf_ctx* ctx, const pcf_job* jobs, int32_t njobs, pcf_tuning* out);

/* Reproduce reference defects (on != 0; the PCF command line's --ref-bugs),
 * for drop-in comparison with result files the reference wrote.  No
 * counterpart call in the reference, whose GPU code has these defects:
 *   - weighted triplet histograms take the cumulative weight of the triplet
 *     kernels (cuda/kernels/3iso_k.cuh:79,168, 3isoBPC_k.cuh:88,179,
 *     3ani_k.cuh:94,197, 3aniBPC_k.cuh:98): v = w1*w2 (FP32) per (point 1,
 *     point 2), multiplied by w3 of each third point deposited before it in its
 *     node, in the node's point order, the node grid being pcf_job.partitions
 *     (default 35) nodes per side of pcf_job.box (needed, also without -bpc);
 *   - 2ani DR (no periodic boundaries) weights each pair with the data weight
 *     at the random point's slot of the ordered arrays (cuda/kernels/2ani_k.cuh:128);
 *     refused when the random catalog is longer than the data catalog (the
 *     reference reads past its data array), and by pcf_histogram_batch and
 *     pcf_histogram_noalloc;
 *   - the analytic normalisations use 32-bit int np*np (cuda/pcf2isoBPC.cuh:201,
 *     cuda/pcf3isoBPC.cuh:236) and 3i^2+3i+1 (cuda/kernels/2isoBPC_k.cuh:572).
 * Counts are unchanged.  Applies to the context and its serving lanes, whose
 * continuous batching is then off. */
int pcf_set_ref_bugs(pcf_ctx* ctx, int32_t on);
/* Statistics of the last pcf_histogram / pcf_wait on this context. */
int pcf_last_stats(const pcf_ctx* ctx, pcf_stats* out);

/* Analytic RR of the periodic 2iso case with np points in a box of side box
 * (cuda/pcf2isoBPC.cuh:199-202 + cuda/kernels/2isoBPC_k.cuh:559-575),
 * computed on the device in FP64 with 64-bit integer arithmetic. */
int pcf_analytic_rr(pcf_ctx* ctx, int32_t bins, int64_t np, float dmax, float box, double* rr);

/* Analytic random triplets of the periodic 3iso case (PCF 3iso -bpc -a):
 * pcf_3isoBPC_analytic (cuda/pcf3isoBPC.cuh:177-375) minus its DDD, which is
 * pcf_histogram(PCF_3ISO, PCF_XXX, bpc = 1).  From the data catalog's fine
 * periodic pair histograms (DD2iso_forBPC, cuda/kernels/3isoBPC_k.cuh:1694)
 * and their analytic randoms it forms ff_av / ff_av_ref (make_ff_av,
 * make_ff_av_ref, :1984-2019) and per bin cube the analytic RRR and DDR
 * (make_histo_analitic, :2021-2119); the reference writes DDR to both its
 * DDR and DRR files.  np <= 0 means the catalog size; ptt and bn_ref <= 0 mean
 * the reference's 100 and 200.  ddr and rrr receive bins^3 values in the
 * [i][j][k] layout; ff_av (bins) and ff_av_ref (bins*bn_ref) may be NULL.
 * FP64 distances and 64-bit bin arithmetic (the reference's FP32 distances
 * and 32-bit 3i^2+3i+1 are not reproduced). */
int pcf_analytic_3iso(pcf_ctx* ctx, pcf_catalog data, int32_t bins, float dmax, float box,
                      int64_t np, int32_t ptt, int32_t bn_ref, double* ddr, double* rrr,
                      double* ff_av, double* ff_av_ref);

/* Correlation-function estimator from finished histograms, on the device
 * (no counterpart in the reference's C++/CUDA programs: it restates the
 * analysis scripts cpp/2PCFisotropic/normal/graph_2DPF.py:8-14 and
 * cpp/3PCFisotropic/normal/graph_3PCF.py:5-6,21-22).  n values per histogram.
 *   order 2: h0 = DD, h1 = DR, h2 = RR (h3 unused);
 *            xi = (DD - 2 DR + RR) / RR, every zero of DD, DR, RR taken as 1e-11;
 *   order 3: h0 = DDD, h1 = DDR, h2 = DRR, h3 = RRR;
 *            xi = (DDD - 3 DDR + 3 DRR - RRR) / RRR, NaN or infinite -> 0.
 * flags & PCF_EST_NORMALIZE divides each histogram by its number of ordered
 * pairs (triplets) of distinct points first (n_data, n_rand: catalog sizes);
 * without it the raw counts are used, as the scripts do. */
#define PCF_EST_NORMALIZE 1
int pcf_estimator(pcf_ctx* ctx, int32_t order, int64_t n, const double* h0, const double* h1,
                  const double* h2, const double* h3, double n_data, double n_rand,
                  int32_t flags, double* xi);

/* ---- asynchronous form (serving) ------------------------------------- *
 * pcf_submit enqueues a histogram job; the engine admits it when enough HBM
 * (the budget set by pcf_set_hbm_budget) and a stream slot are free, runs
 * admitted jobs concurrently on separate streams, and keeps the gridded
 * catalogs resident for reuse.  pcf_wait blocks until the job has finished
 * and copies its results out (counts/weights may be NULL). */
typedef int64_t pcf_ticket;
/* (Re)starts the engine: bytes <= 0 means 80% of the free HBM; max_inflight
 * lanes (1..16), each with its own stream.  Waits for submitted jobs first. */
int pcf_set_hbm_budget(pcf_ctx* ctx, int64_t bytes, int32_t max_inflight);
/* Continuous batching: a lane taking a pair job (PCF_2ISO / PCF_2ANI, mixed
 * precision) also takes up to max_batch - 1 queued compatible pair jobs (same
 * type, periodicity, box and tuning knobs; any bins and dmax) into one
 * persistent launch.  1 (the default) disables it. */
int pcf_set_max_batch(pcf_ctx* ctx, int32_t max_batch);
int pcf_submit(pcf_ctx* ctx, const pcf_job* job, pcf_ticket* ticket);
int pcf_poll(pcf_ctx* ctx, pcf_ticket ticket, int32_t* done);
int pcf_wait(pcf_ctx* ctx, pcf_ticket ticket, uint64_t* counts, double* weights,
             pcf_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* PCF_ABI_H */
