// pcf_internal.h -- types shared by the gfx950 kernels and the host side of
// libpcf_hip.so.  Not part of the public ABI (include/pcf_abi.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace pcf {

constexpr int kWave = 64;        // CDNA wavefront
constexpr int kTile = 64;        // points of one i-tile (one per lane)
constexpr int kBlock = 256;      // threads per workgroup of the counting kernels
constexpr int kMaxBinsIso = 1024;      // bins with exact FP64 thresholds (BinRules::T)
constexpr int kMaxBinsIsoMixed = 253;  // u8 table codes: bins, out, ambiguous
constexpr int kMaxBinsAniMixed = 120;
constexpr int64_t kMaxFlatBins = (int64_t)1 << 27;  // pair histogram entries (2 GB of output)
constexpr size_t kLdsPerCu = 160 * 1024;
// Plain FP64 pair kernel: histograms (counts + weights) above this many LDS
// bytes deposit straight into global memory.
constexpr size_t kPlainLdsHistMax = 96 * 1024;

// Cell grid of one job.  Cells are ordered x-major, z-fastest
// (cell = (cx*Py + cy)*Pz + cz), the order the reference linearises its nodes
// in (cuda/main.cu:274-317), so the cells of one (cx,cy) column are contiguous
// in memory and a z-range of a column is one contiguous point range.
struct GridGeom {
  int P[3];
  float lo[3];
  float cs[3];
  float inv_cs[3];
  int bpc;
  float L;
  __host__ __device__ int64_t ncell() const { return (int64_t)P[0] * P[1] * P[2]; }
  __host__ bool same(const GridGeom& o) const {
    for (int a = 0; a < 3; ++a)
      if (P[a] != o.P[a] || lo[a] != o.lo[a] || cs[a] != o.cs[a]) return false;
    return bpc == o.bpc && L == o.L;
  }
};

// Resident grids order the points of a cell by z sub-slab: a cell is cut into
// zq slabs along z and the counting sort runs over (cell, slab), so a column's
// points are in z order up to one slab and a 16-point block spans a thin range
// of z.  zq is a power of two, 16 unless the grid has more than 2^18 cells.
inline int grid_zq(const GridGeom& g) {
  int q = 16;
  while (q > 1 && g.ncell() * q > ((int64_t)1 << 22)) q >>= 1;
  return q;
}

// Index of the z sub-slab of coordinate z within its column: z cell * zq +
// slab (the z cell clamped as the grid clamps it).  Monotonic in z, and the
// same function places the points (grid build) and bounds the step windows
// (work list).
__host__ __device__ inline int zkey(float z, const GridGeom& g, int zq) {
  const float t = (z - g.lo[2]) * g.inv_cs[2];
  int cz = (int)floorf(t);
  cz = cz < 0 ? 0 : (cz >= g.P[2] ? g.P[2] - 1 : cz);
  if (zq == 1) return cz;
  int sub = (int)floorf((t - (float)cz) * (float)zq);
  sub = sub < 0 ? 0 : (sub >= zq ? zq - 1 : sub);
  return cz * zq + sub;
}

// One work item of the flattened pair work list: the points [i0, i0+ni) of
// one i-tile (ni <= 64, one per lane) against the contiguous j-range
// [j0, j0+nj) of one neighbour column segment, seen through periodic image
// `img` = (kx+1)*9 + (ky+1)*3 + (kz+1).  Bit 5 of `code`: the range overlaps
// the tile in an auto histogram, so only j > i is counted.  Bits 8..: slot of
// a batched launch.
struct PairItem {
  int i0, ni, j0, nj, code;
  // Step windows of the tile's four 16-point blocks (byte b of each word):
  // block b can pair in range only with the j of steps [ws_b, we_b), a step
  // being 16 consecutive j from j0.  kCodeNoWin in code: no windows (more
  // than 255 steps; every block takes every step).
  unsigned ws, we;
};

constexpr int kImgZero = 13;  // kx = ky = kz = 0
constexpr int kCodeMask = 32;
constexpr int kCodeNoWin = 64;

// Exact re-evaluation rules, uploaded once per job (device copy in Job).
struct BinRules {
  double R;           // range limit on d^2 ((double)(float)(dmax*dmax))
  double ds;          // bin scale
  double L;           // box side
  int le;             // range test uses <=
  int zero;           // 0 none, 1 iso regular d2>0, 2 ani both >0
  int bins;           // bins per dimension
  int ani;
  double T[kMaxBinsIso + 1];  // T[k] = min d2 with (int)(sqrt(d2)*ds) >= k, k=1..bins
};

// Table codes of the FP32 bulk path for 2ani (per component; iso tables hold
// the histogram row directly).  Out is the largest code and ambiguous the
// next, so the max of a pair's two codes is out when either component is out
// (the pair needs no exact check even if the other is ambiguous), ambiguous
// when one is and neither is out, else the larger bin.
constexpr uint8_t kAniOut8 = 255;
constexpr uint8_t kAniAmb8 = 254;

struct DevGrid {
  GridGeom g;
  float4* pts = nullptr;     // cell-ordered points
  double4* pts64 = nullptr;  // FP64 copy for the plain path (lazily built)
  int* cell_start = nullptr; // ncell+1
  int* fine_start = nullptr; // ncell*zq+1: start of each (cell, z sub-slab), when zq > 1
  int zq = 1;                // z sub-slabs per cell of the point order (grid_zq)
  int64_t n = 0;
  bool valid = false;
  hipEvent_t ready = nullptr;  // recorded after the last (re)build, on the building stream
  int users = 0;               // jobs holding the grid (no eviction or rebuild meanwhile)
  // Unique per build of the point order (next_grid_build_id): what was derived
  // from the grid (a triplet pair list) is still valid while it is unchanged.
  uint64_t build_id = 0;
};

// A process-wide unique id for each grid build.
uint64_t next_grid_build_id();

struct Catalog {
  float4* raw = nullptr;
  int64_t n = 0;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  float maxabs = 0;
  bool all_w1 = true;
  bool finite = true;
  uint64_t serial = 0;  // unique per upload within a context (ids are reused)
  std::vector<std::unique_ptr<DevGrid>> grids;  // resident grids by geometry
};

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t b = need + need / 4 + 256;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// A DevBuf freed when it leaves scope (per-call temporaries; hipFree waits
// for the device work that may still use it).
struct ScopedDevBuf : DevBuf {
  ScopedDevBuf() = default;
  ScopedDevBuf(const ScopedDevBuf&) = delete;
  ScopedDevBuf& operator=(const ScopedDevBuf&) = delete;
  ~ScopedDevBuf() { release(); }
};

// Page-locked host staging memory (small uploads and result read-backs that
// stay asynchronous on the stream).
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    size_t b = need + need / 4 + 256;
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// ---- device entry points (defined in grid.hip / pairs.hip) -------------
hipError_t launch_catalog_stats(const float4* raw, int64_t n, float* partials, int nblocks,
                                hipStream_t s);
hipError_t build_grid(const float4* raw, int64_t n, const GridGeom& g, DevGrid* out,
                      DevBuf& scratch_cells, DevBuf& scratch_perm, hipStream_t s);
// The same counting sort into caller-provided memory: pts[n], cell_start[ncell+1],
// with zq > 1 fine_start[ncell*zq+1] (the sort then runs over the z sub-slabs),
// and temporaries cells_tmp (grid_cells_tmp_bytes(ncell * zq)) and perm_tmp
// (2n ints).
size_t grid_cells_tmp_bytes(int64_t ncell);
hipError_t build_grid_into(const float4* raw, int64_t n, const GridGeom& g, float4* pts,
                           int* cell_start, int* fine_start, int zq, void* cells_tmp,
                           int* perm_tmp, hipStream_t s);
// Up to kMultiGridMax catalogs gridded on one geometry by one counting sort
// (build_grids_into): catalog k is raw[k][0 .. off[k+1] - off[k]), off[0] = 0,
// and its grid goes to pts[k], cell_start[k] and, with zq > 1, fine_start[k],
// as build_grid_into would write them.  Temporaries: cells_tmp of
// grid_cells_tmp_bytes(m * ncell * zq), perm_tmp of 2 * off[m] ints.
constexpr int kMultiGridMax = 8;
struct MultiGridArgs {
  const float4* raw[kMultiGridMax];
  float4* pts[kMultiGridMax];
  int* cell_start[kMultiGridMax];
  int* fine_start[kMultiGridMax];
  int64_t off[kMultiGridMax + 1];
  int m;
};
hipError_t build_grids_into(const MultiGridArgs& mg, const GridGeom& g, int zq, void* cells_tmp,
                            int* perm_tmp, hipStream_t s);
hipError_t convert_grid64(const float4* pts, int64_t n, double4* out, hipStream_t s);
hipError_t build_grid64(DevGrid* grid, hipStream_t s);

// One histogram of a batched pair launch (k_pairs_mixed<..., BATCH = 1>).
// Jobs of one batch share the histogram type, periodicity, box and grid; each
// has its own bins, range (dmax), d2 -> bin table, exact rules and scale.
struct BatchJob {
  const float4* xs;  // i side, scaled copy (this job's scale)
  const float4* ys;  // j side, scaled copy
  const float4* xo;  // i side, cell-ordered originals
  const float4* yo;  // j side, cell-ordered originals
  int hoff;          // first row of this job's histogram in the batch output
  int zero;          // zero-separation rule of this job (BinRules::zero)
  const uint8_t* table;    // this job's d2 -> bin table (device)
  const BinRules* rules;   // and exact rules (device)
  int nt, bins, nflat;
  float Ls;          // box side in this job's scaled units
  float tflip;       // 2iso: flipped-table top
  float rho2;        // Gram path bound (scaled units)
  float glo[3], gcs[3], ginv[3];  // grid geometry in this job's scaled units
};
// Behind a batch's histograms: 4 statistics words, 4 spare, then one chunk
// cursor per XCD range (8 ranges, 64 bytes apart).
constexpr size_t kBatchTailBytes = 64 + 8 * 64;
// A run of work items [begin, end) that all belong to job `job`.
struct BatchChunk {
  int begin, end, job;
};

struct PairLaunch {
  const float4* xp;   // i side, original coordinates
  const float4* yp;   // j side, original coordinates
  const float4* xs;   // i side scaled by sigma (bulk FP32 path)
  const float4* ys;   // j side scaled by sigma
  const double4* xp64;
  const double4* yp64;
  const PairItem* items;
  int nitems;
  const int* nitems_dev = nullptr;  // item count in device memory (set: nitems unused)
  int nitems_cap = 0;               // items allocated when nitems_dev is set
  int* counter;
  const uint8_t* table;
  int nt;
  float inv_h;
  float L;        // box side in the bulk path's (scaled) units
  int bins;
  int nflat;      // flat histogram bins (bins or bins^2)
  int cshift;     // log2 of histogram copies per workgroup
  unsigned long long* hist;
  double* whist;
  unsigned long long* stats;  // [0] evaluations, [1] rechecks, [2] work items
  const BinRules* rules;      // device copy
  int ani;
  int weighted;
  int plain64;
  int nblocks;
  double R;     // exact range limit on d^2
  double Lx;    // box side, unscaled
  int le;
  int zero;
  // batched launch (mixed path only): chunks of the concatenated work list
  int batch = 0;
  int mfma = 0;       // 2iso unweighted: Gram-form distances on the matrix cores
  float rho2 = 0.f;   // ... for tiles whose box half-diagonal^2 (scaled) is <= rho2
  float tflip = 0.f;  // 2iso mixed path: the table is indexed by cvt_rz(tflip - d2)
  GridGeom geo{};     // grid of the launch (Gram path: tile boxes from cells)
  float sigma = 1.f;  // scale of the bulk path's coordinates
  int prepare_only = 0;  // only set the kernel's LDS attribute (before graph capture)
  int global_hist = 0;   // plain path: deposit into the global histogram (no LDS copy)
  const BatchJob* jobs = nullptr;
  const BatchChunk* chunks = nullptr;
  int nchunks = 0;
};
constexpr int kMixedBlock = 512;       // threads per workgroup of the mixed kernel
// Recheck queue of the mixed pair kernel: per wave up to 2 x 64 entries of
// (i index, j index | image << 27), so j-side catalogs need n < 2^27 (larger
// ones take the plain FP64 kernel, which gives the same results).
constexpr int kRecheckCap = 128;
constexpr int kRecheckImgShift = 27;
constexpr int kRecheckIndexMask = (1 << kRecheckImgShift) - 1;
constexpr int64_t kRecheckMaxPoints = (int64_t)1 << kRecheckImgShift;
constexpr size_t kRecheckQueueBytes = (size_t)(kMixedBlock / 64) * kRecheckCap * 8;
constexpr int kMixedTableMax = 28672;  // static LDS bytes of its d2 -> bin table
constexpr int kGramZsub = 4;  // Gram pair path: thin z cells per lateral cell edge

// Launch shape of the pair kernels: 2^cshift LDS copies of each histogram row
// (lanes spread their atomics over the copies) and the resident workgroups per
// CU of the persistent grid.  Defaults: as many copies as fit 48 KB (16 KB
// for unweighted 2ani), as many
// workgroups as the 160 KB of LDS per CU hold.  The overrides (the job's
// lds_copies_log2 = log2 + 1 and blocks_per_cu; 0 = default) are the
// autotuner's knobs; no choice changes a result.
struct PairShape {
  int cshift;
  int per_cu;
  size_t lds_bytes;
};
inline PairShape pair_shape(int nflat, int bins, bool weighted, bool plain, int lds_log2p1,
                            int per_cu_override) {
  const int rows = nflat + 2;
  const int per = 4 + (weighted ? 8 : 0);
  // Unweighted 2ani (nflat = bins^2) deposits only pairs in range, so a few
  // copies of its 900-odd rows suffice and leave LDS for more workgroups.
  const size_t budget = (nflat != bins && !weighted) ? 16 * 1024 : 48 * 1024;
  int copies = 64;
  while (copies > 1 && (size_t)rows * copies * per > budget) copies >>= 1;
  int cs = 0;
  while ((1 << cs) < copies) ++cs;
  if (lds_log2p1 > 0) {
    cs = lds_log2p1 - 1;
    if (cs > 6) cs = 6;
    while (cs > 0 && (size_t)rows * ((size_t)1 << cs) * per > 120 * 1024) --cs;
  }
  PairShape sh;
  sh.cshift = cs;
  sh.lds_bytes = (plain ? 0 : (size_t)kMixedTableMax + kRecheckQueueBytes) +
                 (size_t)rows * ((size_t)1 << cs) * per + (size_t)(bins + 2) * 8;
  const int threads = plain ? 256 : kMixedBlock;
  int per_cu = (int)((160 * 1024) / (sh.lds_bytes + 512));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 2048 / threads) per_cu = 2048 / threads;
  if (per_cu_override > 0 && per_cu_override < per_cu) per_cu = per_cu_override;
  sh.per_cu = per_cu;
  return sh;
}
// Whether the mixed (FP32 + FP64 recheck) pair kernel serves a histogram: its
// u8 table codes, exact thresholds and LDS histogram must fit.  Other shapes
// run the plain FP64 kernel, which gives the same results.
inline bool pair_mixed_fits(int bins, bool ani, bool weighted) {
  if (bins <= 0 || bins > (ani ? kMaxBinsAniMixed : kMaxBinsIsoMixed)) return false;
  const int nflat = ani ? bins * bins : bins;
  return pair_shape(nflat, bins, weighted, false, 0, 0).lds_bytes <= kLdsPerCu;
}
hipError_t launch_pairs(const PairLaunch& pl, hipStream_t s);
hipError_t launch_scale_points(const float4* in, int64_t n, float sigma, float4* out,
                               hipStream_t s);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// Exclusive prefix sum over the device: out[0..n) and the total in out[n];
// a total of 2^31 or more is written as -1 (and out[0..n) is then not usable).
// tsum holds scan_tiles(n) tile sums.  out may alias in.
inline int64_t scan_tiles(int64_t n) { return (n + 4095) / 4096 + 1; }
hipError_t launch_scan(const int* in, int64_t n, int* out, long long* tsum, hipStream_t s);

struct WorkListArgs {
  const int* x_cell_start;  // i side (tiles)
  const int* y_cell_start;  // j side
  GridGeom g;
  int K[3];                 // reach in cells per axis
  float reach;              // rmax plus slack
  int auto_;
  int cylinder;             // 2ani: |dz| < rmax and d_perp < rmax (not a sphere)
  const int* map;           // optional: j index -> position (pair-list offsets)
  // Optional (null: windows from whole z cells): the j grid's sub-slab starts
  // and the i grid's points, both ordered with zq sub-slabs per cell; a
  // block's window then covers only the sub-slabs within reach of its own.
  const int* y_fine_start;
  const float4* x_pts;
  int zq;
};
// Two-pass (count, scan, fill) work list on the device.  Returns the item
// count through *nitems (a synchronous read of one integer).
hipError_t build_pair_worklist(const WorkListArgs& a, DevBuf& counts, DevBuf& items,
                               int* nitems, hipStream_t s);
// Upper bound of the work list's length: one item per (tile, neighbour column,
// z segment), tiles = ceil(column points / 64) summed over the columns.
int64_t worklist_bound(const WorkListArgs& a, int64_t n_tile_points);
// The same work list without the host round trip: `items` is sized for `cap`
// items (cap >= the count, e.g. worklist_bound) and the count stays on the
// device at *total_dev for the counting kernel to read.  The fill skips any
// cell whose items would end beyond cap, so memory stays in bounds even if
// cap were short; the kernels clamp the count to cap and report it.
hipError_t build_pair_worklist_async(const WorkListArgs& a, int64_t cap, DevBuf& counts,
                                     DevBuf& items, const int** total_dev, hipStream_t s);
// The same into caller-provided memory: `counts` of worklist_count_bytes(a)
// bytes and `items` of cap entries.
size_t worklist_count_bytes(const WorkListArgs& a);
hipError_t build_pair_worklist_into(const WorkListArgs& a, int64_t cap, void* counts,
                                    PairItem* items, const int** total_dev, hipStream_t s);

}  // namespace pcf
