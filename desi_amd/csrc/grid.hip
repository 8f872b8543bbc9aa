// grid.hip -- on-device cell grid ("create_grid" of the reference).
//
// The reference classifies points into nodes on the host, one point at a time,
// re-allocating the node's array on every insert (cuda/create_grid.cuh:81-138),
// then linearises the non-empty nodes and copies them to the device
// (cuda/main.cu:215-329).  Here the grid is a counting sort on the device:
//   1. k_cell_index  : sort key of every point (its cell's z sub-slab), per-key
//                      counts (global atomics) and the point's rank among its
//                      key's points (the atomic's return)
//   2. launch_scan   : key starts = exclusive prefix sum of the counts; the
//                      cell starts are every zq-th of them
//   3. k_scatter     : point i copied to start[key] + rank, float4
// The resulting layout is cell-ordered (x-major, z fastest) so that every z-range
// of a (cx,cy) column is one contiguous, coalesced point range, and within a
// cell the points are in z order up to one sub-slab (cs_z / zq).
#include <algorithm>

#include "pcf_internal.h"

namespace pcf {

namespace {

constexpr int kStatsBlock = 256;

// Per-block reduction of the catalog: min/max per axis, max |coordinate|,
// whether every weight is exactly 1 and whether every value is finite.
// partials[b*10 + 0..2] = min, [3..5] = max, [6] = max|c|, [7] = all_w1,
// [8] = finite, [9] unused.
__global__ __launch_bounds__(kStatsBlock) void k_catalog_stats(const float4* __restrict__ raw,
                                                               int64_t n,
                                                               float* __restrict__ partials) {
  __shared__ float red[10][kStatsBlock];
  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  float ma = 0.f, w1 = 1.f, fin = 1.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float4 p = raw[i];
    float v[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], v[a]);
      mx[a] = fmaxf(mx[a], v[a]);
      ma = fmaxf(ma, fabsf(v[a]));
    }
    if (p.w != 1.0f) w1 = 0.f;
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w))) fin = 0.f;
  }
  const int t = threadIdx.x;
  for (int a = 0; a < 3; ++a) {
    red[a][t] = mn[a];
    red[3 + a][t] = mx[a];
  }
  red[6][t] = ma;
  red[7][t] = w1;
  red[8][t] = fin;
  __syncthreads();
  for (int s = kStatsBlock / 2; s > 0; s >>= 1) {
    if (t < s) {
      for (int a = 0; a < 3; ++a) {
        red[a][t] = fminf(red[a][t], red[a][t + s]);
        red[3 + a][t] = fmaxf(red[3 + a][t], red[3 + a][t + s]);
      }
      red[6][t] = fmaxf(red[6][t], red[6][t + s]);
      red[7][t] = fminf(red[7][t], red[7][t + s]);
      red[8][t] = fminf(red[8][t], red[8][t + s]);
    }
    __syncthreads();
  }
  if (t < 9) partials[(int64_t)blockIdx.x * 10 + t] = red[t][0];
}

__device__ __forceinline__ int cell_axis(float v, float lo, float inv_cs, int P) {
  int c = (int)floorf((v - lo) * inv_cs);
  c = c < 0 ? 0 : c;
  return c >= P ? P - 1 : c;
}

// Sort key of every point (its cell, or with zq > 1 its cell's z sub-slab:
// column * Pz * zq + zkey) and its rank among the points of that key (the
// counting atomic's return value), so the scatter needs no second atomic pass.
__global__ void k_cell_index(const float4* __restrict__ raw, int64_t n, GridGeom g, int zq,
                             int* __restrict__ cell_of, int* __restrict__ rank,
                             int* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = raw[i];
  int cx = cell_axis(p.x, g.lo[0], g.inv_cs[0], g.P[0]);
  int cy = cell_axis(p.y, g.lo[1], g.inv_cs[1], g.P[1]);
  int c = (cx * g.P[1] + cy) * g.P[2] * zq + zkey(p.z, g, zq);
  cell_of[i] = c;
  rank[i] = atomicAdd(&counts[c], 1);
}

// cell_start[c] = fine_start[c * zq] for c = 0 .. ncell (the last is the total).
__global__ void k_coarse_starts(const int* __restrict__ fine_start, int64_t ncell, int zq,
                                int* __restrict__ cell_start) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= ncell) cell_start[c] = fine_start[c * zq];
}


__global__ void k_scatter(const float4* __restrict__ raw, int64_t n, const int* __restrict__ cell_of,
                          const int* __restrict__ rank, const int* __restrict__ cell_start,
                          float4* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[cell_start[cell_of[i]] + rank[i]] = raw[i];
}

// Several catalogs on one geometry in one counting sort: catalog k's keys
// are offset by k * nk, so one scan gives the starts of every catalog, each
// offset by the points of the catalogs before it (off[k]).
__device__ __forceinline__ int multi_cat(const MultiGridArgs& mg, int64_t t) {
  int k = 0;
  while (k + 1 < mg.m && t >= mg.off[k + 1]) ++k;
  return k;
}

__global__ void k_cell_index_multi(MultiGridArgs mg, GridGeom g, int zq, int64_t nk,
                                   int* __restrict__ cell_of, int* __restrict__ rank,
                                   int* __restrict__ counts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= mg.off[mg.m]) return;
  const int k = multi_cat(mg, t);
  const float4 p = mg.raw[k][t - mg.off[k]];
  const int cx = cell_axis(p.x, g.lo[0], g.inv_cs[0], g.P[0]);
  const int cy = cell_axis(p.y, g.lo[1], g.inv_cs[1], g.P[1]);
  const int c = (int)((int64_t)k * nk) + (cx * g.P[1] + cy) * g.P[2] * zq + zkey(p.z, g, zq);
  cell_of[t] = c;
  rank[t] = atomicAdd(&counts[c], 1);
}

// Catalog k's fine starts (zq > 1) and cell starts from the joint scan.
__global__ void k_split_starts(MultiGridArgs mg, const int* __restrict__ starts, int64_t nc,
                               int zq) {
  const int64_t nk = nc * zq;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mg.m * (nk + 1)) return;
  const int k = (int)(t / (nk + 1));
  const int64_t q = t - (int64_t)k * (nk + 1);
  const int v = starts[(int64_t)k * nk + q] - (int)mg.off[k];
  if (zq > 1) mg.fine_start[k][q] = v;
  if (q % zq == 0) mg.cell_start[k][q / zq] = v;
}

__global__ void k_scatter_multi(MultiGridArgs mg, const int* __restrict__ cell_of,
                                const int* __restrict__ rank, const int* __restrict__ starts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= mg.off[mg.m]) return;
  const int k = multi_cat(mg, t);
  mg.pts[k][starts[cell_of[t]] + rank[t] - (int)mg.off[k]] = mg.raw[k][t - mg.off[k]];
}

__global__ void k_to_double(const float4* __restrict__ in, int64_t n, double4* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = in[i];
  out[i] = make_double4(p.x, p.y, p.z, p.w);
}

__global__ void k_zero_words(uint32_t* p, int64_t nwords) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) p[i] = 0u;
}

}  // namespace

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  int64_t words = (int64_t)((bytes + 3) / 4);
  if (words == 0) return hipSuccess;
  int blocks = (int)((words + 255) / 256);
  hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(256), 0, s, (uint32_t*)p, words);
  return hipGetLastError();
}

hipError_t launch_catalog_stats(const float4* raw, int64_t n, float* partials, int nblocks,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_catalog_stats, dim3(nblocks), dim3(kStatsBlock), 0, s, raw, n, partials);
  return hipGetLastError();
}

hipError_t build_grid(const float4* raw, int64_t n, const GridGeom& g, DevGrid* out,
                      DevBuf& scratch_cells, DevBuf& scratch_perm, hipStream_t s) {
  hipError_t e;
  const int64_t nc = g.ncell();
  if (!out->pts || out->n != n) {
    if (out->pts) (void)hipFree(out->pts);
    out->pts = nullptr;
    if ((e = hipMalloc(&out->pts, sizeof(float4) * (n > 0 ? n : 1))) != hipSuccess) return e;
  }
  const int zq = grid_zq(g);
  if (!out->cell_start || !out->g.same(g)) {
    if (out->cell_start) (void)hipFree(out->cell_start);
    if (out->fine_start) (void)hipFree(out->fine_start);
    out->cell_start = nullptr;
    out->fine_start = nullptr;
    if ((e = hipMalloc(&out->cell_start, sizeof(int) * (nc + 1))) != hipSuccess) return e;
    if (zq > 1 && (e = hipMalloc(&out->fine_start, sizeof(int) * (nc * zq + 1))) != hipSuccess)
      return e;
  }
  if (out->pts64) {
    (void)hipFree(out->pts64);
    out->pts64 = nullptr;
  }
  out->g = g;
  out->n = n;
  out->zq = zq;
  if ((e = scratch_cells.ensure(grid_cells_tmp_bytes(nc * zq))) != hipSuccess) return e;
  if ((e = scratch_perm.ensure(2 * sizeof(int) * (n > 0 ? n : 1))) != hipSuccess) return e;
  if ((e = build_grid_into(raw, n, g, out->pts, out->cell_start, out->fine_start, zq,
                           scratch_cells.p, scratch_perm.as<int>(), s)) != hipSuccess)
    return e;
  out->valid = true;
  return hipSuccess;
}

size_t grid_cells_tmp_bytes(int64_t ncell) {
  const size_t cbytes = (sizeof(int) * (size_t)(ncell + 1) + 15) & ~(size_t)15;
  return cbytes + sizeof(long long) * (size_t)scan_tiles(ncell);
}

hipError_t build_grid_into(const float4* raw, int64_t n, const GridGeom& g, float4* pts,
                           int* cell_start, int* fine_start, int zq, void* cells_tmp,
                           int* perm_tmp, hipStream_t s) {
  hipError_t e;
  if (zq < 1 || (zq > 1 && !fine_start)) return hipErrorInvalidValue;
  const int64_t nc = g.ncell();
  const int64_t nk = nc * zq;  // sort keys
  int* starts = zq > 1 ? fine_start : cell_start;
  const size_t cbytes = (sizeof(int) * (size_t)(nk + 1) + 15) & ~(size_t)15;
  int* counts = reinterpret_cast<int*>(cells_tmp);
  int* cell_of = perm_tmp;
  int* rank = cell_of + (n > 0 ? n : 1);
  if ((e = launch_zero(counts, sizeof(int) * (nk + 1), s)) != hipSuccess) return e;
  const int blocks = (int)((n + 255) / 256);
  if (n > 0) {
    hipLaunchKernelGGL(k_cell_index, dim3(blocks), dim3(256), 0, s, raw, n, g, zq, cell_of, rank,
                       counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if ((e = launch_scan(counts, nk, starts,
                       reinterpret_cast<long long*>(reinterpret_cast<char*>(cells_tmp) + cbytes),
                       s)) != hipSuccess)
    return e;
  if (zq > 1) {
    hipLaunchKernelGGL(k_coarse_starts, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, s,
                       fine_start, nc, zq, cell_start);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, raw, n, cell_of, rank, starts,
                       pts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t build_grids_into(const MultiGridArgs& mg, const GridGeom& g, int zq, void* cells_tmp,
                            int* perm_tmp, hipStream_t s) {
  hipError_t e;
  if (zq < 1 || mg.m < 1 || mg.m > kMultiGridMax) return hipErrorInvalidValue;
  const int64_t nc = g.ncell();
  const int64_t nk = nc * zq;
  const int64_t keys = (int64_t)mg.m * nk;
  const int64_t n = mg.off[mg.m];
  if (keys + 1 > (int64_t)INT32_MAX || n > (int64_t)INT32_MAX) return hipErrorInvalidValue;
  for (int k = 0; k < mg.m; ++k)
    if (!mg.cell_start[k] || (zq > 1 && !mg.fine_start[k])) return hipErrorInvalidValue;
  int* counts = reinterpret_cast<int*>(cells_tmp);
  const size_t cbytes = (sizeof(int) * (size_t)(keys + 1) + 15) & ~(size_t)15;
  int* cell_of = perm_tmp;
  int* rank = cell_of + (n > 0 ? n : 1);
  if ((e = launch_zero(counts, sizeof(int) * (size_t)(keys + 1), s)) != hipSuccess) return e;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (n > 0) {
    hipLaunchKernelGGL(k_cell_index_multi, dim3(blocks), dim3(256), 0, s, mg, g, zq, nk, cell_of,
                       rank, counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // In place: counts become the joint starts.
  if ((e = launch_scan(counts, keys, counts,
                       reinterpret_cast<long long*>(reinterpret_cast<char*>(cells_tmp) + cbytes),
                       s)) != hipSuccess)
    return e;
  const int64_t nsplit = (int64_t)mg.m * (nk + 1);
  hipLaunchKernelGGL(k_split_starts, dim3((unsigned)((nsplit + 255) / 256)), dim3(256), 0, s, mg,
                     counts, nc, zq);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter_multi, dim3(blocks), dim3(256), 0, s, mg, cell_of, rank, counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t convert_grid64(const float4* pts, int64_t n, double4* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts, n,
                     out);
  return hipGetLastError();
}

hipError_t build_grid64(DevGrid* grid, hipStream_t s) {
  if (grid->pts64) return hipSuccess;
  hipError_t e = hipMalloc(&grid->pts64, sizeof(double4) * (grid->n > 0 ? grid->n : 1));
  if (e != hipSuccess) return e;
  return convert_grid64(grid->pts, grid->n, grid->pts64, s);
}

// ---------------------------------------------------------------------------
// Device-wide exclusive prefix sum of int counts: out[i] = sum(in[0..i)),
// out[n] = total, or -1 when the total is 2^31 or more.  Three passes over tiles of 4096:
// tile sums, a one-workgroup scan of the tile sums, then each tile scanned in
// LDS with its offset.  out may alias in.
// ---------------------------------------------------------------------------
namespace {
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_tile_sums(const int* __restrict__ in, int64_t n,
                                                            long long* __restrict__ tsum) {
  __shared__ long long red[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  long long s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) tsum[blockIdx.x] = red[0];
}

constexpr int kSumsBlock = 1024;
__global__ __launch_bounds__(kSumsBlock) void k_scan_tile_sums(long long* __restrict__ tsum,
                                                               int64_t ntiles,
                                                               int* __restrict__ total) {
  __shared__ long long part[kSumsBlock];
  const int t = threadIdx.x;
  const int64_t chunk = (ntiles + kSumsBlock - 1) / kSumsBlock;
  const int64_t b = t * chunk;
  const int64_t e = b + chunk < ntiles ? b + chunk : ntiles;
  long long s = 0;
  for (int64_t i = b; i < e; ++i) s += tsum[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < kSumsBlock; off <<= 1) {
    long long v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - s;
  for (int64_t i = b; i < e; ++i) {
    const long long c = tsum[i];
    tsum[i] = run;
    run += c;
  }
  // Totals of 2^31 or more are flagged as -1 (the tile offsets of such a scan
  // wrap, so its consumers must not use them); callers test for a negative total.
  if (t == kSumsBlock - 1) *total = part[t] > 0x7fffffffLL ? -1 : (int)part[t];
}

__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const int* in, int64_t n,
                                                            const long long* __restrict__ tsum,
                                                            int* out) {
  __shared__ int tile[kScanTile];
  __shared__ long long wsum[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int k = 0; k < kScanItems; ++k) {
    const int j = k * kScanThreads + threadIdx.x;
    const int64_t i = base + j;
    tile[j] = i < n ? in[i] : 0;
  }
  __syncthreads();
  const int t = threadIdx.x;
  long long s = 0;
  for (int k = 0; k < kScanItems; ++k) s += tile[t * kScanItems + k];
  wsum[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    long long v = t >= off ? wsum[t - off] : 0;
    __syncthreads();
    wsum[t] += v;
    __syncthreads();
  }
  long long run = tsum[blockIdx.x] + wsum[t] - s;
  for (int k = 0; k < kScanItems; ++k) {
    const int c = tile[t * kScanItems + k];
    tile[t * kScanItems + k] = (int)run;
    run += c;
  }
  __syncthreads();
  for (int k = 0; k < kScanItems; ++k) {
    const int j = k * kScanThreads + threadIdx.x;
    const int64_t i = base + j;
    if (i < n) out[i] = tile[j];
  }
}
}  // namespace

hipError_t launch_scan(const int* in, int64_t n, int* out, long long* tsum, hipStream_t s) {
  hipError_t e;
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, n,
                       tsum);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kSumsBlock), 0, s, tsum, ntiles, out + n);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, n, tsum,
                       out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace pcf
