// pairs.hip -- flattened cell-pair work list and the gfx950 pair-histogram
// kernels (2iso, 2ani; periodic or not).
//
// Reference behaviour restated (cuda/kernels/2iso_k.cuh:8-120,
// 2isoBPC_k.cuh:8-557, 2ani_k.cuh:7-135, 2aniBPC_k.cuh:8-682): every ordered
// pair of points whose separation passes the range test adds w_i*w_j to bin
// (int)(sqrt(d2)*ds).  The reference runs one thread per ordered node pair with
// a global FP64 atomicAdd per counted pair and evaluates periodic images with
// seven branchy copies of the point loop.  Here:
//
//   * the work list holds (i-tile of <= 64 points, contiguous j-range of one
//     neighbour column segment, periodic image) items, built on the device;
//     periodic boundaries are one image offset per item, not per point;
//   * auto histograms (DD, RR) visit each unordered pair once (j > i in cell
//     order) and double the result at the end;
//   * each lane holds one i point in VGPRs; the j points stream through scalar
//     loads (uniform across the wave), so no LDS is spent on point staging;
//   * the squared distance is formed in FP32 (3 sub + mul + 2 fma) and binned
//     through a d2 -> bin table in LDS built on the host from exact FP64 bin
//     thresholds; table cells that a bin edge or the range limit could cross
//     within the FP32 error bound are marked ambiguous, and the lanes that hit
//     them re-evaluate their pair in FP64 exactly as the oracle does;
//   * counts go to a per-lane-column LDS histogram (no bank conflicts among the
//     lanes of a wave), merged into the global uint64 / FP64 histogram once per
//     workgroup of a persistent grid that pulls items from an atomic counter.
#include <type_traits>
#include <vector>

#include "pcf_internal.h"
#include "worklist.h"
#include "exact.h"


// Wave-wide min / max (ROCm device library, DPP-based).
extern "C" __device__ __attribute__((const)) float __ockl_wfred_min_f32(float);
extern "C" __device__ __attribute__((const)) float __ockl_wfred_max_f32(float);

namespace pcf {

namespace {

// One thread per (cell, s), s < S; entry t = cell * S + s of counts/offsets.
__global__ void k_worklist_count(WorkListArgs a, int S, int* __restrict__ counts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.g.ncell() * S) return;
  counts[t] = enumerate_cell<false>(a, (int)(t / S), (int)(t % S), S, nullptr);
}

__global__ void k_worklist_fill(WorkListArgs a, int S, const int* __restrict__ offsets,
                                PairItem* __restrict__ items, int64_t cap) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.g.ncell() * S) return;
  // Beyond the list's allocation, or a count that overflowed the scan.
  if (offsets[t] < 0 || offsets[t + 1] < offsets[t] || (int64_t)offsets[t + 1] > cap) return;
  enumerate_cell<true>(a, (int)(t / S), (int)(t % S), S, items + offsets[t]);
}


// ---------------------------------------------------------------------------
// Counting kernels
// ---------------------------------------------------------------------------
struct KArgs {
  int nitems;
  int nt;
  float inv_h;
  float L;      // box side in the bulk path's units (scaled by sigma)
  int bins;
  int nflat;
  int cshift;
  double R;     // exact rules (mirror BinRules) for the in-kernel FP64 path
  double Lx;    // box side, unscaled
  int le;
  int zero;
  int claim;    // work items a wave claims per atomic (0: derived from the count)
  const int* nitems_dev;  // item count in device memory (nullptr: nitems)
  int cap;                // items allocated (the count is clamped to it)
  float rho2;             // Gram path: largest squared half-diagonal of a tile's
                          // bounding box (scaled units) its error bound covers
  float tflip;            // the table is indexed by cvt_rz(tflip - d2) (2ani: per component)
  // Grid geometry in the bulk path's (scaled) units: the Gram path bounds a
  // tile by its cells.
  float glo[3], gcs[3], ginv[3];
  int gP[3];
};

// Item count of a launch (from device memory when the list was built without
// a host round trip) and the claim granularity: up to 8 items per atomic
// while every wave still gets about 8 claims (tail balance).
__device__ __forceinline__ void launch_items(const KArgs& k, int* nitems, int* claim) {
  int n = k.nitems;
  if (k.nitems_dev) n = min(__builtin_amdgcn_readfirstlane(*k.nitems_dev), k.cap);
  *nitems = n;
  int cl = k.claim;
  if (cl <= 0) {
    const long long waves = (long long)gridDim.x * (blockDim.x / kWave);
    const long long c = (long long)n / (waves * 8);
    cl = (int)(c < 1 ? 1 : (c > 8 ? 8 : c));
  }
  *claim = cl;
}

// Loads of catalog points through the global address space.  The batched
// launch takes its point arrays from the job table, where the compiler cannot
// tell their address space and would emit flat loads, whose completion the
// LDS waits of the inner loop would then also wait for.
typedef float gv4f __attribute__((ext_vector_type(4)));
typedef const gv4f __attribute__((address_space(1)))* GV4;
__device__ __forceinline__ float4 gld(const float4* p, int i) {
  const gv4f v = ((GV4)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
// The Gram loops' j points: x, y, z only (a 12-byte load), so that no
// register of the prefetch in flight is free for the compiler to reuse (an
// unused w would be, forcing a wait on the prefetch within the step that
// issued it), through a buffer resource over the run [0, n): the lane's
// offset is fixed and the step's is scalar (no per-lane address arithmetic),
// and points at or past n read as zeros (the loops mask those lanes).
typedef float gv3f __attribute__((ext_vector_type(3)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t point_rsrc(const float4* p, int n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)((unsigned)n * 16u),
                                           0x00020000);
}
__device__ __forceinline__ gv3f bld3(__amdgpu_buffer_rsrc_t r, int lane_off, int j) {
  return __builtin_amdgcn_raw_buffer_load_b96(r, lane_off, j * 16, 0);
}

// FP32 bulk + FP64 edge re-evaluation.
//
// Table codes (u8): iso -> the histogram row (bin, nflat = out, nflat+1 =
// ambiguous); ani -> per component bin, kAniAmb8 / kAniOut8.  For both the
// ambiguous row is the largest row index, so one max over four pairs tells
// whether any of them needs the exact path.
template <int ANI>
__device__ __forceinline__ unsigned pair_row(const float4 q, float xs, float ys, float zs,
                                             unsigned ntm1, const uint8_t* __restrict__ tab,
                                             unsigned bins, unsigned out_row, unsigned amb_row,
                                             float tflip) {
  const float dx = q.x - xs;
  const float dy = q.y - ys;
  const float dz = q.z - zs;
  if (!ANI) {
    // Flipped index: d2 beyond the table saturates to cell 0 (out).
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    return tab[__float2uint_rz(tflip - d2)];
  } else {
    // 2ani reads two cells of the flipped table: lanes with a component
    // beyond the table (cell 0, out of range) skip both reads.
    const float dz2 = dz * dz;
    const float do2 = __builtin_fmaf(dy, dy, dx * dx);
    const unsigned tz = __float2uint_rz(tflip - dz2);
    const unsigned to = __float2uint_rz(tflip - do2);
    unsigned r = out_row;
    if (tz != 0u && to != 0u) {
      const unsigned cz = tab[tz];
      const unsigned co = tab[to];
      const unsigned m = max(cz, co);
      r = m < bins ? cz * bins + co : (m == kAniOut8 ? out_row : amb_row);
    }
    return r;
  }
}

// The oracle's arithmetic for one (pair, image): FP64 from the unscaled FP32
// inputs, no contraction.  Returns the histogram row (out_row if excluded).
template <int ANI>
__device__ __forceinline__ unsigned exact_row(const float4 pi, const float4 pj, int img,
                                              const double* T, int bins, double R, int le,
                                              int zero, double L, unsigned out_row) {
  const int kx = img / 9 - 1, ky = (img / 3) % 3 - 1, kz = img % 3 - 1;
  double tx, ty, tz;
  if (!axis_sq(pi.x, pj.x, kx, L, &tx)) return out_row;
  if (!axis_sq(pi.y, pj.y, ky, L, &ty)) return out_row;
  if (!axis_sq(pi.z, pj.z, kz, L, &tz)) return out_row;
  if (!ANI) {
    const double d2 = __dadd_rn(__dadd_rn(tx, ty), tz);
    if (zero == 1 && img == kImgZero && !(d2 > 0.0)) return out_row;
    const int b = class_lds(d2, T, bins, R, le);
    return b < 0 ? out_row : (unsigned)b;
  }
  const double do2 = __dadd_rn(tx, ty);
  if (zero == 2 && !(tz > 0.0 && do2 > 0.0)) return out_row;
  const int bz = class_lds(tz, T, bins, R, le);
  if (bz < 0) return out_row;
  const int bo = class_lds(do2, T, bins, R, le);
  if (bo < 0) return out_row;
  return (unsigned)(bz * bins + bo);
}

#ifndef PCF_BATCH_CLAIM
#define PCF_BATCH_CLAIM 4
#endif
constexpr int kBatchClaim = PCF_BATCH_CLAIM;  // items a wave takes per claim in a batch chunk
constexpr int kBlockMixed = kMixedBlock;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2
constexpr int kMaxTable = kMixedTableMax;

// xp/yp: points scaled by sigma (bulk arithmetic); xp_orig/yp_orig: the same
// points unscaled (exact path, weights).
//
// BATCH = 1: many histograms in one persistent launch.  The work list is the
// concatenation of the jobs' lists, cut into chunks that never straddle two
// jobs; a workgroup takes a chunk (global counter), its waves take the chunk's
// items (LDS counter), and at the end of the chunk the workgroup flushes its
// LDS histogram into that job's rows and clears it.  Point arrays and the zero
// rule come from the chunk's job descriptor; xp..yp_orig are unused.
//
// MF = 1 (unweighted; 2ani: see run_item_mf): squared distances of an item
// come from the matrix cores in Gram form, d2 = |xi|^2 + |xj|^2 - 2 xi.xj, 16 j x 16 i per
// v_mfma_f32_16x16x4f32 (rows = j: A operand (x, y, z, |x|^2) of j;
// columns = i: B operand (-2x, -2y, -2z, 1) of i; C = |x_i|^2 broadcast down
// the column).  Coordinates are taken relative to the centre of the i-tile's
// bounding box, so the FP32 error is bounded by the box size; that bound is
// part of the bin table's ambiguity margin.  An item whose tile box exceeds
// the launch's bound (k.rho2) takes the direct per-lane path instead.  The
// instruction is exactly a chain of four fmaf in k order on gfx950 (measured
// bit for bit); the bound allows four roundings of partial sums of at most
// (2 rho + d)^2 in any order, with a factor 2 of margin.
template <int ANI, int WEIGHTED, int CS, int BATCH, int MF>
__global__ __launch_bounds__(kBlockMixed) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_pairs_mixed(
    const float4* __restrict__ xp, const float4* __restrict__ yp,
    const float4* __restrict__ xp_orig, const float4* __restrict__ yp_orig,
    const PairItem* __restrict__ items, int* __restrict__ counter,
    const uint8_t* __restrict__ table, const BinRules* __restrict__ rules,
    unsigned long long* __restrict__ hist, double* __restrict__ whist,
    unsigned long long* __restrict__ stats, KArgs k_in, const BatchJob* __restrict__ jobs,
    const BatchChunk* __restrict__ chunks) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint8_t tab[kMaxTable];
  // The job parameters in force; a batched launch switches them per chunk
  // (jobs of one batch may differ in bins, range and scale).  k_in's nflat and
  // bins are the largest of the launch and size the LDS layout.
  KArgs k = k_in;
  const int nrows = k_in.nflat + 2;
  constexpr int ncopy = 1 << CS;
  uint32_t* hc = reinterpret_cast<uint32_t*>(smem);
  const int hc_bytes = (nrows * ncopy * 4 + 15) & ~15;
  double* sT = reinterpret_cast<double*>(smem + hc_bytes);
  double* hw = sT + ((k_in.bins + 2) & ~1);
  // Per-wave queue of the pairs whose table cell straddles an edge.
  int2* const wq =
      reinterpret_cast<int2*>(hw + (WEIGHTED ? nrows * ncopy : 0)) + (threadIdx.x >> 6) * kRecheckCap;
  for (int t = threadIdx.x; t < k.nt; t += kBlockMixed) tab[t] = table[t];
  for (int t = threadIdx.x; t < nrows * ncopy; t += kBlockMixed) {
    hc[t] = 0u;
    if (WEIGHTED) hw[t] = 0.0;
  }
  for (int t = threadIdx.x; t <= k.bins; t += kBlockMixed) sT[t] = rules->T[t];
  __syncthreads();

  const int lane = threadIdx.x & (kWave - 1);
  // Gram path: lane group grp = lane / 16 carries component grp of its point
  // in the matrix operands; col = lane % 16 is its point within a block of 16.
  const int grp = lane >> 4, col = lane & 15;
  const float sel_x = grp == 0 ? 1.0f : 0.0f, sel_y = grp == 1 ? 1.0f : 0.0f;
  const float sel_z = grp == 2 ? 1.0f : 0.0f, sel_w = grp == 3 ? 1.0f : 0.0f;
  uint32_t* const hc_lane = hc + (lane & (ncopy - 1));
  double* const hw_lane = hw + (lane & (ncopy - 1));
  unsigned ntm1 = (unsigned)(k.nt - 1);
  unsigned out_row = (unsigned)k.nflat;
  unsigned amb_row = (unsigned)k.nflat + 1;
  unsigned bins = (unsigned)k.bins;
  unsigned long long evals = 0;
  unsigned long long nre = 0;
  unsigned long long ndirect = 0;  // Gram path: items run by the direct path
  int qn = 0;  // entries in this wave's recheck queue (wave-uniform)

  // Exact FP64 bins of the newest `nd` queued pairs, one pair per lane (the
  // oracle's arithmetic on the unscaled inputs), deposited like the bulk.
  auto drain = [&](const int nd, const float4* __restrict__ xo,
                   const float4* __restrict__ yo, const int zero) {
    __builtin_amdgcn_wave_barrier();  // the queue was written by other lanes
    if (lane < nd) {
      const int2 e = wq[qn - nd + lane];
      const float4 pa = gld(xo, e.x);
      const float4 pb = gld(yo, e.y & kRecheckIndexMask);
      const unsigned r = exact_row<ANI>(pa, pb, (int)((unsigned)e.y >> kRecheckImgShift), sT,
                                        k.bins, k.R, k.le, zero, k.Lx, out_row);
      if (r < out_row) {
        atomicAdd(hc_lane + (r << CS), 1u);
        if (WEIGHTED) atomicAdd(hw_lane + (r << CS), (double)pa.w * (double)pb.w);
      }
    }
    qn -= nd;
    __builtin_amdgcn_wave_barrier();
  };

  auto run_item = [&](const int it, const float4* __restrict__ xp,
                      const float4* __restrict__ yp, const float4* __restrict__ xp_orig,
                      const float4* __restrict__ yp_orig, const int zero) {
    const int i0 = __builtin_amdgcn_readfirstlane(items[it].i0);
    const int ni = __builtin_amdgcn_readfirstlane(items[it].ni);
    const int j0 = __builtin_amdgcn_readfirstlane(items[it].j0);
    const int nj = __builtin_amdgcn_readfirstlane(items[it].nj);
    const int code = __builtin_amdgcn_readfirstlane(items[it].code);
    const int img = code & 31;
    const bool masked = (code & kCodeMask) != 0;
    // The j range any of the tile's blocks can reach (the union of the step
    // windows, see PairItem): the rest of the run is out of range.
    int jlo = j0, jhi = j0 + nj;
    if (!(code & kCodeNoWin)) {
      const unsigned ws = (unsigned)__builtin_amdgcn_readfirstlane((int)items[it].ws);
      const unsigned we = (unsigned)__builtin_amdgcn_readfirstlane((int)items[it].we);
      int t_first = 0x7fffffff, t_last = 0;
      for (int b = 0; b < 4 && 16 * b < ni; ++b) {
        const int s0 = (int)((ws >> (8 * b)) & 0xffu), e0 = (int)((we >> (8 * b)) & 0xffu);
        if (e0 > s0) {
          t_first = min(t_first, s0);
          t_last = max(t_last, e0);
        }
      }
      jlo = t_first < t_last ? j0 + 16 * t_first : j0;
      jhi = t_first < t_last ? min(j0 + nj, j0 + 16 * t_last) : j0;
    }
    evals += (unsigned long long)ni * (unsigned long long)(jhi - jlo);

    float4 ps, pi;
    if (lane < ni) {
      ps = xp[i0 + lane];
      pi = xp_orig[i0 + lane];
    } else {
      ps = make_float4(1e18f, 1e18f, 1e18f, 0.f);
      pi = ps;
    }
    // Image offset applied once to the (scaled) i point: xi' = xi - k*L.
    const float xs = ps.x - (float)(img / 9 - 1) * k.L;
    const float ys = ps.y - (float)((img / 3) % 3 - 1) * k.L;
    const float zs = ps.z - (float)(img % 3 - 1) * k.L;
    const int my_i = i0 + lane;
    const double wi = (double)pi.w;

    auto deposit = [&](unsigned row, float wj) {
      // 2iso deposits every lane (a branch costs more than the LDS op);
      // 2ani, with most candidates out of range, skips excluded pairs.
      if (!ANI || row < out_row) {
        atomicAdd(hc_lane + (row << CS), 1u);
        if (WEIGHTED) atomicAdd(hw_lane + (row << CS), wi * (double)wj);
      }
    };
    auto fix = [&](unsigned& r, int j) {
      // Lanes whose table cell straddles an edge: the pair goes to the wave's
      // queue (counted out here) and is binned exactly when 64 are queued,
      // one per lane, instead of running the FP64 path for a lane or two.
      const unsigned long long amb = __ballot(r == amb_row);
      if (amb) {
        const int na = __popcll(amb);
        nre += na;
        if (r == amb_row) {
          const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi(
                                   (unsigned)(amb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)amb, 0u));
          wq[pos] = make_int2(my_i, j | (img << kRecheckImgShift));
          r = out_row;
        }
        qn += na;
        if (qn >= kWave) drain(kWave, xp_orig, yp_orig, zero);
      }
    };
    auto quad = [&](int j, const float4 q0, const float4 q1, const float4 q2, const float4 q3) {
      unsigned r0 = pair_row<ANI>(q0, xs, ys, zs, ntm1, tab, bins, out_row, amb_row, k.tflip);
      unsigned r1 = pair_row<ANI>(q1, xs, ys, zs, ntm1, tab, bins, out_row, amb_row, k.tflip);
      unsigned r2 = pair_row<ANI>(q2, xs, ys, zs, ntm1, tab, bins, out_row, amb_row, k.tflip);
      unsigned r3 = pair_row<ANI>(q3, xs, ys, zs, ntm1, tab, bins, out_row, amb_row, k.tflip);
      if (masked) {
        r0 = (j <= my_i) ? out_row : r0;
        r1 = (j + 1 <= my_i) ? out_row : r1;
        r2 = (j + 2 <= my_i) ? out_row : r2;
        r3 = (j + 3 <= my_i) ? out_row : r3;
      }
      const unsigned m = max(max(r0, r1), max(r2, r3));
      if (__builtin_expect(__ballot(m == amb_row) != 0, 0)) {
        fix(r0, j);
        fix(r1, j + 1);
        fix(r2, j + 2);
        fix(r3, j + 3);
      }
      deposit(r0, q0.w);
      deposit(r1, q1.w);
      deposit(r2, q2.w);
      deposit(r3, q3.w);
    };

    int j = jlo;
    const int jend = jhi;
    for (; j + 8 <= jend; j += 8) {
      const float4 q0 = yp[j], q1 = yp[j + 1], q2 = yp[j + 2], q3 = yp[j + 3];
      const float4 q4 = yp[j + 4], q5 = yp[j + 5], q6 = yp[j + 6], q7 = yp[j + 7];
      quad(j, q0, q1, q2, q3);
      quad(j + 4, q4, q5, q6, q7);
    }
    for (; j < jend; ++j) {
      const float4 q = yp[j];
      unsigned r = pair_row<ANI>(q, xs, ys, zs, ntm1, tab, bins, out_row, amb_row, k.tflip);
      if (masked) r = (j <= my_i) ? out_row : r;
      fix(r, j);
      deposit(r, q.w);
    }
  };

  // Gram path: the i side of the wave's current tile -- its origin, whether
  // its box fits the table's error bound, and the B operands and C columns of
  // its four blocks -- kept from one item to the next.  The work list emits a
  // tile's neighbour columns one after another and a wave takes runs of
  // consecutive items, so most items find their tile here and skip the tile
  // box, the i loads and the operand set-up.  Keyed by (i0, ni, image);
  // t_i0 = -1 when the point arrays change (a new chunk of a batch).
  int t_i0 = -1, t_ni = 0, t_img = -1;
  bool t_ok = false;
  float t_ox = 0.f, t_oy = 0.f, t_oz = 0.f;
  float t_b[4], t_c[4];  // 2iso: B, C; 2ani: perpendicular B, line-of-sight B
#pragma unroll
  for (int b = 0; b < 4; ++b) t_b[b] = t_c[b] = 0.f;

  // The i side of a tile (see t_i0 above): false when its box is larger than
  // the table's error bound allows.
  auto tile_setup = [&](const int i0, const int ni, const int img,
                        const float4* __restrict__ xp) -> bool {
    // Image offset on the i side, the same FP32 operation as the direct path.
    const float shx = (float)(img / 9 - 1) * k.L;
    const float shy = (float)((img / 3) % 3 - 1) * k.L;
    const float shz = (float)(img % 3 - 1) * k.L;
    auto uni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
    // Tile box from the grid, with no reduction over the lanes: the tile's
    // points lie in the (cx, cy) column of its first point, between the z
    // cells of its first and last points.  Cell indices are formed as the
    // grid build forms them (the scaled units scale both operands by a power
    // of two, so the same cells result); a point can sit outside its cell by
    // FP32 rounding only, far inside the 2% margin of the bound k.rho2.
    const float4 pf = gld(xp, i0), pe = gld(xp, i0 + ni - 1);
    auto cell = [&](float v, int a) {
      const int c = (int)floorf((uni(v) - k.glo[a]) * k.ginv[a]);
      return min(max(c, 0), k.gP[a] - 1);
    };
    const int cx = cell(pf.x, 0), cy = cell(pf.y, 1), czf = cell(pf.z, 2), czl = cell(pe.z, 2);
    const float ex = k.gcs[0], ey = k.gcs[1], ez = (float)(czl - czf + 1) * k.gcs[2];
    float ox = k.glo[0] + ((float)cx + 0.5f) * k.gcs[0] - shx;
    float oy = k.glo[1] + ((float)cy + 0.5f) * k.gcs[1] - shy;
    float oz = k.glo[2] + 0.5f * (float)(czf + czl + 1) * k.gcs[2] - shz;
    if (!(0.25f * (ex * ex + ey * ey + ez * ez) <= k.rho2)) {
      // Cells too tall for the bound (a sparse column): the points' own box,
      // by wave-wide reductions, may still fit.
      const float4 pl = gld(xp, i0 + min(lane, ni - 1));
      const float tx = pl.x - shx, ty = pl.y - shy, tz = pl.z - shz;
      const float lox = uni(__ockl_wfred_min_f32(tx)), hix = uni(__ockl_wfred_max_f32(tx));
      const float loy = uni(__ockl_wfred_min_f32(ty)), hiy = uni(__ockl_wfred_max_f32(ty));
      const float loz = uni(__ockl_wfred_min_f32(tz)), hiz = uni(__ockl_wfred_max_f32(tz));
      const float bx = hix - lox, by = hiy - loy, bz = hiz - loz;
      if (!(0.25f * (bx * bx + by * by + bz * bz) <= k.rho2)) return false;
      ox = 0.5f * (lox + hix);
      oy = 0.5f * (loy + hiy);
      oz = 0.5f * (loz + hiz);
    }
    t_ox = ox;
    t_oy = oy;
    t_oz = oz;
    if constexpr (!ANI) {
      // i side: B operand (component grp of point i = 16b + col, times 2;
      // grp 3: -1) and the C column Tf - |x_i|^2, so the product is
      // Tf - d2, the flipped table index (beyond the table: negative, cell
      // 0).  Lanes past the tile get a hugely negative C (always out).
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int ir = 16 * b + col;
        const float4 p = gld(xp, i0 + min(ir, ni - 1));
        const float x = (p.x - shx) - ox, y = (p.y - shy) - oy, z = (p.z - shz) - oz;
        // Exact select by 0/1 coefficients (products by 0 and 1 are exact).
        t_b[b] = __builtin_fmaf(sel_x, 2.0f * x,
                                __builtin_fmaf(sel_y, 2.0f * y, __builtin_fmaf(sel_z, 2.0f * z, -sel_w)));
        t_c[b] = ir < ni ? k.tflip - __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x)) : -1e30f;
      }
    } else {
      // i side, 2ani: two products per block, both giving Tf - d2 (the
      // flipped table's index).  Perpendicular: B operand (2x, 2y, -1,
      // Tf - |x_perp|^2) of i against A (x, y, |x_perp|^2, 1) of j; line of
      // sight: B (2z, -1, Tf - z^2, 0) against A (z, z^2, 1, 0).  C = 0 for
      // both (the i side's terms ride in the K dimension).
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int ir = 16 * b + col;
        const float4 p = gld(xp, i0 + min(ir, ni - 1));
        const float x = (p.x - shx) - ox, y = (p.y - shy) - oy, z = (p.z - shz) - oz;
        // A lane past the tile gets hugely negative products (cell 0, out)
        // in both, so neither component can read as ambiguous (its pair
        // does not exist).
        const bool iv = ir < ni;
        const float tp = iv ? k.tflip - __builtin_fmaf(y, y, x * x) : -1e30f;
        const float tz = iv ? k.tflip - z * z : -1e30f;
        t_b[b] = __builtin_fmaf(sel_x, 2.0f * x,
                                 __builtin_fmaf(sel_y, 2.0f * y, __builtin_fmaf(sel_w, tp, -sel_z)));
        t_c[b] = __builtin_fmaf(sel_x, 2.0f * z, __builtin_fmaf(sel_z, tz, -sel_y));
      }
    }
    return true;
  };

  // Gram-form item on the matrix cores (MF = 1).  Returns false, having done
  // nothing, when the tile's bounding box is larger than the error bound of
  // the table covers; the caller then runs the direct path.
  auto run_item_mf = [&](const int it, const float4* __restrict__ xp,
                         const float4* __restrict__ yp, const float4* __restrict__ xp_orig,
                         const float4* __restrict__ yp_orig, const int zero) -> bool {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const int i0 = __builtin_amdgcn_readfirstlane(items[it].i0);
    const int ni = __builtin_amdgcn_readfirstlane(items[it].ni);
    const int j0 = __builtin_amdgcn_readfirstlane(items[it].j0);
    const int nj = __builtin_amdgcn_readfirstlane(items[it].nj);
    const int code = __builtin_amdgcn_readfirstlane(items[it].code);
    const int img = code & 31;
    const bool masked = (code & kCodeMask) != 0;
    if (i0 != t_i0 || ni != t_ni || img != t_img) {
      t_i0 = i0;
      t_ni = ni;
      t_img = img;
      t_ok = tile_setup(i0, ni, img, xp);
    }
    if (!t_ok) return false;
    const float ox = t_ox, oy = t_oy, oz = t_oz;
    // Step windows of the four blocks (see PairItem).
    const bool nowin = (code & kCodeNoWin) != 0;
    const unsigned wsb = nowin ? 0u : (unsigned)__builtin_amdgcn_readfirstlane((int)items[it].ws);
    const unsigned web = nowin ? 0xffffffffu : (unsigned)__builtin_amdgcn_readfirstlane((int)items[it].we);

    const int nb = (ni + 15) >> 4;
    const int jend = j0 + nj;
    // Self-pair mask (items of a cell with itself): only the j steps that
    // start before the tile's end hold pairs with j <= i.  A uniform integer,
    // so the test per step is a scalar compare.
    const int mask_end = masked ? i0 + ni : j0;
    // Block skip: block b takes only the steps of its window (the work list's
    // z-cell reach rule applied to the block's own cells), so the loop runs
    // from the first window's start to the last one's end, and a block
    // outside its window is not evaluated (all its pairs there are out of
    // range).  Windows are in steps of 16 j from j0; without windows every
    // block takes every step.
    auto windows = [&](unsigned* wlo, unsigned* wlen, int* jstart, int* jstop) {
      const int nbk = nb;
      int t_first = 0x7fffffff, t_last = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (b >= nbk) break;
        const unsigned s0 = (wsb >> (8 * b)) & 0xffu;
        const unsigned e0 = nowin ? 0xffffffu : ((web >> (8 * b)) & 0xffu);
        // Uniform: kept in scalar registers, so the per-step test is scalar.
        wlo[b] = (unsigned)__builtin_amdgcn_readfirstlane((int)s0);
        wlen[b] = (unsigned)__builtin_amdgcn_readfirstlane((int)(e0 > s0 ? e0 - s0 : 0u));
        if (e0 > s0) {
          t_first = min(t_first, (int)s0);
          t_last = max(t_last, (int)e0);
        }
      }
      *jstart = t_first < t_last ? j0 + 16 * t_first : j0;
      *jstop = t_first < t_last ? min(jend, j0 + 16 * t_last) : j0;
    };
    // Blocks run at step jb (bit b), and their pairs counted as evaluations
    // (the last block may hold fewer than 16 points).
    auto step_blocks = [&](int jb, const unsigned* wlo, const unsigned* wlen) -> unsigned {
      const unsigned t = (unsigned)((jb - j0) >> 4);
      unsigned act = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (b >= nb) break;
        if (t - wlo[b] < wlen[b]) act |= 1u << b;
      }
      if (act) {
        const unsigned njs = (unsigned)min(16, jend - jb);
        const unsigned last = 1u << (nb - 1);
        unsigned nip = 16u * (unsigned)__builtin_popcount(act & (last - 1u));
        if (act & last) nip += (unsigned)(ni - 16 * (nb - 1));
        evals += (unsigned long long)(njs * nip);
      }
      return act;
    };
    auto fix2 = [&](unsigned& r, bool is_amb, int ii, int jj) {
      const unsigned long long amb = __ballot(is_amb);
      if (amb) {
        const int na = __popcll(amb);
        nre += na;
        if (is_amb) {
          const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi(
                                   (unsigned)(amb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)amb, 0u));
          wq[pos] = make_int2(ii, jj | (img << kRecheckImgShift));
          r = out_row;
        }
        qn += na;
        if (qn >= kWave) drain(kWave, xp_orig, yp_orig, zero);
      }
    };
    if constexpr (!ANI) {
      // The j loop for a tile of NB i-blocks (NB*16 >= ni), 16 j per step.
      auto jloop = [&](auto nbc) {
        constexpr int NB = decltype(nbc)::value;
        unsigned wlo[NB], wlen[NB];
        int jstart, jstop;
        windows(wlo, wlen, &jstart, &jstop);
        // The next step's j point is loaded one step ahead (its latency hides
        // behind this step's matrix, LDS and histogram work).
        const __amdgpu_buffer_rsrc_t jrs = point_rsrc(yp, jend);
        gv3f qn4 = bld3(jrs, 16 * col, jstart);
        for (int jb = jstart; jb < jstop; jb += 16) {
          // j side: A operand (component grp of point j = jb + col; grp 3:
          // |x_j|^2, huge past the run's end).
          const int jj = jb + col;
          const gv3f q = qn4;
          qn4 = bld3(jrs, 16 * col, jb + 16);
          const unsigned act = step_blocks(jb, wlo, wlen);
          if (act == 0) continue;
          const float z = q.z - oz;
          const float x = q.x - ox, y = q.y - oy;
          const float bj = jj < jend ? __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x)) : 1e30f;
          const float aop =
              __builtin_fmaf(sel_x, x, __builtin_fmaf(sel_y, y, __builtin_fmaf(sel_z, z, sel_w * bj)));
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (!(act & (1u << b))) continue;
            const f4v cv = {t_c[b], t_c[b], t_c[b], t_c[b]};
            const f4v d = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, t_b[b], cv, 0, 0, 0);
            unsigned r[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) r[v] = tab[__float2uint_rz(d[v])];
            // Element v of this lane: i = i0 + 16b + col, j = jb + 4*grp + v.
            if (jb < mask_end) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                r[v] = (jb + 4 * grp + v <= i0 + 16 * b + col) ? out_row : r[v];
            }
            const unsigned m = max(max(r[0], r[1]), max(r[2], r[3]));
            if (__builtin_expect(__ballot(m == amb_row) != 0, 0)) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                fix2(r[v], r[v] == amb_row, i0 + 16 * b + col, jb + 4 * grp + v);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) atomicAdd(hc_lane + (r[v] << CS), 1u);
          }
        }
      };
      switch (nb) {
        case 4: jloop(std::integral_constant<int, 4>{}); break;
        case 3: jloop(std::integral_constant<int, 3>{}); break;
        case 2: jloop(std::integral_constant<int, 2>{}); break;
        default: jloop(std::integral_constant<int, 1>{}); break;
      }
    } else {
      // The j loop for a tile of NB i-blocks (NB*16 >= ni), 16 j per step.
      auto jloop = [&](auto nbc) {
        constexpr int NB = decltype(nbc)::value;
        const f4v zero4 = {0.f, 0.f, 0.f, 0.f};
        unsigned wlo[NB], wlen[NB];
        int jstart, jstop;
        windows(wlo, wlen, &jstart, &jstop);
        const __amdgpu_buffer_rsrc_t jrs = point_rsrc(yp, jend);
        gv3f qn4 = bld3(jrs, 16 * col, jstart);
        for (int jb = jstart; jb < jstop; jb += 16) {
          // j side: A operands (component grp of point j = jb + col; a j past
          // the run's end gets huge squared norms in both).
          const int jj = jb + col;
          const gv3f q = qn4;
          qn4 = bld3(jrs, 16 * col, jb + 16);
          const unsigned act = step_blocks(jb, wlo, wlen);
          if (act == 0) continue;
          const float x = q.x - ox, y = q.y - oy, z = q.z - oz;
          const bool jv = jj < jend;
          const float np2 = jv ? __builtin_fmaf(y, y, x * x) : 1e30f;
          const float nz2 = jv ? z * z : 1e30f;
          const float aopp =
              __builtin_fmaf(sel_x, x, __builtin_fmaf(sel_y, y, __builtin_fmaf(sel_z, np2, sel_w)));
          const float aopz = __builtin_fmaf(sel_x, z, __builtin_fmaf(sel_y, nz2, sel_z));
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (!(act & (1u << b))) continue;
            const f4v dp = __builtin_amdgcn_mfma_f32_16x16x4f32(aopp, t_b[b], zero4, 0, 0, 0);
            const f4v dz = __builtin_amdgcn_mfma_f32_16x16x4f32(aopz, t_c[b], zero4, 0, 0, 0);
            // Per element: line-of-sight and perpendicular codes (bin,
            // ambiguous = 254 or out = 255); the row is cz * bins + co when
            // both are bins, and their max is out when either is out.
            unsigned r[4], mc[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const unsigned cz = tab[__float2uint_rz(dz[v])];
              const unsigned co = tab[__float2uint_rz(dp[v])];
              // The row; it counts only when both codes are bins (mc < bins).
              mc[v] = max(cz, co);
              r[v] = __umul24(cz, bins) + co;
            }
            if (jb < mask_end) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (jb + 4 * grp + v <= i0 + 16 * b + col) mc[v] = kAniOut8;
            }
            bool any_amb = false;
#pragma unroll
            for (int v = 0; v < 4; ++v) any_amb = any_amb || mc[v] == kAniAmb8;
            if (__builtin_expect(__ballot(any_amb) != 0, 0)) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                fix2(r[v], mc[v] == kAniAmb8, i0 + 16 * b + col, jb + 4 * grp + v);
            }
            // Most 2ani candidates are out of range, and their deposits would
            // all land on one row per LDS copy (8 lanes per address): only
            // pairs with both codes bins deposit (ambiguous ones go through
            // the queue; excluded self pairs read as out).
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (mc[v] < bins) atomicAdd(hc_lane + (r[v] << CS), 1u);
          }
        }
      };
      switch (nb) {
        case 4: jloop(std::integral_constant<int, 4>{}); break;
        case 3: jloop(std::integral_constant<int, 3>{}); break;
        case 2: jloop(std::integral_constant<int, 2>{}); break;
        default: jloop(std::integral_constant<int, 1>{}); break;
      }
    }
    return true;
  };
  // Sum of the LDS copies of each row into rows [hoff, hoff + nflat) of the
  // global histogram; `clear` zeroes the LDS rows for the next chunk.
  auto flush = [&](int hoff, bool clear) {
    for (int r = threadIdx.x; r < nrows; r += kBlockMixed) {
      unsigned long long c = 0;
      double w = 0.0;
      for (int cpy = 0; cpy < ncopy; ++cpy) {
        c += hc[(r << CS) + cpy];
        if (WEIGHTED) w += hw[(r << CS) + cpy];
        if (clear) {
          hc[(r << CS) + cpy] = 0u;
          if (WEIGHTED) hw[(r << CS) + cpy] = 0.0;
        }
      }
      if (c && r < k.nflat) {
        atomicAdd(&hist[hoff + r], c);
        if (WEIGHTED) atomicAdd(&whist[hoff + r], w);
      }
    }
  };

  if constexpr (!BATCH) {
    // The item list (cell order) is cut into kXcds contiguous ranges, one per
    // XCD: workgroups are dealt to the XCDs round robin, so the waves of
    // workgroup b first drain range b % kXcds (neighbouring cells, one L2),
    // then help the other ranges.  A wave claims `claim` consecutive items per
    // atomic: the claim's round trip (the wave waits on it) is paid once per
    // `claim` items.  counter[r * 16] is range r's cursor (own cache line).
    int nitems, claim;
    launch_items(k, &nitems, &claim);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      stats[2] = (unsigned long long)(k.nitems_dev ? *k.nitems_dev : k.nitems);
    const int home = blockIdx.x % kXcds;
    for (int rr = 0; rr < kXcds; ++rr) {
      const int r = (home + rr) % kXcds;
      const int rb = (int)(((long long)nitems * r) / kXcds);
      const int re = (int)(((long long)nitems * (r + 1)) / kXcds);
      for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(counter + r * 16, claim);
        base = rb + __builtin_amdgcn_readfirstlane(base);
        if (base >= re) break;
        const int end = min(base + claim, re);
        for (int it = base; it < end; ++it) {
          if constexpr (MF) {
            if (run_item_mf(it, xp, yp, xp_orig, yp_orig, k.zero)) continue;
            ++ndirect;
          }
          run_item(it, xp, yp, xp_orig, yp_orig, k.zero);
        }
      }
    }
    if (qn > 0) drain(qn, xp_orig, yp_orig, k.zero);
    __syncthreads();
    flush(0, false);
  } else {
    __shared__ int s_chunk, s_next;
    const uint8_t* loaded = table;  // the table (and rules) now in LDS
    // The chunk list (jobs in order, each in cell order) is cut into kXcds
    // contiguous ranges, one per XCD: workgroups are dealt to the XCDs round
    // robin, so workgroup b first takes the chunks of range b % kXcds
    // (neighbouring cells of one job, whose j columns the XCD's L2 then
    // shares), then helps the other ranges.  counter[r * 16] is range r's
    // cursor; thread 0 keeps the ranges it has found empty.
    const int home = blockIdx.x % kXcds;
    int nspent = 0;
    for (;;) {
      if (threadIdx.x == 0) {
        int ch = -1;
        while (nspent < kXcds) {
          const int r = (home + nspent) % kXcds;
          const int rb = (int)(((long long)k.nitems * r) / kXcds);
          const int re = (int)(((long long)k.nitems * (r + 1)) / kXcds);
          const int v = rb + atomicAdd(counter + r * 16, 1);
          if (v < re) {
            ch = v;
            break;
          }
          ++nspent;
        }
        s_chunk = ch;
        s_next = 0;
      }
      __syncthreads();
      const int c = s_chunk;
      if (c < 0) break;  // every range done (nitems = number of chunks in batch mode)
      const int cb = chunks[c].begin, ce = chunks[c].end;
      const BatchJob jd = jobs[chunks[c].job];
      // The chunk's job parameters (the same for every thread of the
      // workgroup); a job with another table reloads it and its thresholds.
      if (jd.table != loaded) {
        for (int t = threadIdx.x; t < jd.nt; t += kBlockMixed) tab[t] = jd.table[t];
        for (int t = threadIdx.x; t <= jd.bins; t += kBlockMixed) sT[t] = jd.rules->T[t];
        loaded = jd.table;
        __syncthreads();
      }
      k.nt = jd.nt;
      k.bins = jd.bins;
      k.nflat = jd.nflat;
      k.R = jd.rules->R;
      k.le = jd.rules->le;
      k.L = jd.Ls;
      k.tflip = jd.tflip;
      k.rho2 = jd.rho2;
      for (int a = 0; a < 3; ++a) {
        k.glo[a] = jd.glo[a];
        k.gcs[a] = jd.gcs[a];
        k.ginv[a] = jd.ginv[a];
      }
      ntm1 = (unsigned)(k.nt - 1);
      out_row = (unsigned)k.nflat;
      amb_row = (unsigned)k.nflat + 1;
      bins = (unsigned)k.bins;
      t_i0 = -1;  // another job's points (and table) may follow
      for (;;) {
        // Runs of kBatchClaim consecutive items per wave: consecutive items
        // mostly share their tile (the Gram path then keeps its i side).
        int it0 = 0;
        if (lane == 0) it0 = cb + atomicAdd(&s_next, kBatchClaim);
        it0 = __builtin_amdgcn_readfirstlane(it0);
        if (it0 >= ce) break;
        const int it1 = min(it0 + kBatchClaim, ce);
        for (int it = it0; it < it1; ++it) {
          if constexpr (MF) {
            if (run_item_mf(it, jd.xs, jd.ys, jd.xo, jd.yo, jd.zero)) continue;
            ++ndirect;
          }
          run_item(it, jd.xs, jd.ys, jd.xo, jd.yo, jd.zero);
        }
      }
      if (qn > 0) drain(qn, jd.xo, jd.yo, jd.zero);
      __syncthreads();
      flush(jd.hoff, true);
      __syncthreads();
    }
  }
  if (lane == 0) atomicAdd(&stats[0], evals);
  if (lane == 0 && nre) atomicAdd(&stats[1], nre);
  if (MF && lane == 0 && ndirect) atomicAdd(&stats[3], ndirect);
}

// Scaled copy of a point array: x, y, z multiplied by sigma (a power of two,
// so the scaling is exact), w kept.
__global__ void k_scale_points(const float4* __restrict__ in, int64_t n, float sigma,
                               float4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = in[i];
  out[i] = make_float4(p.x * sigma, p.y * sigma, p.z * sigma, p.w);
}

// Plain FP64 path: every pair formed in FP64, the range test and
// bin = (int)(sqrt(d2)*ds) exactly as the reference kernels write them.
__device__ __forceinline__ bool axis_sq64(double ai, double aj, int k, double L, double* out) {
  double d = __dsub_rn(aj, ai);
  if (k == 0) {
    *out = __dmul_rn(d, d);
    return true;
  }
  if ((k > 0 && d >= 0.0) || (k < 0 && d < 0.0)) return false;
  double t = __dsub_rn(L, fabs(d));
  *out = __dmul_rn(t, t);
  return true;
}

// GLOBAL = 1: histograms too large for LDS (the reference's any-size -b):
// every counted pair goes straight to the global histogram with 64-bit
// atomics, as the reference kernels deposit (cuda/kernels/2iso_k.cuh:51-55).
template <int ANI, int WEIGHTED, int GLOBAL>
__global__ __launch_bounds__(kBlock) void k_pairs_fp64(
    const double4* __restrict__ xp, const double4* __restrict__ yp,
    const PairItem* __restrict__ items, int* __restrict__ counter,
    const BinRules* __restrict__ rules, unsigned long long* __restrict__ hist,
    double* __restrict__ whist, unsigned long long* __restrict__ stats, KArgs k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nrows = k.nflat + 1;
  const int ncopy = 1 << k.cshift;
  uint32_t* hc = reinterpret_cast<uint32_t*>(smem);
  double* hw = reinterpret_cast<double*>(smem + ((nrows * ncopy * 4 + 15) & ~15));
  if (!GLOBAL) {
    for (int t = threadIdx.x; t < nrows * ncopy; t += kBlock) {
      hc[t] = 0u;
      if (WEIGHTED) hw[t] = 0.0;
    }
  }
  __syncthreads();
  const double R = rules->R;
  const double ds = rules->ds;
  const double L = rules->L;
  const int le = rules->le;
  const int zero = rules->zero;
  const int bins = k.bins;
  const int lane = threadIdx.x & (kWave - 1);
  const int copy = lane & (ncopy - 1);
  unsigned long long evals = 0;
  // The same XCD-aware ranges and multi-item claims as the mixed kernel.
  int nitems, claim;
  launch_items(k, &nitems, &claim);
  if (blockIdx.x == 0 && threadIdx.x == 0)
      stats[2] = (unsigned long long)(k.nitems_dev ? *k.nitems_dev : k.nitems);
  const int home = blockIdx.x % kXcds;
  int rr = 0, r = home;
  int rb = (int)(((long long)nitems * r) / kXcds);
  int re = (int)(((long long)nitems * (r + 1)) / kXcds);
  int it = 0, claim_end = 0;
  for (;; ++it) {
    while (it >= claim_end) {
      int base = 0;
      if (lane == 0) base = atomicAdd(counter + r * 16, claim);
      it = rb + __builtin_amdgcn_readfirstlane(base);
      claim_end = min(it + claim, re);
      if (it < re) break;
      if (++rr == kXcds) break;  // every range drained
      r = (home + rr) % kXcds;
      rb = (int)(((long long)nitems * r) / kXcds);
      re = (int)(((long long)nitems * (r + 1)) / kXcds);
    }
    if (rr == kXcds) break;
    const int i0 = __builtin_amdgcn_readfirstlane(items[it].i0);
    const int ni = __builtin_amdgcn_readfirstlane(items[it].ni);
    const int j0 = __builtin_amdgcn_readfirstlane(items[it].j0);
    const int nj = __builtin_amdgcn_readfirstlane(items[it].nj);
    const int code = __builtin_amdgcn_readfirstlane(items[it].code);
    const int img = code & 31;
    const int kx = img / 9 - 1, ky = (img / 3) % 3 - 1, kz = img % 3 - 1;
    const bool masked = (code & kCodeMask) != 0;
    evals += (unsigned long long)ni * (unsigned long long)nj;
    const bool active = lane < ni;
    double4 pi = active ? xp[i0 + lane] : make_double4(0, 0, 0, 0);
    const int my_i = i0 + lane;
    for (int j = j0; j < j0 + nj; ++j) {
      const double4 q = yp[j];
      double tx, ty, tz;
      bool ok = active && !(masked && j <= my_i);
      ok = ok && axis_sq64(pi.x, q.x, kx, L, &tx);
      ok = ok && axis_sq64(pi.y, q.y, ky, L, &ty);
      ok = ok && axis_sq64(pi.z, q.z, kz, L, &tz);
      int row = -1;
      if (ok) {
        if (!ANI) {
          const double d2 = __dadd_rn(__dadd_rn(tx, ty), tz);
          const bool in = le ? (d2 <= R) : (d2 < R);
          if (in && !(zero == 1 && img == kImgZero && !(d2 > 0.0))) {
            const int b = (int)(sqrt(d2) * ds);
            if (b <= bins - 1) row = b;
          }
        } else {
          const double do2 = __dadd_rn(tx, ty);
          const bool in = le ? (tz <= R && do2 <= R) : (tz < R && do2 < R);
          if (in && !(zero == 2 && !(tz > 0.0 && do2 > 0.0))) {
            const int bz = (int)(sqrt(tz) * ds);
            const int bo = (int)(sqrt(do2) * ds);
            if (bz <= bins - 1 && bo <= bins - 1) row = bz * bins + bo;
          }
        }
      }
      if (GLOBAL) {
        if (row >= 0) {
          atomicAdd(&hist[row], 1ull);
          if (WEIGHTED) atomicAdd(&whist[row], pi.w * q.w);
        }
        continue;
      }
      const unsigned r = row >= 0 ? (unsigned)row : (unsigned)k.nflat;
      atomicAdd(&hc[(r << k.cshift) + copy], 1u);
      if (WEIGHTED) atomicAdd(&hw[(r << k.cshift) + copy], pi.w * q.w);
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; !GLOBAL && r < k.nflat; r += kBlock) {
    unsigned long long c = 0;
    double w = 0.0;
    for (int cpy = 0; cpy < ncopy; ++cpy) {
      c += hc[(r << k.cshift) + cpy];
      if (WEIGHTED) w += hw[(r << k.cshift) + cpy];
    }
    if (c) {
      atomicAdd(&hist[r], c);
      if (WEIGHTED) atomicAdd(&whist[r], w);
    }
  }
  if (lane == 0) atomicAdd(&stats[0], evals);
}

template <class K>
hipError_t set_lds(K kern, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace

hipError_t launch_scale_points(const float4* in, int64_t n, float sigma, float4* out,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, n,
                     sigma, out);
  return hipGetLastError();
}

hipError_t build_pair_worklist(const WorkListArgs& a, DevBuf& counts, DevBuf& items, int* nitems,
                               hipStream_t s) {
  hipError_t e;
  const int S = worklist_split(a, a.g.ncell());
  const int64_t nc = a.g.ncell() * S;
  const size_t ibytes = (sizeof(int) * (size_t)(2 * nc + 2) + 15) & ~(size_t)15;
  if ((e = counts.ensure(ibytes + sizeof(long long) * (size_t)scan_tiles(nc))) != hipSuccess)
    return e;
  int* cnt = counts.as<int>();
  int* off = cnt + nc + 1;
  long long* tsum = reinterpret_cast<long long*>(counts.as<char>() + ibytes);
  const int blocks = (int)((nc + 255) / 256);
  hipLaunchKernelGGL(k_worklist_count, dim3(blocks), dim3(256), 0, s, a, S, cnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan(cnt, nc, off, tsum, s)) != hipSuccess) return e;
  int total = 0;
  if ((e = hipMemcpyAsync(&total, off + nc, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (total < 0) return hipErrorInvalidValue;  // 2^31 items or more
  *nitems = total;
  if (total == 0) return hipSuccess;
  if ((e = items.ensure(sizeof(PairItem) * (size_t)total)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_worklist_fill, dim3(blocks), dim3(256), 0, s, a, S, off,
                     items.as<PairItem>(), (int64_t)total);
  return hipGetLastError();
}

int64_t worklist_bound(const WorkListArgs& a, int64_t n_tile_points) {
  const int64_t tiles = (n_tile_points + kTile - 1) / kTile + (int64_t)a.g.P[0] * a.g.P[1];
  const int64_t cols = (int64_t)(2 * a.K[0] + 1) * (2 * a.K[1] + 1);
  return tiles * cols * (a.g.bpc ? 3 : 1);
}

size_t worklist_count_bytes(const WorkListArgs& a) {
  const int S = worklist_split(a, a.g.ncell());
  const int64_t nc = a.g.ncell() * S;
  const size_t ibytes = (sizeof(int) * (size_t)(2 * nc + 2) + 15) & ~(size_t)15;
  return ibytes + sizeof(long long) * (size_t)scan_tiles(nc);
}

hipError_t build_pair_worklist_into(const WorkListArgs& a, int64_t cap, void* counts,
                                    PairItem* items, const int** total_dev, hipStream_t s) {
  hipError_t e;
  const int S = worklist_split(a, a.g.ncell());
  const int64_t nc = a.g.ncell() * S;
  const size_t ibytes = (sizeof(int) * (size_t)(2 * nc + 2) + 15) & ~(size_t)15;
  int* cnt = reinterpret_cast<int*>(counts);
  int* off = cnt + nc + 1;
  long long* tsum = reinterpret_cast<long long*>(reinterpret_cast<char*>(counts) + ibytes);
  const int blocks = (int)((nc + 255) / 256);
  hipLaunchKernelGGL(k_worklist_count, dim3(blocks), dim3(256), 0, s, a, S, cnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan(cnt, nc, off, tsum, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_worklist_fill, dim3(blocks), dim3(256), 0, s, a, S, off, items, cap);
  *total_dev = off + nc;
  return hipGetLastError();
}

hipError_t build_pair_worklist_async(const WorkListArgs& a, int64_t cap, DevBuf& counts,
                                     DevBuf& items, const int** total_dev, hipStream_t s) {
  hipError_t e;
  if ((e = counts.ensure(worklist_count_bytes(a))) != hipSuccess) return e;
  if ((e = items.ensure(sizeof(PairItem) * (size_t)std::max<int64_t>(cap, 1))) != hipSuccess)
    return e;
  return build_pair_worklist_into(a, cap, counts.p, items.as<PairItem>(), total_dev, s);
}

__global__ void k_gather_offsets(const int* __restrict__ off, int64_t stride, int njobs,
                                 int* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= njobs) out[k] = off[(int64_t)k * stride];
}

// Work lists of several jobs over one geometry, concatenated in job order:
// one count pass per job, one scan over all (job, cell) counts, one fill pass
// per job.  job_offsets[k] is the first item of job k, job_offsets[njobs] the
// total (one synchronous read of njobs + 1 integers).
hipError_t build_batch_worklists(const std::vector<WorkListArgs>& args, DevBuf& counts,
                                 DevBuf& items, std::vector<int>* job_offsets, hipStream_t s) {
  hipError_t e;
  const int nj = (int)args.size();
  // One split for all jobs (their counts share one scan); jobs of a batch
  // share the neighbour reach, else every cell takes one thread.
  bool same_k = true;
  for (const WorkListArgs& w : args)
    for (int a = 0; a < 3; ++a) same_k = same_k && w.K[a] == args[0].K[a];
  const int S = same_k ? worklist_split(args[0], args[0].g.ncell() * nj) : 1;
  const int64_t nc = args[0].g.ncell() * S;
  const int64_t m = nc * nj;
  if (m + 2 > INT32_MAX) return hipErrorInvalidValue;
  const size_t ibytes = (sizeof(int) * (size_t)(2 * m + nj + 4) + 15) & ~(size_t)15;
  if ((e = counts.ensure(ibytes + sizeof(long long) * (size_t)scan_tiles(m))) != hipSuccess)
    return e;
  int* cnt = counts.as<int>();
  int* off = cnt + m + 1;
  int* jo = off + m + 1;
  long long* tsum = reinterpret_cast<long long*>(counts.as<char>() + ibytes);
  const int blocks = (int)((nc + 255) / 256);
  for (int k = 0; k < nj; ++k) {
    hipLaunchKernelGGL(k_worklist_count, dim3(blocks), dim3(256), 0, s, args[k], S, cnt + k * nc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if ((e = launch_scan(cnt, m, off, tsum, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather_offsets, dim3((nj + 256) / 256), dim3(256), 0, s, off, nc, nj, jo);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  job_offsets->assign(nj + 1, 0);
  if ((e = hipMemcpyAsync(job_offsets->data(), jo, sizeof(int) * (nj + 1), hipMemcpyDeviceToHost,
                          s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const int total = (*job_offsets)[nj];
  if (total < 0) return hipErrorInvalidValue;  // 2^31 items or more
  if (total == 0) return hipSuccess;
  if ((e = items.ensure(sizeof(PairItem) * (size_t)total)) != hipSuccess) return e;
  for (int k = 0; k < nj; ++k) {
    hipLaunchKernelGGL(k_worklist_fill, dim3(blocks), dim3(256), 0, s, args[k], S, off + k * nc,
                       items.as<PairItem>(), (int64_t)total);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_pairs(const PairLaunch& pl, hipStream_t s) {
  KArgs k;
  k.nitems = pl.batch ? pl.nchunks : pl.nitems;
  k.nt = pl.nt;
  k.inv_h = pl.inv_h;
  k.L = pl.L;
  k.bins = pl.bins;
  k.nflat = pl.nflat;
  k.cshift = pl.cshift;
  k.R = pl.R;
  k.Lx = pl.Lx;
  k.le = pl.le;
  k.zero = pl.zero;
  // Claim granularity: derived in the kernel from the item count.
  k.claim = 0;
  k.nitems_dev = pl.batch ? nullptr : pl.nitems_dev;
  k.cap = pl.nitems_cap;
  k.rho2 = pl.rho2;
  k.tflip = pl.tflip;
  for (int a = 0; a < 3; ++a) {
    // Scaling by sigma (a power of two) is exact.
    k.glo[a] = pl.geo.lo[a] * pl.sigma;
    k.gcs[a] = pl.geo.cs[a] * pl.sigma;
    k.ginv[a] = pl.geo.inv_cs[a] / pl.sigma;
    k.gP[a] = pl.geo.P[a];
  }
  const int ncopy = 1 << pl.cshift;
  dim3 grid(pl.nblocks), block(kBlock);
  hipError_t e;
  if (!pl.plain64) {
    if (pl.nt > kMaxTable) return hipErrorInvalidValue;
    const size_t hc_bytes = (size_t)(((pl.nflat + 2) * ncopy * 4 + 15) & ~15);
    const size_t t_bytes = (size_t)((pl.bins + 2) & ~1) * 8;
    const size_t hw_bytes = pl.weighted ? (size_t)(pl.nflat + 2) * ncopy * 8 : 0;
    const size_t lds = hc_bytes + t_bytes + hw_bytes + kRecheckQueueBytes;
    const dim3 mblock(kBlockMixed);
#define PCF_LAUNCH_MIXED2(A, W, C, B, M)                                                          \
  do {                                                                                            \
    if ((e = set_lds(k_pairs_mixed<A, W, C, B, M>, lds)) != hipSuccess) return e;                 \
    if (pl.prepare_only) return hipSuccess;                                                       \
    hipLaunchKernelGGL((k_pairs_mixed<A, W, C, B, M>), grid, mblock, lds, s, pl.xs, pl.ys, pl.xp, \
                       pl.yp, pl.items, pl.counter, pl.table, pl.rules, pl.hist, pl.whist,        \
                       pl.stats, k, pl.jobs, pl.chunks);                                          \
  } while (0)
#define PCF_LAUNCH_MIXED(A, W, C)                                       \
  do {                                                                  \
    const bool mf_ = W == 0 && pl.mfma;                                 \
    if (pl.batch) {                                                     \
      if (mf_) PCF_LAUNCH_MIXED2(A, 0, C, 1, 1);                        \
      else PCF_LAUNCH_MIXED2(A, W, C, 1, 0);                            \
    } else if (mf_) PCF_LAUNCH_MIXED2(A, 0, C, 0, 1);                   \
    else PCF_LAUNCH_MIXED2(A, W, C, 0, 0);                              \
  } while (0)
#define PCF_LAUNCH_CS(A, W)                                   \
  switch (pl.cshift) {                                        \
    case 6: PCF_LAUNCH_MIXED(A, W, 6); break;                 \
    case 5: PCF_LAUNCH_MIXED(A, W, 5); break;                 \
    case 4: PCF_LAUNCH_MIXED(A, W, 4); break;                 \
    case 3: PCF_LAUNCH_MIXED(A, W, 3); break;                 \
    case 2: PCF_LAUNCH_MIXED(A, W, 2); break;                 \
    case 1: PCF_LAUNCH_MIXED(A, W, 1); break;                 \
    default: PCF_LAUNCH_MIXED(A, W, 0); break;                \
  }
    if (!pl.ani && !pl.weighted) { PCF_LAUNCH_CS(0, 0) }
    else if (!pl.ani && pl.weighted) { PCF_LAUNCH_CS(0, 1) }
    else if (pl.ani && !pl.weighted) { PCF_LAUNCH_CS(1, 0) }
    else { PCF_LAUNCH_CS(1, 1) }
#undef PCF_LAUNCH_CS
#undef PCF_LAUNCH_MIXED
#undef PCF_LAUNCH_MIXED2
  } else {
    const size_t hc_bytes = (size_t)(((pl.nflat + 1) * ncopy * 4 + 15) & ~15);
    const size_t hw_bytes = pl.weighted ? (size_t)(pl.nflat + 1) * ncopy * 8 : 0;
    const size_t lds = pl.global_hist ? 0 : hc_bytes + hw_bytes;
#define PCF_LAUNCH_64(A, W, G)                                                                    \
  do {                                                                                            \
    if ((e = set_lds(k_pairs_fp64<A, W, G>, lds)) != hipSuccess) return e;                        \
    hipLaunchKernelGGL((k_pairs_fp64<A, W, G>), grid, block, lds, s, pl.xp64, pl.yp64, pl.items,  \
                       pl.counter, pl.rules, pl.hist, pl.whist, pl.stats, k);                     \
  } while (0)
#define PCF_LAUNCH_64G(A, W)                 \
  do {                                       \
    if (pl.global_hist) PCF_LAUNCH_64(A, W, 1); \
    else PCF_LAUNCH_64(A, W, 0);             \
  } while (0)
    if (!pl.ani && !pl.weighted) PCF_LAUNCH_64G(0, 0);
    else if (!pl.ani && pl.weighted) PCF_LAUNCH_64G(0, 1);
    else if (pl.ani && !pl.weighted) PCF_LAUNCH_64G(1, 0);
    else PCF_LAUNCH_64G(1, 1);
#undef PCF_LAUNCH_64G
#undef PCF_LAUNCH_64
  }
  return hipGetLastError();
}

}  // namespace pcf
