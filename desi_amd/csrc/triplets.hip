// triplets.hip -- 3iso / 3ani triplet histograms on gfx950.
//
// Reference behaviour restated (cuda/kernels/3iso_k.cuh:8-180,
// 3isoBPC_k.cuh:8-1692, 3ani_k.cuh:8-209, 3aniBPC_k.cuh:8-1709): every ordered
// triplet (1, 2, 3) with all three sides in range adds w1*w2*w3 to bin
// [b(d12)][b(d23)][b(d31)] (3iso) or [b(d12)][b(d31)][b(d23)][t][p] (3ani, t and
// p the binned cosines of the 1-2 and 1-3 directions with the line of sight z).
// The reference runs one thread per ordered node triple with a triple point
// loop.  Here:
//
//   1. pair list: the pair work list over X emits every pair within range with
//      its exact FP64 bin b12 (and, for 3ani, the exact cosine bin t), the
//      periodic image of point 2 relative to point 1 and the two points'
//      coordinates pre-shifted into point 1's frame; records are grouped by
//      point 1 in cell order (count, scan, scatter);
//   2. triplet kernel: a wave owns one point 1; its third points are binned
//      one per lane against point 1 (the 1-3 side) and the ones in range are
//      listed in LDS, then the wave's lanes take point 1's pairs and walk the
//      list (the 2-3 side).  Both sides go through the FP32 d2 -> bin table;
//      lanes in an ambiguous cell (and 3ani's cosine bins) take the exact FP64
//      path;
//   3. 3iso: pairs are unordered (i < j).  XXY: the third point is any point
//      of Y and the ordered histogram is H[a][b][c] = V[a][b][c] + V[a][c][b]
//      (exchanging points 1 and 2 swaps the d23 and d31 sides).  XXX: the
//      third point has k > i as well, so point 1 is the lowest of the three
//      and an unordered triplet is met twice; XXX launches take one b12 row
//      each and only third points with b31 >= b12, so just the meeting(s)
//      with b12 <= b31 are deposited, and k_sym_cyc rebuilds the six orders.
//      3ani: the
//      cosine of the 1-2 direction is not symmetric, so pairs are emitted in
//      both orders and every third point is taken.
//
// 3iso accumulates in LDS (u32 counts, FP64 weights) over slices of b12 so the
// slice fits (or, when one b12 row does not, in global memory); 3ani's 5-D
// histogram (bins^5) is privatised in LDS per (b12, t) slice where one t value
// of a b12 row fits, and goes to global atomics otherwise.
#include <cstring>

#include "exact.h"
#include "pcf_ctx.h"
#include "worklist.h"

namespace pcf {

namespace {

constexpr int kMaxTripBins = 253;
constexpr int kTripBlock = 1024;    // 3iso: one workgroup per CU owns the LDS histogram
constexpr int kAniTripBlock = 256;  // 3ani: global histogram, many workgroups per CU
constexpr int kEmitBlock = 256;

// One pair of the pair list, held by one lane.  The list is grouped by
// (point 1, b12): key = point1 * bins + b12 (counting sort), so the pairs of
// one point 1 and one b12 slice are contiguous.
struct PairRec {
  float xj, yj, zj;  // point 2 scaled by sigma, in point 1's frame (image applied)
  float wj;          // weight of point 2
  int j;             // cell-ordered index of point 2 in the X grid
  int meta;          // image code of point 2 (bits 0-4) | b12 << 8 | t << 20 (3ani)
};
static_assert(sizeof(PairRec) == 24, "pair record layout");

struct EmitArgs {
  int nitems;
  int nt;
  float tf;       // the flipped tables' top: cell cvt_rz(tf - d2) holds d2
  float Ls;       // box side times sigma
  int bins;
  int bpc;
  int ani;
  double R;
  double L;
  double ds_th;
  int key_t;      // 3ani list: key = point1 * bins^2 + b12 * bins + t (else point1 * bins + b12)
};

__device__ __forceinline__ int img_neg(int img) { return 26 - img; }

// Exact b12 of pair (p1, p2) with point 2 at image `img` relative to point 1,
// or -1 when the pair is out of range (triplet pairs need 0 < d12,
// 3iso_k.cuh:58, 3isoBPC_k.cuh:62).  For 3ani, *t_out receives the cosine bin
// t = (int)((1 + dz12/d12) * bn/2) (3ani_k.cuh:75-90; periodic dz12 =
// |dz| - L*f, 3aniBPC_k.cuh:57), or -1 when t > bn-1.
__device__ __forceinline__ int exact_pair(const float4 p1, const float4 p2, int img, int ani,
                                          int bpc, const double* T, int bins, double R, double L,
                                          double ds_th, int* t_out) {
  const int f2 = image_flags(p1, p2, img, false);
  if (f2 < 0) return -1;
  const float a1[3] = {p1.x, p1.y, p1.z}, a2[3] = {p2.x, p2.y, p2.z};
  double t[3];
  for (int a = 0; a < 3; ++a) {
    const double d = __dsub_rn((double)a2[a], (double)a1[a]);
    t[a] = bpc ? proj_term(d, L, (f2 >> a) & 1) : d;
  }
  const double s = __dadd_rn(__dadd_rn(__dmul_rn(t[0], t[0]), __dmul_rn(t[1], t[1])),
                             __dmul_rn(t[2], t[2]));
  if (!(s > 0.0)) return -1;
  const int b = class_lds(s, T, bins, R, 0);
  if (b < 0) return -1;
  if (ani) {
    const double cth = __dadd_rn(1.0, t[2] / sqrt(s));
    const int tb = (int)__dmul_rn(cth, ds_th);
    *t_out = tb > bins - 1 ? -1 : tb;
  }
  return b;
}

template <int PASS>
__global__ __launch_bounds__(kEmitBlock) void k_emit_pairs(
    const float4* __restrict__ xs, const float4* __restrict__ xo,
    const PairItem* __restrict__ items, const uint8_t* __restrict__ table,
    const double* __restrict__ T, int* __restrict__ cnt, const int* __restrict__ off,
    int* __restrict__ cursor, PairRec* __restrict__ out, EmitArgs e) {
  __shared__ double sT[kMaxTripBins + 2];
  __shared__ uint8_t tab[kMixedTableMax];  // the bin table (e.nt <= kMixedTableMax bytes)
  for (int t = threadIdx.x; t <= e.bins; t += kEmitBlock) sT[t] = T[t];
  for (int t = threadIdx.x; t < e.nt; t += kEmitBlock) tab[t] = table[t];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = (int)((blockIdx.x * (unsigned)kEmitBlock + threadIdx.x) >> 6);
  const int nwaves = (int)((gridDim.x * (unsigned)kEmitBlock) >> 6);
  const unsigned bins = (unsigned)e.bins;
  const unsigned amb_row = bins + 1;
  for (int it = wave; it < e.nitems; it += nwaves) {
    const int i0 = __builtin_amdgcn_readfirstlane(items[it].i0);
    const int ni = __builtin_amdgcn_readfirstlane(items[it].ni);
    const int j0 = __builtin_amdgcn_readfirstlane(items[it].j0);
    const int nj = __builtin_amdgcn_readfirstlane(items[it].nj);
    const int code = __builtin_amdgcn_readfirstlane(items[it].code);
    const int img = code & 31;
    const bool masked = (code & kCodeMask) != 0;
    const bool active = lane < ni;
    const int my_i = i0 + lane;
    float4 ps = make_float4(1e18f, 1e18f, 1e18f, 0.f), po = ps;
    if (active) {
      ps = xs[my_i];
      po = xo[my_i];
    }
    const float kxL = (float)(img / 9 - 1) * e.Ls, kyL = (float)((img / 3) % 3 - 1) * e.Ls,
                kzL = (float)(img % 3 - 1) * e.Ls;
    const float sx = ps.x - kxL, sy = ps.y - kyL, sz = ps.z - kzL;
    for (int j = j0; j < j0 + nj; ++j) {
      const float4 q = xs[j];
      const float dx = q.x - sx, dy = q.y - sy, dz = q.z - sz;
      const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
      unsigned row = tab[__float2uint_rz(e.tf - d2)];
      if (masked && j <= my_i) row = bins;
      if (!active) row = bins;
      // Everything emitted gets the exact bin: the pairs are few compared
      // with the triplets they feed.
      if (row < bins || row == amb_row) {
        const float4 qo = xo[j];
        int t12 = 0, t21 = -1;
        const int b12 = exact_pair(po, qo, img, e.ani, e.bpc, sT, e.bins, e.R, e.L, e.ds_th, &t12);
        if (b12 >= 0) {
          const bool fwd = !e.ani || t12 >= 0;
          // the same pair seen from point 2 (3ani needs both orders)
          if (e.ani &&
              exact_pair(qo, po, img_neg(img), 1, e.bpc, sT, e.bins, e.R, e.L, e.ds_th, &t21) < 0)
            t21 = -1;
          const int key_i = e.key_t ? (my_i * e.bins + b12) * e.bins + (t12 < 0 ? 0 : t12)
                                    : my_i * e.bins + b12;
          const int key_j = e.key_t ? (j * e.bins + b12) * e.bins + (t21 < 0 ? 0 : t21)
                                    : j * e.bins + b12;
          if (PASS == 0) {
            if (fwd) atomicAdd(&cnt[key_i], 1);
            if (e.ani && t21 >= 0) atomicAdd(&cnt[key_j], 1);
          } else {
            // Both passes take identical decisions, so a slot never passes
            // the end of its key's range; the bound is kept as a guard.
            if (fwd) {
              const int slot = off[key_i] + atomicAdd(&cursor[key_i], 1);
              if (slot < off[key_i + 1]) {
                PairRec r;
                r.xj = q.x + kxL; r.yj = q.y + kyL; r.zj = q.z + kzL;
                r.wj = qo.w;
                r.j = j;
                r.meta = img | (b12 << 8) | ((e.ani ? t12 : 0) << 20);
                out[slot] = r;
              }
            }
            if (e.ani && t21 >= 0) {
              const int slot = off[key_j] + atomicAdd(&cursor[key_j], 1);
              if (slot < off[key_j + 1]) {
                PairRec r;
                r.xj = sx; r.yj = sy; r.zj = sz;
                r.wj = po.w;
                r.j = my_i;
                r.meta = img_neg(img) | (b12 << 8) | (t21 << 20);
                out[slot] = r;
              }
            }
          }
        }
      }
    }
  }
}

// Exact 1-3 side of a cross pair (point 3 from the other catalog): bin of d31
// and the cosine bin p of the 1-3 direction, or -1.  XXY takes d31 = 0 (the
// cosine is then NaN and lands in bin 0, 3ani_k.cuh:150-170 under the PTX
// NaN -> 0 conversion).
__device__ __forceinline__ int exact_cross(const float4 p1, const float4 p3, int img, int bpc,
                                           const double* T, int bins, double R, double L,
                                           double ds_th, int* p_out) {
  const int f3 = image_flags(p1, p3, img, false);
  if (f3 < 0) return -1;
  const float a1[3] = {p1.x, p1.y, p1.z}, a3[3] = {p3.x, p3.y, p3.z};
  double t[3];
  for (int a = 0; a < 3; ++a) {
    const double d = __dsub_rn((double)a3[a], (double)a1[a]);
    t[a] = bpc ? proj_term(d, L, (f3 >> a) & 1) : d;
  }
  const double s = __dadd_rn(__dadd_rn(__dmul_rn(t[0], t[0]), __dmul_rn(t[1], t[1])),
                             __dmul_rn(t[2], t[2]));
  const int b = class_lds(s, T, bins, R, 0);
  if (b < 0) return -1;
  const double pp = __dmul_rn(__dadd_rn(1.0, t[2] / sqrt(s)), ds_th);
  const int pb = isnan(pp) ? 0 : (int)pp;
  if (pb > bins - 1) return -1;
  *p_out = pb;
  return b;
}

// The third-point list of 3ani XXY: every (point 1 of X, point 3 of K) in
// range, keyed by point 1, in the PairRec layout (point 3 in point 1's frame,
// meta = image | b31 << 8 | p << 20).
template <int PASS>
__global__ __launch_bounds__(kEmitBlock) void k_emit_cross(
    const float4* __restrict__ xs, const float4* __restrict__ xo,
    const float4* __restrict__ ys, const float4* __restrict__ yo,
    const PairItem* __restrict__ items, const uint8_t* __restrict__ table,
    const double* __restrict__ T, int* __restrict__ cnt, const int* __restrict__ off,
    int* __restrict__ cursor, PairRec* __restrict__ out, EmitArgs e) {
  __shared__ double sT[kMaxTripBins + 2];
  __shared__ uint8_t tab[kMixedTableMax];  // the bin table (e.nt <= kMixedTableMax bytes)
  for (int t = threadIdx.x; t <= e.bins; t += kEmitBlock) sT[t] = T[t];
  for (int t = threadIdx.x; t < e.nt; t += kEmitBlock) tab[t] = table[t];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = (int)((blockIdx.x * (unsigned)kEmitBlock + threadIdx.x) >> 6);
  const int nwaves = (int)((gridDim.x * (unsigned)kEmitBlock) >> 6);
  const unsigned bins = (unsigned)e.bins;
  const unsigned amb_row = bins + 1;
  for (int it = wave; it < e.nitems; it += nwaves) {
    const int i0 = __builtin_amdgcn_readfirstlane(items[it].i0);
    const int ni = __builtin_amdgcn_readfirstlane(items[it].ni);
    const int j0 = __builtin_amdgcn_readfirstlane(items[it].j0);
    const int nj = __builtin_amdgcn_readfirstlane(items[it].nj);
    const int img = __builtin_amdgcn_readfirstlane(items[it].code) & 31;
    const bool active = lane < ni;
    const int my_i = i0 + lane;
    float4 ps = make_float4(1e18f, 1e18f, 1e18f, 0.f), po = ps;
    if (active) {
      ps = xs[my_i];
      po = xo[my_i];
    }
    const float kxL = (float)(img / 9 - 1) * e.Ls, kyL = (float)((img / 3) % 3 - 1) * e.Ls,
                kzL = (float)(img % 3 - 1) * e.Ls;
    const float sx = ps.x - kxL, sy = ps.y - kyL, sz = ps.z - kzL;
    for (int j = j0; j < j0 + nj; ++j) {
      const float4 q = ys[j];
      const float dx = q.x - sx, dy = q.y - sy, dz = q.z - sz;
      const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
      unsigned row = tab[__float2uint_rz(e.tf - d2)];
      if (!active) row = bins;
      if (row < bins || row == amb_row) {
        const float4 qo = yo[j];
        int pb = 0;
        const int b31 = exact_cross(po, qo, img, e.bpc, sT, e.bins, e.R, e.L, e.ds_th, &pb);
        if (b31 >= 0) {
          if (PASS == 0) {
            atomicAdd(&cnt[my_i], 1);
          } else {
            const int slot = off[my_i] + atomicAdd(&cursor[my_i], 1);
            if (slot < off[my_i + 1]) {
              PairRec r;
              r.xj = q.x + kxL; r.yj = q.y + kyL; r.zj = q.z + kzL;
              r.wj = qo.w;
              r.j = j;
              r.meta = img | (b31 << 8) | (pb << 20);
              out[slot] = r;
            }
          }
        }
      }
    }
  }
}

struct TripArgs {
  int nx;         // points of X (point-1 tasks)
  int nt;
  float tf;       // the flipped tables' top: cell cvt_rz(tf - d2) holds d2
  float Ls;
  int bins;
  int auto_;      // XXX: all sides > 0, point 3 is neither point 1 nor point 2
  int bpc;
  double R, L, ds_th;
  int lo, hi;     // b12 slice
  int glob;       // 3iso: V rows in global memory (one b12 row exceeds LDS)
  int cs;         // 3iso LDS slice: 2^cs interleaved copies of each V row (lane & (2^cs - 1))
  float reach;    // rmax plus slack (unscaled)
  int K[3];       // neighbour reach in cells
  GridGeom g;
};

__device__ __forceinline__ int cell_axis_dev(float v, float lo, float inv_cs, int P) {
  int c = (int)floorf((v - lo) * inv_cs);
  c = c < 0 ? 0 : c;
  return c >= P ? P - 1 : c;
}

// Calls f(k0, k1, sx, sy, sz) for every contiguous range [k0, k1) of the
// K grid that may hold a point within `reach` of point p, with the periodic
// image s (the range's points are seen at k + s*L).  Columns are pruned by
// the lateral distance from p, z-ranges by the remaining reach.
template <class F>
__device__ __forceinline__ void for_each_segment(const float4 p, const TripArgs& a,
                                                 const int* __restrict__ cs_k, F f) {
  const GridGeom& g = a.g;
  const int Px = g.P[0], Py = g.P[1], Pz = g.P[2];
  const int cx = cell_axis_dev(p.x, g.lo[0], g.inv_cs[0], Px);
  const int cy = cell_axis_dev(p.y, g.lo[1], g.inv_cs[1], Py);
  const int cz = cell_axis_dev(p.z, g.lo[2], g.inv_cs[2], Pz);
  const float r2 = a.reach * a.reach;
  for (int dx = -a.K[0]; dx <= a.K[0]; ++dx) {
    const int vx = cx + dx;
    int ccx = vx, sx = 0;
    if (ccx < 0) {
      if (!g.bpc) continue;
      ccx += Px;
      sx = -1;
    } else if (ccx >= Px) {
      if (!g.bpc) continue;
      ccx -= Px;
      sx = 1;
    }
    const float x0 = g.lo[0] + (float)vx * g.cs[0];
    const float gx = fmaxf(0.f, fmaxf(x0 - p.x, p.x - (x0 + g.cs[0])));
    for (int dy = -a.K[1]; dy <= a.K[1]; ++dy) {
      const int vy = cy + dy;
      int ccy = vy, sy = 0;
      if (ccy < 0) {
        if (!g.bpc) continue;
        ccy += Py;
        sy = -1;
      } else if (ccy >= Py) {
        if (!g.bpc) continue;
        ccy -= Py;
        sy = 1;
      }
      const float y0 = g.lo[1] + (float)vy * g.cs[1];
      const float gy = fmaxf(0.f, fmaxf(y0 - p.y, p.y - (y0 + g.cs[1])));
      const float rho2 = gx * gx + gy * gy;
      if (rho2 >= r2) continue;
      const float hz = sqrtf(r2 - rho2);
      int zlo = (int)floorf((p.z - hz - g.lo[2]) * g.inv_cs[2]);
      int zhi = (int)floorf((p.z + hz - g.lo[2]) * g.inv_cs[2]);
      zlo = max(zlo, cz - a.K[2]);
      zhi = min(zhi, cz + a.K[2]);
      const int col = (ccx * Py + ccy) * Pz;
      if (!g.bpc) {
        zlo = max(zlo, 0);
        zhi = min(zhi, Pz - 1);
        if (zhi >= zlo) {
          const int k0 = cs_k[col + zlo], k1 = cs_k[col + zhi + 1];
          if (k1 > k0) f(k0, k1, sx, sy, 0);
        }
      } else {
        if (zlo < 0) {
          const int k0 = cs_k[col + zlo + Pz], k1 = cs_k[col + Pz];
          if (k1 > k0) f(k0, k1, sx, sy, -1);
        }
        const int a0 = max(zlo, 0), a1 = min(zhi, Pz - 1);
        if (a1 >= a0) {
          const int k0 = cs_k[col + a0], k1 = cs_k[col + a1 + 1];
          if (k1 > k0) f(k0, k1, sx, sy, 0);
        }
        if (zhi > Pz - 1) {
          const int k0 = cs_k[col], k1 = cs_k[col + zhi - Pz + 1];
          if (k1 > k0) f(k0, k1, sx, sy, 1);
        }
      }
    }
  }
}

// One wave per point-1 task, in two alternating phases.
//   A (lanes = third points): point 1's neighbour segments are scanned 64
//     third points at a time; each lane bins its 1-3 side (FP32 table, exact
//     FP64 when ambiguous, always exact for 3ani's cosine) and the third points
//     in range are appended to the wave's list in LDS (ballot compaction):
//     image-shifted scaled coordinates and weight, index, b31 | p << 8 |
//     image << 16.
//   B (lanes = pairs): when the list is full (and at the end), the wave walks
//     point 1's pairs in the b12 slice, 64 per pass, against every list entry
//     (an LDS broadcast): the 2-3 side per lane, then the deposit.
// So the 1-3 side costs one lane-evaluation per third point instead of one
// wave-wide evaluation per third point and pass over the pairs.  ANI = 0: V
// histogram [b12 - lo][b23][b31] in LDS.  ANI = 1: the ordered 5-D histogram
// in global memory.  PLAIN = 1: every side through the FP64 arithmetic (the
// plain FP64 path).
template <int ANI>
struct ListCap {
  static constexpr int value = ANI ? 256 : 64;  // third-point entries per wave
};

// MF = 1 (3iso, mixed precision): phase B forms the 2-3 sides on the matrix
// cores, 16 list entries x 16 pairs per v_mfma_f32_16x16x4f32 in point 1's
// frame: rows = list entries (A operand (x3, y3, z3, |x3|^2), all relative to
// point 1, written by phase A), columns = pairs (B operand (-2x2, -2y2, -2z2,
// 1), C = |x2|^2), so d23 = |x2|^2 + |x3|^2 - 2 x2.x3.  The FP32 error of the
// form is part of the third-point table's ambiguity margin (every coordinate
// is within reach of point 1).  A lane reads one list component, the rows'
// b31 and w3 four at a time, per 16 entries, instead of three broadcast reads
// per entry.  GLOB = 1 (3iso): the V rows go to global memory (one b12 row
// exceeds the LDS budget), a separate instantiation so that the LDS kernels
// carry no global-atomic code (its registers would spill them).
template <int ANI, int WEIGHTED, int PLAIN, int MF, int GLOB>
__global__ __launch_bounds__(ANI ? kAniTripBlock : kTripBlock) void k_triplets(
    const float4* __restrict__ xs, const float4* __restrict__ xo,
    const float4* __restrict__ ks_all, const float4* __restrict__ ko_all,
    const int* __restrict__ k_cell_start, const PairRec* __restrict__ pairs,
    const int* __restrict__ poff, int* __restrict__ counter, const uint8_t* __restrict__ table,
    const double* __restrict__ T, unsigned long long* __restrict__ hist,
    double* __restrict__ whist, unsigned long long* __restrict__ stats, TripArgs a) {
  constexpr int BLOCK = ANI ? kAniTripBlock : kTripBlock;
  constexpr int CAP = ListCap<ANI>::value;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint8_t tab[kMixedTableMax];
  __shared__ double sT[kMaxTripBins + 2];
  const int bins = a.bins;
  const int bins2 = bins * bins;
  const int nrows = (ANI || GLOB) ? 0 : (a.hi - a.lo) * bins2;
  // Each V row has 2^cs copies side by side; a lane deposits into copy
  // lane & (2^cs - 1), so lanes meeting on one row hit distinct words.
  const int ncopy = nrows << a.cs;
  uint32_t* hc = reinterpret_cast<uint32_t*>(smem);
  const size_t hc_bytes = ((size_t)ncopy * 4 + 15) & ~(size_t)15;
  double* hw = reinterpret_cast<double*>(smem + hc_bytes);
  const size_t hist_bytes = (hc_bytes + (WEIGHTED ? (size_t)ncopy * 8 : 0) + 15) & ~(size_t)15;
  const int wave = threadIdx.x / kWave;
  float4* Lq = reinterpret_cast<float4*>(smem + hist_bytes) + wave * CAP;
  int* Lk = reinterpret_cast<int*>(smem + hist_bytes + (size_t)(BLOCK / kWave) * CAP * 16) +
            wave * CAP;
  int* Lm = reinterpret_cast<int*>(smem + hist_bytes + (size_t)(BLOCK / kWave) * CAP * 20) +
            wave * CAP;
  // MF: third-point weights, beside the other list arrays.
  float* Lw = reinterpret_cast<float*>(smem + hist_bytes + (size_t)(BLOCK / kWave) * CAP * 24) +
              wave * CAP;
  if (!PLAIN)
    for (int t = threadIdx.x; t < a.nt; t += BLOCK) tab[t] = table[t];
  for (int t = threadIdx.x; t <= bins; t += BLOCK) sT[t] = T[t];
  if (!ANI) {
    for (int t = threadIdx.x; t < ncopy; t += BLOCK) {
      hc[t] = 0u;
      if (WEIGHTED) hw[t] = 0.0;
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & (kWave - 1);
  const unsigned ubins = (unsigned)bins;
  unsigned long long evals = 0, nre = 0;
  // 3iso deposit of V row vrow = (b12 * bins + b23) * bins + b31: into the
  // LDS slice [lo, hi) of b12 rows, or (glob) straight into global memory.
  const unsigned vrow0 = GLOB ? 0u : (unsigned)a.lo * (unsigned)bins2;
  auto deposit_v = [&](unsigned vrow, double w) {
    if constexpr (GLOB) {
      atomicAdd(&hist[vrow], 1ull);
      if (WEIGHTED) atomicAdd(&whist[vrow], w);
    } else {
      const unsigned slot = ((vrow - vrow0) << a.cs) | (threadIdx.x & ((1u << a.cs) - 1u));
      atomicAdd(&hc[slot], 1u);
      if (WEIGHTED) atomicAdd(&hw[slot], w);
    }
  };

  for (;;) {
    int task = 0;
    if (lane == 0) task = atomicAdd(counter, 1);
    const int i = __builtin_amdgcn_readfirstlane(task);
    if (i >= a.nx) break;
    const int pbeg = __builtin_amdgcn_readfirstlane(poff[i * bins + a.lo]);
    const int pend = __builtin_amdgcn_readfirstlane(poff[i * bins + a.hi]);
    if (pend <= pbeg) continue;
    const float4 is = xs[i];   // point 1, scaled
    const float4 io = xo[i];   // point 1, original
    const double wi = (double)io.w;
    int n = 0;  // list entries (uniform)

    // ---- phase B: the wave's pairs against the list
    auto run_list = [&]() {
      __builtin_amdgcn_wave_barrier();
      // Lane l of chunk c takes pair l * nch + c: pairs are sorted by b12, so
      // consecutive pairs share histogram rows; spreading them over the lanes
      // keeps a wave's LDS atomics off one address.
      const int npair = pend - pbeg;
      const int nch = (npair + kWave - 1) / kWave;
      for (int ch = 0; ch < nch; ++ch) {
        const int pi = lane * nch + ch;
        const int p = pbeg + pi;
        const bool active = pi < npair;
        PairRec r;
        r.xj = r.yj = r.zj = 1e18f;
        r.wj = 0.f;
        r.j = -1;
        r.meta = 0;
        if (active) r = pairs[p];
        const int b12 = (r.meta >> 8) & 0xfff;
        const int t12 = (r.meta >> 20) & 0xfff;
        const int im2 = r.meta & 31;
        const int f2 = ((im2 / 9 - 1) != 0 ? 1 : 0) | (((im2 / 3) % 3 - 1) != 0 ? 2 : 0) |
                       ((im2 % 3 - 1) != 0 ? 4 : 0);
        const double w12 = wi * (double)r.wj;
        const unsigned nact = (unsigned)__popcll(__ballot(active));
        evals += (unsigned long long)nact * (unsigned long long)n;
        const unsigned rbase = (unsigned)b12 * (unsigned)bins2;
        const int64_t abase = ((int64_t)b12 * bins * bins * bins) * bins + (int64_t)t12 * bins;
        float4 jo = make_float4(0.f, 0.f, 0.f, 0.f);
        bool jo_loaded = false;
        for (int e = 0; e < n; ++e) {
          const float4 kq = Lq[e];
          const int k = Lk[e];
          const int meta = Lm[e];
          const int b31 = meta & 0xff;
          const int pb = (meta >> 8) & 0xff;
          const int simg = meta >> 16;
          const float cx = kq.x - r.xj, cy = kq.y - r.yj, cz = kq.z - r.zj;
          const float d23 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx));
          unsigned c23 = PLAIN ? ubins + 1 : tab[__float2uint_rz(a.tf - d23)];
          const bool amb = active && c23 > ubins;
          if (__ballot(amb)) {
            if (!PLAIN) nre += __popcll(__ballot(amb));
            if (amb) {
              c23 = ubins;  // excluded unless the exact path finds a bin
              if (!(a.auto_ && k == r.j)) {
                if (!jo_loaded) {
                  jo = xo[r.j];
                  jo_loaded = true;
                }
                const float4 kco = ko_all[k];
                const int g3 = image_flags(io, kco, simg, false);
                if (g3 >= 0) {
                  const float a2[3] = {jo.x, jo.y, jo.z}, a3[3] = {kco.x, kco.y, kco.z};
                  double t23[3];
                  for (int q = 0; q < 3; ++q) {
                    const double d = __dsub_rn((double)a3[q], (double)a2[q]);
                    t23[q] = a.bpc ? proj_term(d, a.L, ((f2 >> q) & 1) != ((g3 >> q) & 1)) : d;
                  }
                  const double s23 = __dadd_rn(
                      __dadd_rn(__dmul_rn(t23[0], t23[0]), __dmul_rn(t23[1], t23[1])),
                      __dmul_rn(t23[2], t23[2]));
                  if (!(a.auto_ && !(s23 > 0.0))) {
                    const int e23 = class_lds(s23, sT, bins, a.R, 0);
                    if (e23 >= 0) c23 = (unsigned)e23;
                  }
                }
              }
            }
          }
          if (active && c23 < ubins) {
            if (!ANI) {
              deposit_v(rbase + c23 * ubins + (unsigned)b31,
                        WEIGHTED ? __dmul_rn(w12, (double)kq.w) : 0.0);
            } else {
              const int64_t idx = abase + ((int64_t)b31 * bins * bins * bins +
                                           (int64_t)c23 * bins * bins) + pb;
              atomicAdd(&hist[idx], 1ull);
              if (WEIGHTED) atomicAdd(&whist[idx], __dmul_rn(w12, (double)kq.w));
            }
          }
        }
      }
      n = 0;
      __builtin_amdgcn_wave_barrier();
    };

    // ---- phase B on the matrix cores (MF): 64 pairs per chunk as four
    // column blocks of 16, the list 16 entries per step, blocks two at a time.
    auto run_list_mf = [&]() {
      typedef float f4v __attribute__((ext_vector_type(4)));
      constexpr int HB = WEIGHTED ? 1 : 2;  // column blocks per pass (registers)
      __builtin_amdgcn_wave_barrier();
      const int grp = lane >> 4, col = lane & 15;
      const float sel_x = grp == 0 ? 1.0f : 0.0f, sel_y = grp == 1 ? 1.0f : 0.0f;
      const float sel_z = grp == 2 ? 1.0f : 0.0f, sel_w = grp == 3 ? 1.0f : 0.0f;
      const int npair = pend - pbeg;
      const int nch = (npair + kWave - 1) / kWave;
      const float* Lqf = reinterpret_cast<const float*>(Lq);
      for (int ch = 0; ch < nch; ++ch) {
        // Column 16b + col of the chunk is pair (16b + col) * nch + ch: pairs are
        // sorted by b12, so spreading them keeps the deposits of a wave apart.
        float bop[4], cin[4], wj[4];
        unsigned rb[4];
        unsigned npact = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int pi = (16 * b + col) * nch + ch;
          const bool act = pi < npair;
          PairRec r;
          r.xj = r.yj = r.zj = 0.f;
          r.wj = 0.f;
          r.meta = 0;
          if (act) r = pairs[pbeg + pi];
          const float x = r.xj - is.x, y = r.yj - is.y, z = r.zj - is.z;
          // B (2x, 2y, 2z, -1) of point 2 and C = tf - |x2|^2: the product is
          // tf - d23, the flipped table's index (an inactive column: hugely
          // negative, cell 0, out).
          bop[b] = __builtin_fmaf(sel_x, 2.0f * x, __builtin_fmaf(sel_y, 2.0f * y,
                                                                  __builtin_fmaf(sel_z, 2.0f * z, -sel_w)));
          cin[b] = act ? a.tf - __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x)) : -1e30f;
          rb[b] = (unsigned)((r.meta >> 8) & 0xfff) * (unsigned)bins2;
          wj[b] = r.wj;
          npact += (unsigned)__popcll(__ballot(act && grp == 0));
        }
        evals += (unsigned long long)npact * (unsigned long long)n;
        for (int jb = 0; jb < n; jb += 16) {
          // A operand: component grp of entry jb + col (|x3|^2 huge past n).
          const int e = jb + col;
          const float comp = Lqf[(size_t)min(e, CAP - 1) * 4 + grp];
          const float aop = e < n ? comp : sel_w * 1e30f;
          // Rows 4grp + v of this step: b31 and weight of entry jb + 4grp + v.
          const int4 mrow = *reinterpret_cast<const int4*>(Lm + jb + 4 * grp);
          const float4 wrow = *reinterpret_cast<const float4*>(Lw + jb + 4 * grp);
          const int mr[4] = {mrow.x, mrow.y, mrow.z, mrow.w};
          const float wr[4] = {wrow.x, wrow.y, wrow.z, wrow.w};
          unsigned ambm = 0;  // (b, v) -> bit 4b + v: ambiguous elements of this lane
#pragma unroll
          for (int h = 0; h < 4; h += HB) {
            f4v d[HB];
#pragma unroll
            for (int u = 0; u < HB; ++u) {
              const f4v cv = {cin[h + u], cin[h + u], cin[h + u], cin[h + u]};
              d[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, bop[h + u], cv, 0, 0, 0);
            }
            unsigned c23[4 * HB];
#pragma unroll
            for (int u = 0; u < HB; ++u)
#pragma unroll
              for (int v = 0; v < 4; ++v) c23[4 * u + v] = tab[__float2uint_rz(d[u][v])];
            unsigned mx = c23[0];
#pragma unroll
            for (int q = 1; q < 4 * HB; ++q) mx = max(mx, c23[q]);
            if (__builtin_expect(__ballot(mx > ubins) != 0, 0)) {
              // Ambiguous cells (the largest code) are set aside for the exact path.
#pragma unroll
              for (int q = 0; q < 4 * HB; ++q)
                if (c23[q] > ubins) {
                  ambm |= 1u << (4 * h + q);
                  c23[q] = ubins;
                }
            }
#pragma unroll
            for (int u = 0; u < HB; ++u) {
              const double w12 = WEIGHTED ? wi * (double)wj[h + u] : 0.0;  // exact
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const unsigned cc = c23[4 * u + v];
                if (cc < ubins) {
                  deposit_v(rb[h + u] + cc * ubins + (unsigned)(mr[v] & 0xff),
                            WEIGHTED ? __dmul_rn(w12, (double)wr[v]) : 0.0);
                }
              }
            }
          }
          // The exact FP64 2-3 side of the ambiguous elements, each lane taking
          // its lowest remaining one per round; everything is re-read by index.
          while (__ballot(ambm != 0)) {
            if (ambm) {
              const int q = __builtin_ctz(ambm);
              ambm &= ambm - 1;
              const int b = q >> 2, v = q & 3;
              ++nre;
              const PairRec r = pairs[pbeg + (16 * b + col) * nch + ch];
              const int ent = jb + 4 * grp + v;
              const int k = Lk[ent];
              const int meta = Lm[ent];
              if (!(a.auto_ && k == r.j)) {
                const int im2 = r.meta & 31;
                const int f2 = ((im2 / 9 - 1) != 0 ? 1 : 0) | (((im2 / 3) % 3 - 1) != 0 ? 2 : 0) |
                               ((im2 % 3 - 1) != 0 ? 4 : 0);
                const float4 jo = xo[r.j];
                const float4 kco = ko_all[k];
                const int g3 = image_flags(io, kco, meta >> 16, false);
                if (g3 >= 0) {
                  const float a2[3] = {jo.x, jo.y, jo.z}, a3[3] = {kco.x, kco.y, kco.z};
                  double t23[3];
                  for (int qq = 0; qq < 3; ++qq) {
                    const double dd = __dsub_rn((double)a3[qq], (double)a2[qq]);
                    t23[qq] = a.bpc ? proj_term(dd, a.L, ((f2 >> qq) & 1) != ((g3 >> qq) & 1)) : dd;
                  }
                  const double s23 = __dadd_rn(
                      __dadd_rn(__dmul_rn(t23[0], t23[0]), __dmul_rn(t23[1], t23[1])),
                      __dmul_rn(t23[2], t23[2]));
                  if (!(a.auto_ && !(s23 > 0.0))) {
                    const int e23 = class_lds(s23, sT, bins, a.R, 0);
                    if (e23 >= 0) {
                      deposit_v((unsigned)((r.meta >> 8) & 0xfff) * (unsigned)bins2 +
                                    (unsigned)e23 * ubins + (unsigned)(meta & 0xff),
                                WEIGHTED ? __dmul_rn(wi * (double)r.wj, (double)Lw[ent]) : 0.0);
                    }
                  }
                }
              }
            }
          }
        }
      }
      n = 0;
      __builtin_amdgcn_wave_barrier();
    };

    // ---- XXX 3iso on the matrix cores: point 1's third points (k > i with
    // b31 >= lo) are its own pairs of b12 >= lo (i < j, in range), already
    // listed in bin order with exact bins and images and in point 1's frame:
    // copied from the pair list instead of walking the neighbour cells again
    // in every row launch.
    if constexpr (MF && !ANI) {
      if (a.auto_) {
        const int qb = __builtin_amdgcn_readfirstlane(poff[i * bins + a.lo]);
        const int qe = __builtin_amdgcn_readfirstlane(poff[i * bins + bins]);
        for (int kb = qb; kb < qe; kb += kWave) {
          const int q = kb + lane;
          const bool v = q < qe;
          PairRec r;
          r.xj = r.yj = r.zj = 0.f;
          r.wj = 0.f;
          r.j = -1;
          r.meta = 0;
          if (v) r = pairs[q];
          const unsigned long long m = __ballot(v);
          const int cnt = __popcll(m);
          if (n + cnt > CAP) run_list_mf();
          const int pos = n + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
          if (v) {
            const float ax = r.xj - is.x, ay = r.yj - is.y, az = r.zj - is.z;
            Lq[pos] = make_float4(ax, ay, az, __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax)));
            Lw[pos] = r.wj;
            Lk[pos] = r.j;
            Lm[pos] = ((r.meta >> 8) & 0xfff) | ((r.meta & 31) << 16);
          }
          n += cnt;
        }
        if (n > 0) run_list_mf();
        continue;
      }
    }

    // ---- phase A: point 1's third points, one per lane
    for_each_segment(io, a, k_cell_start, [&](int k0, int k1, int sx, int sy, int sz) {
      const float sxL = (float)sx * a.Ls, syL = (float)sy * a.Ls, szL = (float)sz * a.Ls;
      const int simg = (sx + 1) * 9 + (sy + 1) * 3 + (sz + 1);
      for (int kb = k0; kb < k1; kb += kWave) {
        const int k = kb + lane;
        // XXX 3iso: only third points above point 1 (point 1 is then the
        // lowest of the three and each unordered triplet is met twice, not
        // three times; k_sym_cyc restores the six orders).  3ani and XXY
        // take every third point but point 1 itself.
        bool v = k < k1 && !(a.auto_ && (ANI ? k == i : k <= i));
        float4 kq = make_float4(1e18f, 1e18f, 1e18f, 0.f);
        if (v) kq = ks_all[k];
        const float kx = kq.x + sxL, ky = kq.y + syL, kz = kq.z + szL;
        const float ax = kx - is.x, ay = ky - is.y, az = kz - is.z;
        const float d31 = __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax));
        unsigned c31 = PLAIN ? ubins + 1 : tab[__float2uint_rz(a.tf - d31)];
        if (c31 == ubins) v = false;  // certainly out of range of point 1
        int b31 = (int)c31, pb = 0;
        if (v && (ANI || c31 > ubins)) {
          // exact 1-3 side (and the 3ani cosine bin)
          const float4 kco = ko_all[k];
          const int f3 = image_flags(io, kco, simg, false);
          if (f3 < 0) {
            v = false;
          } else {
            const float a1[3] = {io.x, io.y, io.z}, a3[3] = {kco.x, kco.y, kco.z};
            double t31[3];
            for (int q = 0; q < 3; ++q) {
              const double d = __dsub_rn((double)a3[q], (double)a1[q]);
              t31[q] = a.bpc ? proj_term(d, a.L, (f3 >> q) & 1) : d;
            }
            const double s31 = __dadd_rn(
                __dadd_rn(__dmul_rn(t31[0], t31[0]), __dmul_rn(t31[1], t31[1])),
                __dmul_rn(t31[2], t31[2]));
            if (a.auto_ && !(s31 > 0.0)) v = false;
            b31 = class_lds(s31, sT, bins, a.R, 0);
            if (b31 < 0) v = false;
            if (ANI && v) {
              const double cth = __dadd_rn(1.0, t31[2] / sqrt(s31));
              const double pp = __dmul_rn(cth, a.ds_th);
              pb = isnan(pp) ? 0 : (int)pp;
              if (pb > bins - 1) v = false;
            }
          }
        }
        // XXX 3iso: slices are one b12 row wide (lo == b12 of every pair of
        // the launch), so dropping third points with b31 < lo keeps exactly
        // the meetings with b12 <= b31.
        if (!ANI && a.auto_ && v && b31 < a.lo) v = false;
        const unsigned long long m = __ballot(v);
        const int cnt = __popcll(m);
        if (cnt == 0) continue;
        if (n + cnt > CAP) {
          if constexpr (MF) run_list_mf();
          else run_list();
        }
        const int pos = n + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (v) {
          if (MF) {
            Lq[pos] = make_float4(ax, ay, az, d31);
            Lw[pos] = kq.w;
          } else {
            Lq[pos] = make_float4(kx, ky, kz, kq.w);
          }
          Lk[pos] = k;
          Lm[pos] = b31 | (pb << 8) | (simg << 16);
        }
        n += cnt;
      }
    });
    if (n > 0) {
      if constexpr (MF) run_list_mf();
      else run_list();
    }
  }
  if (!ANI) {
    __syncthreads();
    const int base = a.lo * bins2;
    const int C = 1 << a.cs;
    for (int t = threadIdx.x; t < nrows; t += BLOCK) {
      unsigned long long cnt = 0;
      double w = 0.0;
      for (int q = 0; q < C; ++q) {
        cnt += hc[t * C + q];
        if (WEIGHTED) w += hw[t * C + q];
      }
      if (cnt) {
        atomicAdd(&hist[base + t], cnt);
        if (WEIGHTED) atomicAdd(&whist[base + t], w);
      }
    }
  }
  if (lane == 0) atomicAdd(&stats[0], evals);
  if (lane == 0 && nre) atomicAdd(&stats[1], nre);
}

// 3ani with the histogram privatised in LDS.  A slice is one b12 row and a t
// range [tlo, tlo + tn); a workgroup serves one slice (its LDS holds
// [b31][b23][t - tlo][p], u32 counts and FP64 weights) and shares the slice's
// point-1 tasks with the other workgroups of that slice; one launch runs all
// slices, workgroups given to slices in proportion to their pairs.  A wave
// owns point 1: its lanes take point 1's third points (the list q, 64 at a
// time), and point 1's pairs in the slice (contiguous in the list p, which is
// keyed by point 1, b12, t) stream through scalar loads, one per step, against
// all lanes.  The third-point list of XXX is the pair list itself (point 3 is
// any point in range of point 1); XXY's comes from k_emit_cross.
// Largest bin table of the LDS 3ani kernel: the rest of its table array holds
// the per-wave recheck queues.
constexpr int kAniTableMax = kMixedTableMax - 8 * kTripBlock;
static_assert(kAniTableMax % 16 == 0 && kAniTableMax > 0, "3ani table and queues");

struct AniSlice {
  int b12, tlo, tn;
  int qstride;  // key stride of the third-point list (bins^2 for XXX, 1 for XXY)
};

// Pairs per (b12, t) over all points 1 (the slices' work estimate).
__global__ void k_key_mass(const int* __restrict__ off, int64_t nkeys, int bins2,
                           unsigned long long* __restrict__ mass) {
  __shared__ unsigned long long m[1600];
  for (int t = threadIdx.x; t < bins2; t += blockDim.x) m[t] = 0ull;
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nkeys;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int c = off[k + 1] - off[k];
    if (c) atomicAdd(&m[k % bins2], (unsigned long long)c);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < bins2; t += blockDim.x)
    if (m[t]) atomicAdd(&mass[t], m[t]);
}

template <int WEIGHTED, int PLAIN>
__global__ __launch_bounds__(kTripBlock) void k_trip_ani_lds(
    const float4* __restrict__ xo, const float4* __restrict__ ko_all,
    const PairRec* __restrict__ P, const int* __restrict__ poff,
    const PairRec* __restrict__ Q, const int* __restrict__ qoff, int* __restrict__ counters,
    const uint8_t* __restrict__ table, const double* __restrict__ T,
    unsigned long long* __restrict__ hist, double* __restrict__ whist,
    unsigned long long* __restrict__ stats, TripArgs a, const AniSlice* __restrict__ slices,
    const int* __restrict__ wg_slice) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int si = wg_slice[blockIdx.x];
  const AniSlice sl = slices[si];
  int* const counter = counters + si;
  // The bin table (a.nt <= kAniTableMax bytes) and behind it the waves'
  // recheck queues (64 entries of 8 bytes each).
  __shared__ __align__(16) uint8_t tab[kMixedTableMax];
  __shared__ double sT[kMaxTripBins + 2];
  uint2* const rq = reinterpret_cast<uint2*>(tab + kAniTableMax);
  const int B = a.bins;
  const int nrows = B * B * sl.tn * B;  // [b31][b23][t - tlo][p]
  uint32_t* hc = reinterpret_cast<uint32_t*>(smem);
  double* hw = reinterpret_cast<double*>(smem + (((size_t)nrows * 4 + 15) & ~(size_t)15));
  if (!PLAIN)
    for (int t = threadIdx.x; t < a.nt; t += kTripBlock) tab[t] = table[t];
  for (int t = threadIdx.x; t <= B; t += kTripBlock) sT[t] = T[t];
  for (int t = threadIdx.x; t < nrows; t += kTripBlock) {
    hc[t] = 0u;
    if (WEIGHTED) hw[t] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned ubins = (unsigned)B;
  const int B2 = B * B;
  unsigned long long evals = 0, nre = 0;
  // The exact FP64 bin of the 2-3 side (point 2 of pair r, third point at
  // kco with image flags g3), or -1 out of range.
  auto exact23 = [&](const PairRec& r, const float4 jo, const float4 kco, int g3) -> int {
    const int im2 = r.meta & 31;
    const int f2 = ((im2 / 9 - 1) != 0 ? 1 : 0) | (((im2 / 3) % 3 - 1) != 0 ? 2 : 0) |
                   ((im2 % 3 - 1) != 0 ? 4 : 0);
    const float a2[3] = {jo.x, jo.y, jo.z}, a3[3] = {kco.x, kco.y, kco.z};
    double t23[3];
    for (int qd = 0; qd < 3; ++qd) {
      const double d = __dsub_rn((double)a3[qd], (double)a2[qd]);
      t23[qd] = a.bpc ? proj_term(d, a.L, ((f2 >> qd) & 1) != ((g3 >> qd) & 1)) : d;
    }
    const double s23 = __dadd_rn(__dadd_rn(__dmul_rn(t23[0], t23[0]), __dmul_rn(t23[1], t23[1])),
                                 __dmul_rn(t23[2], t23[2]));
    if (a.auto_ && !(s23 > 0.0)) return -1;
    return class_lds(s23, sT, B, a.R, 0);
  };
  // Mixed path: a wave's ambiguous (pair, third point) elements wait in its
  // LDS queue (pair index, third-point index) and are evaluated 64 at a time,
  // one per lane, instead of stalling the whole wave on each one's exact
  // FP64 side and its two coordinate loads.
  uint2* const wq = rq + (PLAIN ? 0 : (threadIdx.x & ~(kWave - 1)));
  int qn = 0;
  double wi = 0.0;
  auto drain = [&]() {
    __builtin_amdgcn_wave_barrier();
    if (lane < qn) {
      const uint2 e = wq[lane];
      const PairRec r = P[e.x];
      const PairRec kq = Q[e.y];
      if (!(a.auto_ && kq.j == r.j)) {
        const int img3 = kq.meta & 31;
        const int g3 = ((img3 / 9 - 1) != 0 ? 1 : 0) | (((img3 / 3) % 3 - 1) != 0 ? 2 : 0) |
                       ((img3 % 3 - 1) != 0 ? 4 : 0);
        const int e23 = exact23(r, xo[r.j], ko_all[kq.j], g3);
        if (e23 >= 0) {
          const int b31 = (kq.meta >> 8) & 0xfff, p31 = (kq.meta >> 20) & 0xfff;
          const int t12 = (r.meta >> 20) & 0xfff;
          const unsigned row =
              (((unsigned)b31 * ubins + (unsigned)e23) * (unsigned)sl.tn + (unsigned)(t12 - sl.tlo)) *
                  ubins + (unsigned)p31;
          atomicAdd(&hc[row], 1u);
          if (WEIGHTED) atomicAdd(&hw[row], __dmul_rn(wi * (double)r.wj, (double)kq.wj));
        }
      }
    }
    qn = 0;
    __builtin_amdgcn_wave_barrier();
  };
  for (;;) {
    int task = 0;
    if (lane == 0) task = atomicAdd(counter, 1);
    const int i = __builtin_amdgcn_readfirstlane(task);
    if (i >= a.nx) break;
    const int kb = i * B2 + sl.b12 * B + sl.tlo;
    const int pbeg = __builtin_amdgcn_readfirstlane(poff[kb]);
    const int pend = __builtin_amdgcn_readfirstlane(poff[kb + sl.tn]);
    if (pend <= pbeg) continue;
    const int qbeg = __builtin_amdgcn_readfirstlane(qoff[i * sl.qstride]);
    const int qend = __builtin_amdgcn_readfirstlane(qoff[(i + 1) * sl.qstride]);
    if (qend <= qbeg) continue;
    wi = (double)xo[i].w;
    // The list is cut into 8 regions of nch * 8 entries; chunk c gives the 8
    // lanes of group g the entries g * nch * 8 + c * 8 + [0, 8).  One load
    // then reads 8 runs of 8 consecutive records (a few cache lines, not 64),
    // while the list's sort by (b31, p) still spreads a wave's histogram rows
    // over 8 regions, keeping its LDS atomics off one address.
    const int nq = qend - qbeg;
    const int nch = (nq + kWave - 1) / kWave;
    auto qslot = [&](int ch) { return ((lane >> 3) * nch + ch) * 8 + (lane & 7); };
    auto load_q = [&](int ch) {
      PairRec kr;
      kr.xj = kr.yj = kr.zj = 1e18f;
      kr.wj = 0.f;
      kr.j = -1;
      kr.meta = 0;
      const int qi = qslot(ch);
      if (qi < nq) kr = Q[qbeg + qi];
      return kr;
    };
    // Point 1's pairs in the slice, up to 64 per pass: one vector load per
    // lane (the next pass's in flight while this one runs), each pair then
    // broadcast from its lane by readlane.  Scalar loads here would share the
    // LDS counter with the table reads, so every table wait would also wait
    // for them.
    auto load_p = [&](int pc) {
      PairRec r;
      r.xj = r.yj = r.zj = 0.f;
      r.wj = 0.f;
      r.j = -1;
      r.meta = 0;
      if (pc + lane < pend) r = P[pc + lane];
      return r;
    };
    PairRec pnext = load_p(pbeg);
    for (int pc = pbeg; pc < pend; pc += kWave) {
      const int np = min(kWave, pend - pc);
      const PairRec pmine = pnext;
      if (pc + kWave < pend) pnext = load_p(pc + kWave);
      auto rdl = [](float v, int u) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), u));
      };
      auto bcast = [&](int u) -> PairRec {
        PairRec r;
        r.xj = rdl(pmine.xj, u);
        r.yj = rdl(pmine.yj, u);
        r.zj = rdl(pmine.zj, u);
        r.wj = WEIGHTED ? rdl(pmine.wj, u) : 1.f;
        r.j = PLAIN ? __builtin_amdgcn_readlane(pmine.j, u) : 0;
        r.meta = __builtin_amdgcn_readlane(pmine.meta, u);
        return r;
      };
      PairRec kr_next = load_q(0);
      for (int ch = 0; ch < nch; ++ch) {
        const PairRec kr = kr_next;
        if (ch + 1 < nch) kr_next = load_q(ch + 1);  // in flight while this chunk runs
        const bool active = qslot(ch) < nq;
        const int k = kr.j;
        const int img3 = kr.meta & 31;
        const int b31 = (kr.meta >> 8) & 0xfff;
        const int p31 = (kr.meta >> 20) & 0xfff;
        const int g3 = ((img3 / 9 - 1) != 0 ? 1 : 0) | (((img3 / 3) % 3 - 1) != 0 ? 2 : 0) |
                       ((img3 % 3 - 1) != 0 ? 4 : 0);
        const double wk = (double)kr.wj;
        const unsigned nact = (unsigned)__popcll(__ballot(active));
        float4 kco = make_float4(0.f, 0.f, 0.f, 0.f);
        bool kco_loaded = false;
        const unsigned rb = ((unsigned)b31 * ubins) * (unsigned)sl.tn * ubins + (unsigned)p31;
        // The 2-3 side of pair r against every lane's third point: the table
        // class; an ambiguous cell is queued (mixed path) or evaluated exactly
        // at once (plain path, where every cell is ambiguous).
        auto classify = [&](const PairRec& r, int pidx, unsigned c23) -> unsigned {
          const bool amb = active && c23 > ubins;
          if (!PLAIN) {
            const unsigned long long m = __ballot(amb);
            if (m) {
              const int n = __popcll(m);
              nre += n;
              if (qn + n > kWave) drain();
              if (amb) {
                const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi(
                                         (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                wq[pos] = make_uint2((unsigned)pidx, (unsigned)(qbeg + qslot(ch)));
                c23 = ubins;
              }
              qn += n;
            }
          } else if (amb) {
            c23 = ubins;
            if (!(a.auto_ && k == r.j)) {
              if (!kco_loaded) {
                kco = ko_all[k];
                kco_loaded = true;
              }
              const int e23 = exact23(r, xo[r.j], kco, g3);
              if (e23 >= 0) c23 = (unsigned)e23;
            }
          }
          return c23;
        };
        auto lookup = [&](const PairRec& r) -> unsigned {
          const float cx = kr.xj - r.xj, cy = kr.yj - r.yj, cz = kr.zj - r.zj;
          const float d23 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx));
          return PLAIN ? ubins + 1 : tab[__float2uint_rz(a.tf - d23)];
        };
        auto deposit = [&](const PairRec& r, unsigned c23) {
          if (active && c23 < ubins) {
            const int t12 = (r.meta >> 20) & 0xfff;
            const unsigned row = rb + (c23 * (unsigned)sl.tn + (unsigned)(t12 - sl.tlo)) * ubins;
            atomicAdd(&hc[row], 1u);
            if (WEIGHTED) atomicAdd(&hw[row], __dmul_rn(wi * (double)r.wj, wk));
          }
        };
        int u = 0;
        for (; u + 4 <= np; u += 4) {
          const PairRec r0 = bcast(u), r1 = bcast(u + 1), r2 = bcast(u + 2), r3 = bcast(u + 3);
          unsigned c0 = lookup(r0), c1 = lookup(r1), c2 = lookup(r2), c3 = lookup(r3);
          evals += 4ull * nact;
          c0 = classify(r0, pc + u, c0);
          c1 = classify(r1, pc + u + 1, c1);
          c2 = classify(r2, pc + u + 2, c2);
          c3 = classify(r3, pc + u + 3, c3);
          deposit(r0, c0);
          deposit(r1, c1);
          deposit(r2, c2);
          deposit(r3, c3);
        }
        for (; u < np; ++u) {
          const PairRec r = bcast(u);
          evals += nact;
          deposit(r, classify(r, pc + u, lookup(r)));
        }
      }
    }
    if (!PLAIN && qn) drain();
  }
  __syncthreads();
  // LDS row (b31, b23, dt, p) -> global [b12][b31][b23][t][p]
  for (int t = threadIdx.x; t < nrows; t += kTripBlock) {
    if (!hc[t]) continue;
    const int p = t % B;
    const int dt = (t / B) % sl.tn;
    const int b23 = (t / (B * sl.tn)) % B;
    const int b31 = t / (B * sl.tn * B);
    const int64_t idx =
        ((((int64_t)sl.b12 * B + b31) * B + b23) * B + (sl.tlo + dt)) * B + p;
    atomicAdd(&hist[idx], (unsigned long long)hc[t]);
    if (WEIGHTED) atomicAdd(&whist[idx], hw[t]);
  }
  if (lane == 0) atomicAdd(&stats[0], evals);
  if (lane == 0 && nre) atomicAdd(&stats[1], nre);
}

// Reference-defect triplet weights (pcf_set_ref_bugs): the reference's triplet
// kernels set v = w1 * w2 (an FP32 product) for each (point 1, point 2) and,
// walking node 3's points in order, multiply v by w3 before each deposit and
// deposit v (cuda/kernels/3iso_k.cuh:61,79,168, 3isoBPC_k.cuh:60,88,179,
// 3ani_k.cuh:94,197, 3aniBPC_k.cuh:98), so the weight of a triplet is the
// product over the third points deposited before it in its node (and
// projection) -- a chain that restarts at each node.  Counts are unchanged.
// A wave owns point 1 and walks the reference nodes within reach of it (the
// third-point catalog in the reference's node order, refbugs.hip) in
// lockstep; its lanes are point 1's pairs, each with its own chain, in FP64
// exact arithmetic (the plain path's rules).  3iso pairs are unordered:
// (j, i) deposits the same values into the (b23, b31)-transposed bins (the
// chains of the two orders hold the same third points: relabelling 1 and 2
// maps projection (f2, f3) to (f2, f2 ^ f3)), so k_sym_bc completes the
// histogram; 3ani's list holds both orders.
struct RefArgs {
  int nx;          // points 1
  int kstride;     // pair-list keys per point 1
  int bins, auto_, bpc, P;
  float size_node, reach;
  double R, L, ds_th;
};

template <int ANI>
__global__ __launch_bounds__(256) void k_trip_refbugs(
    const float4* __restrict__ xo, const float4* __restrict__ kref, const int* __restrict__ ns,
    const PairRec* __restrict__ pairs, const int* __restrict__ poff, int* __restrict__ counter,
    const double* __restrict__ T, unsigned long long* __restrict__ hist,
    double* __restrict__ whist, unsigned long long* __restrict__ stats, RefArgs a) {
  __shared__ double sT[kMaxTripBins + 2];
  for (int t = threadIdx.x; t <= a.bins; t += blockDim.x) sT[t] = T[t];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int B = a.bins, P = a.P;
  const float sn = a.size_node;
  unsigned long long evals = 0;
  auto node_axis = [&](float v) {
    const int c = (int)(v / sn);
    return c < 0 ? 0 : (c >= P ? P - 1 : c);
  };
  for (;;) {
    int task = 0;
    if (lane == 0) task = atomicAdd(counter, 1);
    const int i = __builtin_amdgcn_readfirstlane(task);
    if (i >= a.nx) break;
    const int pb = __builtin_amdgcn_readfirstlane(poff[(int64_t)i * a.kstride]);
    const int pe = __builtin_amdgcn_readfirstlane(poff[(int64_t)(i + 1) * a.kstride]);
    if (pe <= pb) continue;
    const float4 p1 = xo[i];
    const float a1[3] = {p1.x, p1.y, p1.z};
    const int cx = node_axis(p1.x), cy = node_axis(p1.y), cz = node_axis(p1.z);
    int K = (int)ceilf(a.reach / sn) + 1;
    if (a.bpc) K = min(K, P);
    const float r2 = a.reach * a.reach;
    for (int pc = pb; pc < pe; pc += kWave) {
      const bool act = pc + lane < pe;
      PairRec r;
      r.j = i;
      r.meta = 13;  // image code of (0, 0, 0)
      r.wj = 0.f;
      if (act) r = pairs[pc + lane];
      const float4 p2 = xo[r.j];
      const int f2 = act ? image_flags(p1, p2, r.meta & 31, false) : 0;
      const int b12 = (r.meta >> 8) & 0xfff, t12 = (r.meta >> 20) & 0xfff;
      const float w12 = __fmul_rn(p1.w, p2.w);  // FP32, as the reference forms v
      const float a2[3] = {p2.x, p2.y, p2.z};
      const unsigned long long nact = (unsigned long long)__popcll(__ballot(act));
      for (int dx = -K; dx <= K; ++dx) {
        const int vx = cx + dx;
        int ccx = vx, sx = 0;
        if (vx < 0 || vx >= P) {
          if (!a.bpc) continue;
          sx = vx < 0 ? -1 : 1;
          ccx = vx - sx * P;
        }
        const float x0 = (float)vx * sn;
        const float gx = fmaxf(0.f, fmaxf(x0 - p1.x, p1.x - (x0 + sn))) * 0.99999f;
        for (int dy = -K; dy <= K; ++dy) {
          const int vy = cy + dy;
          int ccy = vy, sy = 0;
          if (vy < 0 || vy >= P) {
            if (!a.bpc) continue;
            sy = vy < 0 ? -1 : 1;
            ccy = vy - sy * P;
          }
          const float y0 = (float)vy * sn;
          const float gy = fmaxf(0.f, fmaxf(y0 - p1.y, p1.y - (y0 + sn))) * 0.99999f;
          const float lat = gx * gx + gy * gy;
          if (lat >= r2) continue;
          const float hz = sqrtf(r2 - lat) * 1.00001f;
          const int zlo = max(cz - K, (int)floorf((p1.z - hz) / sn) - 1);
          const int zhi = min(cz + K, (int)floorf((p1.z + hz) / sn) + 1);
          for (int vz = zlo; vz <= zhi; ++vz) {
            int ccz = vz, sz = 0;
            if (vz < 0 || vz >= P) {
              if (!a.bpc) continue;
              sz = vz < 0 ? -1 : 1;
              ccz = vz - sz * P;
            }
            const int node = (ccx * P + ccy) * P + ccz;
            const int q0 = ns[node], q1 = ns[node + 1];
            if (q1 <= q0) continue;
            const int img3 = (sx + 1) * 9 + (sy + 1) * 3 + (sz + 1);
            double v = (double)w12;  // the chain restarts at each node (and projection)
            for (int q = q0; q < q1; ++q) {
              const float4 p3 = kref[q];
              const int f3 = image_flags(p1, p3, img3, false);
              if (f3 < 0) continue;
              const float a3[3] = {p3.x, p3.y, p3.z};
              double t31[3], t23[3];
              for (int ax = 0; ax < 3; ++ax) {
                const double d = __dsub_rn((double)a3[ax], (double)a1[ax]);
                t31[ax] = a.bpc ? proj_term(d, a.L, (f3 >> ax) & 1) : d;
              }
              const double s31 = __dadd_rn(
                  __dadd_rn(__dmul_rn(t31[0], t31[0]), __dmul_rn(t31[1], t31[1])),
                  __dmul_rn(t31[2], t31[2]));
              if (a.auto_ && !(s31 > 0.0)) continue;
              const int b31 = class_lds(s31, sT, B, a.R, 0);
              if (b31 < 0) continue;
              int pbin = 0;
              if (ANI) {
                const double pp = __dmul_rn(__dadd_rn(1.0, t31[2] / sqrt(s31)), a.ds_th);
                pbin = isnan(pp) ? 0 : (int)pp;
                if (pbin > B - 1) continue;
              }
              evals += nact;
              for (int ax = 0; ax < 3; ++ax) {
                const double d = __dsub_rn((double)a3[ax], (double)a2[ax]);
                t23[ax] = a.bpc ? proj_term(d, a.L, ((f2 >> ax) & 1) != ((f3 >> ax) & 1)) : d;
              }
              const double s23 = __dadd_rn(
                  __dadd_rn(__dmul_rn(t23[0], t23[0]), __dmul_rn(t23[1], t23[1])),
                  __dmul_rn(t23[2], t23[2]));
              int b23 = -1;
              if (act && !(a.auto_ && !(s23 > 0.0))) b23 = class_lds(s23, sT, B, a.R, 0);
              if (b23 >= 0) {
                v = __dmul_rn(v, (double)p3.w);
                const int64_t idx =
                    ANI ? ((((int64_t)b12 * B + b31) * B + b23) * B + t12) * B + pbin
                        : ((int64_t)b12 * B + b23) * B + b31;
                atomicAdd(&hist[idx], 1ull);
                atomicAdd(&whist[idx], v);
              }
            }
          }
        }
      }
    }
  }
  if (lane == 0) atomicAdd(&stats[0], evals);
}

// XXX 3iso: point 1 is the lowest index of its triplet (pairs i < j, third
// points k > i).  An unordered triplet {p < q < r} with sides A = b(pq),
// B = b(qr), C = b(rp) is met as (A, B, C) (pair (p, q), third r) and as
// (C, B, A) (pair (p, r), third q), and the kernel deposits the one(s) whose
// first index is at most its last (b12 <= b31: one-row b12 slices and third
// points with b31 >= b12): once when A != C, twice (the same cell) when A == C.  So V2[x][y][z] = V[x][y][z] + V[z][y][x] (x != z),
// V[x][y][x] (x == z) holds both meetings, and the six ordered triplets are
// V2's cells and their cyclic rotations:
// H[a][b][c] = V2[a][b][c] + V2[c][a][b] + V2[b][c][a].
__global__ void k_sym_cyc(const unsigned long long* __restrict__ V, const double* __restrict__ W,
                          int bins, unsigned long long* __restrict__ H, double* __restrict__ HW) {
  const int64_t n = (int64_t)bins * bins * bins;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int a = (int)(t / (bins * bins)), b = (int)((t / bins) % bins), c = (int)(t % bins);
  unsigned long long h = 0;
  double hw = 0.0;
  const int cyc[3][3] = {{a, b, c}, {c, a, b}, {b, c, a}};
  for (int r = 0; r < 3; ++r) {
    const int x = cyc[r][0], y = cyc[r][1], z = cyc[r][2];
    const int64_t u = ((int64_t)x * bins + y) * bins + z;
    h += V[u];
    hw += W[u];
    if (x != z) {
      const int64_t v = ((int64_t)z * bins + y) * bins + x;
      h += V[v];
      hw += W[v];
    }
  }
  H[t] = h;
  HW[t] = hw;
}

// XXY 3iso: H[a][b][c] = V[a][b][c] + V[a][c][b]
__global__ void k_sym_bc(const unsigned long long* __restrict__ V, const double* __restrict__ W,
                         int bins, unsigned long long* __restrict__ H, double* __restrict__ HW) {
  const int64_t n = (int64_t)bins * bins * bins;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int a = (int)(t / (bins * bins)), b = (int)((t / bins) % bins), c = (int)(t % bins);
  const int64_t u = ((int64_t)a * bins + c) * bins + b;
  H[t] = V[t] + V[u];
  HW[t] = W[t] + W[u];
}

__global__ void k_fill_int(int* p, int64_t n, int v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}


}  // namespace

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
void make_rules(int type, int bpc, int auto_, int bins, float dmax, float box, BinRules* r);

int run_triplets(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights) {
  const bool auto_ = job->hist == PCF_XXX;
  const bool ani = job->type == PCF_3ANI;
  Catalog* X = get_cat(c, job->cat_x);
  Catalog* K = auto_ ? X : get_cat(c, job->cat_y);
  if (!X || !K) return fail(c, PCF_ERR_ARG, "unknown catalog id");
  const int bins = job->bins;
  if (bins <= 0 || bins > kMaxTripBins)
    return fail(c, PCF_ERR_ARG, "bins out of range");
  if (!(job->dmax > 0) || !std::isfinite(job->dmax)) return fail(c, PCF_ERR_ARG, "dmax must be > 0");
  const int bpc = job->bpc ? 1 : 0;
  const float L = job->box;
  if (bpc) {
    if (!(L > 0)) return fail(c, PCF_ERR_ARG, "box side required for periodic boundaries");
    // 2*dmax >= L is handled (a point may count under its regular and its
    // projected image on an axis); dmax >= L would reach far images, which no
    // reference term is.
    if (!(job->dmax < L))
      return fail(c, PCF_ERR_UNSUPPORTED, "periodic boundaries need dmax < box side");
    for (Catalog* cc : {X, K})
      for (int a = 0; a < 3; ++a)
        if (cc->n && (cc->lo[a] < 0.f || !(cc->hi[a] < L)))
          return fail(c, PCF_ERR_RANGE, "coordinates outside [0, box) with periodic boundaries");
  }
  const int64_t b64 = bins;
  const int64_t nflat = ani ? b64 * b64 * b64 * b64 * b64 : b64 * b64 * b64;
  const bool weighted = !(X->all_w1 && K->all_w1);
  // With 2*dmax >= L the reference's 2-3 side, |x3 - x2| - L*|f2 - f3| per
  // axis, is no longer the separation of the two images in point 1's frame
  // (both points projected toward opposite walls, say), which the FP32 bulk
  // forms; such jobs take the plain FP64 kernels, which evaluate that rule.
  // pcf_set_ref_bugs with weights: the reference's cumulative triplet weight
  // (k_trip_refbugs, FP64 exact rules, reference node order of K).
  const bool refchain = ref_bugs(c) && weighted;
  if (refchain && !(L > 0))
    return fail(c, PCF_ERR_ARG,
                "reference triplet weights need the box side (the node grid is box / partitions)");
  const bool plain = job->precision == PCF_PREC_FP64 || (bpc && !(2.0 * job->dmax < L)) ||
                     refchain;

  GridGeom g;
  int Kc[3];
  float reach;
  int m = 0;
  if (choose_geometry(X, auto_ ? nullptr : K, job->dmax, bpc, L, job->cells_per_rmax, &g, Kc,
                      &reach, &m) != 0)
    return fail(c, PCF_ERR_UNSUPPORTED, "geometry");
  int rc = PCF_OK;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  GridPins pins(c);
  DevGrid* gx = pins.add(get_grid(c, X, g, job->rebuild_grid != 0, &rc));
  if (!gx) return rc;
  DevGrid* gk = auto_ ? gx : pins.add(get_grid(c, K, g, job->rebuild_grid != 0, &rc));
  if (!gk) return rc;

  // Exact rules: floored ds (3iso_k.cuh:44, 3isoBPC_k.cuh:218, 3ani_k.cuh:45),
  // strict range test.  Pair table: 0 < d12.  Third-point table: 0 < d23, d31
  // for XXX only.
  static thread_local BinRules pr, kr;
  make_rules(PCF_2ISO, 0, 1, bins, job->dmax, L, &pr);
  make_rules(PCF_2ISO, 0, auto_ ? 1 : 0, bins, job->dmax, L, &kr);
  pr.L = kr.L = (double)L;
  const double A = bpc ? 2.0 * ((double)std::max(X->maxabs, K->maxabs) + (double)L) : 0.0;
  static thread_local std::vector<uint8_t> ptab, ktab;
  float inv_h = 0.f, inv_h2 = 0.f;
  int nt = 0, nt2 = 0;
  // 3iso mixed: the 2-3 sides come from the matrix cores in point 1's frame,
  // where every coordinate is within reach of point 1; both tables carry that
  // form's error margin (they must share one scale).
  const bool gram = !ani && !plain;
  const double rho = gram ? (double)reach * 1.01 : -1.0;
  // 3ani's LDS kernel keeps its recheck queues behind a smaller table.
  const int tab_max = ani ? kAniTableMax : kMixedTableMax;
  // Both tables are flipped (cell cvt_rz(tf - d2) holds d2: beyond the top
  // reads cell 0, out, with no clamp) and share one scale and top.
  float tf = 0.f, tf2 = 0.f;
  if (build_table(pr, A, tab_max, &ptab, &inv_h, &nt, rho, true, &tf) != 0 ||
      build_table(kr, A, tab_max, &ktab, &inv_h2, &nt2, rho, true, &tf2) != 0 || tf != tf2 ||
      inv_h != inv_h2 ||
      nt != nt2 || nt > tab_max)
    return fail(c, PCF_ERR_UNSUPPORTED, "bin table");
  const float sigma = std::sqrt(inv_h);
  HIPCHK(c, c->table_buf.ensure(2 * (size_t)nt + sizeof(double) * (bins + 2)));
  uint8_t* d_ptab = c->table_buf.as<uint8_t>();
  uint8_t* d_ktab = d_ptab + nt;
  double* d_T = reinterpret_cast<double*>(c->table_buf.as<char>() +
                                          ((2 * (size_t)nt + 15) & ~(size_t)15));
  HIPCHK(c, c->table_buf.ensure(((2 * (size_t)nt + 15) & ~(size_t)15) + sizeof(double) * (bins + 2)));
  d_ptab = c->table_buf.as<uint8_t>();
  d_ktab = d_ptab + nt;
  d_T = reinterpret_cast<double*>(c->table_buf.as<char>() + ((2 * (size_t)nt + 15) & ~(size_t)15));
  HIPCHK(c, hipMemcpyAsync(d_ptab, ptab.data(), nt, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_ktab, ktab.data(), nt, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_T, pr.T, sizeof(double) * (bins + 1), hipMemcpyHostToDevice,
                           c->stream));

  // Scaled copies.
  HIPCHK(c, c->scaled_x.ensure(sizeof(float4) * (size_t)std::max<int64_t>(gx->n, 1)));
  HIPCHK(c, launch_scale_points(gx->pts, gx->n, sigma, c->scaled_x.as<float4>(), c->stream));
  const float4* xs = c->scaled_x.as<float4>();
  const float4* ksc = xs;
  if (!auto_) {
    HIPCHK(c, c->scaled_y.ensure(sizeof(float4) * (size_t)std::max<int64_t>(gk->n, 1)));
    HIPCHK(c, launch_scale_points(gk->pts, gk->n, sigma, c->scaled_y.as<float4>(), c->stream));
    ksc = c->scaled_y.as<float4>();
  }

  // 1. Pair list over X (unordered pairs i < j), counting-sorted by
  //    key = point1 * bins + b12.
  WorkListArgs wa;
  wa.x_cell_start = gx->cell_start;
  wa.y_cell_start = gx->cell_start;
  wa.g = g;
  for (int a = 0; a < 3; ++a) wa.K[a] = Kc[a];
  wa.reach = reach;
  wa.auto_ = 1;
  wa.cylinder = 0;
  wa.map = nullptr;
  wa.y_fine_start = nullptr;  // windows from whole z cells
  wa.x_pts = nullptr;
  wa.zq = 1;
  const int64_t nx = gx->n;
  // 3ani goes through the LDS-privatised kernel when one t value of a b12 row
  // ([b31][b23][p]) fits the LDS budget and the (point 1, b12, t) keys fit an
  // int; its pair list is then keyed by t as well.
  const size_t ani_budget = 160 * 1024 - kMixedTableMax - 8 * (kMaxTripBins + 2) - 1024;
  const size_t ani_per_t = (size_t)bins * bins * bins * (4 + (weighted ? 8 : 0));
  const bool ani_lds = ani && ani_per_t <= ani_budget && nx * b64 * b64 + 1 <= (int64_t)INT32_MAX;
  const int64_t nkeys = ani_lds ? nx * b64 * b64 : nx * bins;
  if (nkeys + 1 > (int64_t)INT32_MAX) return fail(c, PCF_ERR_UNSUPPORTED, "too many pair keys");
  const size_t kbytes = (sizeof(int) * (size_t)(3 * (nkeys + 1) + 8) + 15) & ~(size_t)15;
  EmitArgs ea;
  ea.nitems = 0;
  ea.nt = nt;
  ea.tf = tf;
  ea.Ls = L * sigma;
  ea.bins = bins;
  ea.bpc = bpc;
  ea.ani = ani ? 1 : 0;
  ea.R = pr.R;
  ea.L = (double)L;
  float ds_th_f = (float)bins / 2;  // 3ani_k.cuh:42
  ea.ds_th = (double)ds_th_f;
  ea.key_t = ani_lds ? 1 : 0;
  const int emit_blocks = std::max(1, c->ncu * 8);
  // The list depends on X's grid (its point order), the bins, range, box and
  // the scale and table of the FP32 classification; a job whose inputs match
  // the last list's (DDR after DDD, RRR after DRR) keeps it.
  TripPairKey pk;
  pk.grid_build = gx->build_id;
  pk.bins = bins;
  pk.bpc = bpc;
  pk.ani = ani ? 1 : 0;
  pk.key_t = ea.key_t;
  pk.nt = nt;
  for (int q = 0; q < 3; ++q) pk.K[q] = Kc[q];
  pk.dmax = job->dmax;
  pk.L = L;
  pk.sigma = sigma;
  pk.reach = reach;
  pk.R = pr.R;
  TripPairCache& pc = c->trip_cache;
  int npairs = 0;
  if (pc.valid && pc.key.same(pk) && pc.ptab == ptab) {
    npairs = pc.npairs;
  } else {
    pc.valid = false;
    int npitems = 0;
    HIPCHK(c, build_pair_worklist(wa, c->wl_counts, c->wl_items, &npitems, c->stream));
    HIPCHK(c, c->trip_idx.ensure(kbytes + sizeof(long long) * (size_t)scan_tiles(nkeys)));
    int* d_cnt = c->trip_idx.as<int>();
    int* d_off0 = d_cnt + (nkeys + 1);
    int* d_cur = d_off0 + (nkeys + 1);
    long long* d_tsum = reinterpret_cast<long long*>(c->trip_idx.as<char>() + kbytes);
    HIPCHK(c, launch_zero(d_cnt, sizeof(int) * (size_t)(nkeys + 1), c->stream));
    HIPCHK(c, launch_zero(d_cur, sizeof(int) * (size_t)(nkeys + 1), c->stream));
    ea.nitems = npitems;
    if (npitems > 0) {
      hipLaunchKernelGGL(k_emit_pairs<0>, dim3(emit_blocks), dim3(kEmitBlock), 0, c->stream, xs,
                         gx->pts, c->wl_items.as<PairItem>(), d_ptab, d_T, d_cnt, d_off0, d_cur,
                         (PairRec*)nullptr, ea);
      HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, launch_scan(d_cnt, nkeys, d_off0, d_tsum, c->stream));
    HIPCHK(c, hipMemcpyAsync(&npairs, d_off0 + nkeys, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (npairs < 0) return fail(c, PCF_ERR_UNSUPPORTED, "pair list exceeds 2^31 records");
    HIPCHK(c, c->trip_pairs.ensure(sizeof(PairRec) * (size_t)std::max(npairs, 1)));
    if (npitems > 0 && npairs > 0) {
      hipLaunchKernelGGL(k_emit_pairs<1>, dim3(emit_blocks), dim3(kEmitBlock), 0, c->stream, xs,
                         gx->pts, c->wl_items.as<PairItem>(), d_ptab, d_T, d_cnt, d_off0, d_cur,
                         c->trip_pairs.as<PairRec>(), ea);
      HIPCHK(c, hipGetLastError());
    }
    pc.key = pk;
    pc.ptab = ptab;
    pc.npairs = npairs;
    pc.valid = true;
  }
  int* const d_off = c->trip_idx.as<int>() + (nkeys + 1);
  PairRec* const d_pairs = c->trip_pairs.as<PairRec>();

  // 2. Histograms.
  const size_t hbytes = sizeof(unsigned long long) * (size_t)nflat;
  const size_t wbytes = sizeof(double) * (size_t)nflat;
  HIPCHK(c, c->trip_hist.ensure(2 * (hbytes + wbytes) + 64));
  unsigned long long* d_V = c->trip_hist.as<unsigned long long>();
  double* d_VW = reinterpret_cast<double*>(c->trip_hist.as<char>() + hbytes);
  unsigned long long* d_H = reinterpret_cast<unsigned long long*>(c->trip_hist.as<char>() + hbytes + wbytes);
  double* d_HW = reinterpret_cast<double*>(c->trip_hist.as<char>() + 2 * hbytes + wbytes);
  unsigned long long* d_stats =
      reinterpret_cast<unsigned long long*>(c->trip_hist.as<char>() + 2 * (hbytes + wbytes));
  int* d_counter = reinterpret_cast<int*>(d_stats + 4);
  HIPCHK(c, launch_zero(d_V, hbytes + wbytes, c->stream));
  HIPCHK(c, launch_zero(d_stats, 64, c->stream));

  TripArgs t;
  t.nx = (int)nx;
  t.nt = nt;
  t.tf = tf;
  t.Ls = L * sigma;
  t.bins = bins;
  t.auto_ = auto_ ? 1 : 0;
  t.bpc = bpc;
  t.R = kr.R;
  t.L = (double)L;
  t.ds_th = (double)ds_th_f;
  t.reach = reach;
  t.glob = 0;
  t.cs = 0;
  for (int q = 0; q < 3; ++q) t.K[q] = Kc[q];
  t.g = g;
  // The reference's node order of the third-point catalog (refchain only).
  ScopedDevBuf ref_order, ref_start, ref_pts;
  if (refchain) {
    int rr = build_ref_order(c, K, L, job->partitions, &ref_order, &ref_start, &ref_pts);
    if (rr != PCF_OK) return rr;
  }
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  if (npairs > 0 && refchain) {
    RefArgs ra;
    ra.nx = (int)nx;
    ra.kstride = ani_lds ? bins * bins : bins;
    ra.bins = bins;
    ra.auto_ = auto_ ? 1 : 0;
    ra.bpc = bpc;
    ra.P = ref_partitions(job->partitions);
    ra.size_node = L / (float)ra.P;
    ra.reach = reach;
    ra.R = kr.R;
    ra.L = (double)L;
    ra.ds_th = (double)ds_th_f;
    HIPCHK(c, launch_zero(d_counter, sizeof(int), c->stream));
    if (ani)
      hipLaunchKernelGGL(k_trip_refbugs<1>, dim3(c->ncu * 8), dim3(256), 0, c->stream, gx->pts,
                         ref_pts.as<float4>(), ref_start.as<int>(), d_pairs, d_off, d_counter,
                         d_T, d_V, d_VW, d_stats, ra);
    else
      hipLaunchKernelGGL(k_trip_refbugs<0>, dim3(c->ncu * 8), dim3(256), 0, c->stream, gx->pts,
                         ref_pts.as<float4>(), ref_start.as<int>(), d_pairs, d_off, d_counter,
                         d_T, d_V, d_VW, d_stats, ra);
    HIPCHK(c, hipGetLastError());
  } else if (npairs > 0) {
    if (!ani) {
      // LDS budget of one 1024-thread workgroup per CU: the slice of b12 rows
      // whose counts (and FP64 weights) fit beside the table.
      const size_t lists = (size_t)(kTripBlock / kWave) * ListCap<0>::value * (gram ? 28 : 24);
      const size_t budget = 160 * 1024 - kMixedTableMax - 8 * (kMaxTripBins + 2) - 1024 - lists;
      const size_t per_b12 = (size_t)bins * bins * (4 + (weighted ? 8 : 0));
      // A b12 row larger than the LDS budget (about 90 bins weighted, 160
      // unweighted) deposits straight into the global V histogram instead.
      const bool glob = per_b12 + 16 > budget;
      t.glob = glob ? 1 : 0;
      // XXX takes one b12 row per launch: its third points are then cut to
      // b31 >= b12 (each unordered triplet deposited once, k_sym_cyc).
      // Copies of each V row in LDS: the FP64 weight atomics of lanes that
      // meet on one row are what conflict.  XXX's one row leaves LDS to spare
      // for 8; weighted XXY takes 4 at the price of more, narrower slices
      // (measured at 30 bins, rmax 80: weighted DDD 166 -> 129 ms, DDR
      // 362 -> 350 ms; unweighted XXY slower with copies, XXX unchanged).
      // The plain FP64 variant keeps one copy: copies made its weighted step
      // slower (4.5 -> 5.1 s).
      int cs = 0;
      if (!glob && !plain) {
        int cmax = 0;
        while (cmax < 3 && (per_b12 << (cmax + 1)) + 16 <= budget) ++cmax;
        cs = auto_ ? cmax : (weighted ? std::min(cmax, 2) : 0);
      }
      t.cs = cs;
      const int S =
          auto_ ? 1
                : (glob ? bins
                        : (int)std::max<size_t>(
                              1, std::min<size_t>(bins, (budget - 16) / (per_b12 << cs))));
      for (int lo = 0; lo < bins; lo += S) {
        t.lo = lo;
        t.hi = std::min(bins, lo + S);
        const int nrows = glob ? 0 : ((t.hi - t.lo) * bins * bins) << cs;
        const size_t lds =
            ((((size_t)nrows * 4 + 15) / 16 * 16 + (weighted ? (size_t)nrows * 8 : 0) + 15) &
             ~(size_t)15) + lists;
        HIPCHK(c, launch_zero(d_counter, sizeof(int), c->stream));
#define PCF_TRIP_ISO1(W, P, M, G)                                                              \
  do {                                                                                         \
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_triplets<0, W, P, M, G>),    \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));      \
    hipLaunchKernelGGL((k_triplets<0, W, P, M, G>), dim3(c->ncu), dim3(kTripBlock), lds,       \
                       c->stream, xs, gx->pts, ksc, gk->pts, gk->cell_start, d_pairs, d_off,   \
                       d_counter, d_ktab, d_T, d_V, d_VW, d_stats, t);                         \
  } while (0)
#define PCF_TRIP_ISO(W, P, M)               \
  do {                                      \
    if (glob) PCF_TRIP_ISO1(W, P, M, 1);    \
    else PCF_TRIP_ISO1(W, P, M, 0);         \
  } while (0)
        if (weighted && plain) PCF_TRIP_ISO(1, 1, 0);
        else if (weighted) PCF_TRIP_ISO(1, 0, 1);
        else if (plain) PCF_TRIP_ISO(0, 1, 0);
        else PCF_TRIP_ISO(0, 0, 1);
#undef PCF_TRIP_ISO
#undef PCF_TRIP_ISO1
        HIPCHK(c, hipGetLastError());
      }
    } else if (ani_lds) {
      // Third-point list: the pair list itself (XXX) or the X x K cross list.
      const PairRec* d_q = d_pairs;
      const int* d_qoff = d_off;
      int qstride = bins * bins;
      if (!auto_) {
        WorkListArgs wq = wa;
        wq.y_cell_start = gk->cell_start;
        wq.auto_ = 0;
        int nqitems = 0;
        HIPCHK(c, build_pair_worklist(wq, c->wl_counts, c->wl_items2, &nqitems, c->stream));
        const size_t qbytes = (sizeof(int) * (size_t)(3 * (nx + 1) + 8) + 15) & ~(size_t)15;
        HIPCHK(c, c->trip_qidx.ensure(qbytes + sizeof(long long) * (size_t)scan_tiles(nx)));
        int* q_cnt = c->trip_qidx.as<int>();
        int* q_off = q_cnt + (nx + 1);
        int* q_cur = q_off + (nx + 1);
        long long* q_tsum = reinterpret_cast<long long*>(c->trip_qidx.as<char>() + qbytes);
        HIPCHK(c, launch_zero(q_cnt, sizeof(int) * (size_t)(nx + 1), c->stream));
        HIPCHK(c, launch_zero(q_cur, sizeof(int) * (size_t)(nx + 1), c->stream));
        EmitArgs eq = ea;
        eq.nitems = nqitems;
        eq.key_t = 0;
        if (nqitems > 0) {
          hipLaunchKernelGGL(k_emit_cross<0>, dim3(emit_blocks), dim3(kEmitBlock), 0, c->stream,
                             xs, gx->pts, ksc, gk->pts, c->wl_items2.as<PairItem>(), d_ktab, d_T,
                             q_cnt, q_off, q_cur, (PairRec*)nullptr, eq);
          HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, launch_scan(q_cnt, nx, q_off, q_tsum, c->stream));
        int nq = 0;
        HIPCHK(c, hipMemcpyAsync(&nq, q_off + nx, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (nq < 0) return fail(c, PCF_ERR_UNSUPPORTED, "third-point list exceeds 2^31 records");
        HIPCHK(c, c->trip_q.ensure(sizeof(PairRec) * (size_t)std::max(nq, 1)));
        if (nqitems > 0 && nq > 0) {
          hipLaunchKernelGGL(k_emit_cross<1>, dim3(emit_blocks), dim3(kEmitBlock), 0, c->stream,
                             xs, gx->pts, ksc, gk->pts, c->wl_items2.as<PairItem>(), d_ktab, d_T,
                             q_cnt, q_off, q_cur, c->trip_q.as<PairRec>(), eq);
          HIPCHK(c, hipGetLastError());
        }
        d_q = c->trip_q.as<PairRec>();
        d_qoff = q_off;
        qstride = 1;
      }
      // Slices and their workgroups (proportional to the pairs in the slice).
      const int tn = (int)std::max<size_t>(1, std::min<size_t>(bins, ani_budget / ani_per_t));
      HIPCHK(c, c->misc_buf.ensure(sizeof(unsigned long long) * (size_t)bins * bins));
      unsigned long long* d_mass = c->misc_buf.as<unsigned long long>();
      HIPCHK(c, launch_zero(d_mass, sizeof(unsigned long long) * (size_t)bins * bins, c->stream));
      hipLaunchKernelGGL(k_key_mass, dim3(c->ncu * 4), dim3(256), 0, c->stream, d_off, nkeys,
                         bins * bins, d_mass);
      HIPCHK(c, hipGetLastError());
      std::vector<unsigned long long> mass((size_t)bins * bins);
      HIPCHK(c, hipMemcpyAsync(mass.data(), d_mass, sizeof(unsigned long long) * mass.size(),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      std::vector<AniSlice> slices;
      std::vector<double> wslice;
      double wtot = 0;
      for (int b12 = 0; b12 < bins; ++b12)
        for (int tlo = 0; tlo < bins; tlo += tn) {
          AniSlice sl;
          sl.b12 = b12;
          sl.tlo = tlo;
          sl.tn = std::min(tn, bins - tlo);
          sl.qstride = qstride;
          double w = 0;
          for (int q = tlo; q < tlo + sl.tn; ++q) w += (double)mass[(size_t)b12 * bins + q];
          if (w <= 0) continue;
          slices.push_back(sl);
          wslice.push_back(w);
          wtot += w;
        }
      std::vector<int> wg;
      const int G = c->ncu * 4;  // workgroups of the launch (about 4 rounds of one per CU)
      for (size_t q = 0; q < slices.size(); ++q) {
        const int nwg = std::max(1, (int)std::lround(G * wslice[q] / wtot));
        for (int r = 0; r < nwg; ++r) wg.push_back((int)q);
      }
      const size_t slice_bytes = sizeof(AniSlice) * std::max<size_t>(1, slices.size());
      const size_t sbytes = (slice_bytes + 15) & ~(size_t)15;
      HIPCHK(c, c->rules_buf.ensure(sbytes + sizeof(int) * std::max<size_t>(1, wg.size()) +
                                    sizeof(int) * (slices.size() + 1)));
      AniSlice* d_slices = c->rules_buf.as<AniSlice>();
      int* d_wg = reinterpret_cast<int*>(c->rules_buf.as<char>() + sbytes);
      int* d_ctr = d_wg + wg.size();
      if (!slices.empty()) {
        HIPCHK(c, hipMemcpyAsync(d_slices, slices.data(), sizeof(AniSlice) * slices.size(),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(d_wg, wg.data(), sizeof(int) * wg.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, launch_zero(d_ctr, sizeof(int) * slices.size(), c->stream));
        const size_t nrows = (size_t)bins * bins * tn * bins;
        const size_t lds = ((nrows * 4 + 15) & ~(size_t)15) + (weighted ? nrows * 8 : 0);
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
#define PCF_TRIP_ANI_LDS(W, P)                                                                 \
  do {                                                                                         \
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_trip_ani_lds<W, P>),         \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));      \
    hipLaunchKernelGGL((k_trip_ani_lds<W, P>), dim3((unsigned)wg.size()), dim3(kTripBlock),    \
                       lds, c->stream, gx->pts, gk->pts, d_pairs, d_off, d_q, d_qoff, d_ctr,   \
                       d_ktab, d_T, d_V, d_VW, d_stats, t, d_slices, d_wg);                    \
  } while (0)
        if (weighted && plain) PCF_TRIP_ANI_LDS(1, 1);
        else if (weighted) PCF_TRIP_ANI_LDS(1, 0);
        else if (plain) PCF_TRIP_ANI_LDS(0, 1);
        else PCF_TRIP_ANI_LDS(0, 0);
#undef PCF_TRIP_ANI_LDS
        HIPCHK(c, hipGetLastError());
      }
    } else {
      t.lo = 0;
      t.hi = bins;
      HIPCHK(c, launch_zero(d_counter, sizeof(int), c->stream));
      const int blocks = c->ncu * 4;
      const size_t lds = (size_t)(kAniTripBlock / kWave) * ListCap<1>::value * 24;
#define PCF_TRIP_ANI(W, P)                                                                     \
  hipLaunchKernelGGL((k_triplets<1, W, P, 0, 0>), dim3(blocks), dim3(kAniTripBlock), lds, c->stream, \
                     xs, gx->pts, ksc, gk->pts, gk->cell_start, d_pairs, d_off, d_counter,     \
                     d_ktab, d_T, d_V, d_VW, d_stats, t)
      if (weighted && plain) PCF_TRIP_ANI(1, 1);
      else if (weighted) PCF_TRIP_ANI(1, 0);
      else if (plain) PCF_TRIP_ANI(0, 1);
      else PCF_TRIP_ANI(0, 0);
#undef PCF_TRIP_ANI
      HIPCHK(c, hipGetLastError());
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  const unsigned long long* outH = d_V;
  const double* outW = d_VW;
  if (!ani) {
    // (The reference-weight kernel takes every third point: its V is
    // completed by the (b23, b31) transpose even for XXX.)
    if (auto_ && !refchain)
      hipLaunchKernelGGL(k_sym_cyc, dim3((unsigned)((nflat + 255) / 256)), dim3(256), 0,
                         c->stream, d_V, d_VW, bins, d_H, d_HW);
    else
      hipLaunchKernelGGL(k_sym_bc, dim3((unsigned)((nflat + 255) / 256)), dim3(256), 0,
                         c->stream, d_V, d_VW, bins, d_H, d_HW);
    HIPCHK(c, hipGetLastError());
    outH = d_H;
    outW = d_HW;
  }
  // Straight into the caller's arrays (3ani histograms reach gigabytes): an
  // unweighted job's weights are its counts, copied as integers and converted
  // in place.
  unsigned long long st2[2] = {0ull, 0ull};
  if (counts) HIPCHK(c, hipMemcpyAsync(counts, outH, hbytes, hipMemcpyDeviceToHost, c->stream));
  if (weights)
    HIPCHK(c, hipMemcpyAsync(weights, weighted ? (const void*)outW : (const void*)outH,
                             weighted ? wbytes : hbytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(st2, d_stats, sizeof(st2), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (weights && !weighted)
    for (int64_t b = 0; b < nflat; ++b) {
      uint64_t v;
      std::memcpy(&v, weights + b, sizeof(v));
      weights[b] = (double)v;
    }
  pcf_stats st{};
  st.evaluations = st2[0];
  st.rechecks = st2[1];
  st.work_items = (uint64_t)npairs;
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
  st.kernel_ms = ms;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
  st.total_ms = ms;
  st.cells_per_rmax = m;
  st.pairs_listed = (int32_t)npairs;
  c->last = st;
  return PCF_OK;
}

}  // namespace pcf
