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
//   2. triplet work list: tiles of 64 third points (one per lane) against the
//      pair-list range of each neighbour column segment of point-1 cells;
//   3. triplet kernel: lanes hold the third point, the pair records stream
//      through scalar loads; both remaining sides are binned through the FP32
//      d2 -> bin table, lanes in an ambiguous cell (or at a cosine-bin edge)
//      take the exact FP64 path;
//   4. 3iso: pairs are unordered (i < j) and the third point is any point; the
//      ordered histogram is H[a][b][c] = V[a][b][c] + V[a][c][b] (exchanging
//      points 1 and 2 swaps the d23 and d31 sides).  3ani: the cosine of the
//      1-2 direction is not symmetric, so pairs are emitted in both orders.
//
// 3iso accumulates in LDS (u32 counts, FP64 weights) over slices of b12 so the
// slice fits; 3ani's 5-D histogram (bins^5) goes to global atomics.
#include "exact.h"
#include "pcf_ctx.h"
#include "worklist.h"

namespace pcf {

namespace {

constexpr int kMaxTripBins = 253;
constexpr int kTripBlock = 1024;    // 3iso: one workgroup per CU owns the LDS histogram
constexpr int kAniTripBlock = 256;  // 3ani: global histogram, many workgroups per CU
constexpr int kEmitBlock = 256;

struct PairRec {
  float xi, yi, zi;  // point 1, scaled by sigma
  float xj, yj, zj;  // point 2, scaled, shifted into point 1's frame
  float wi, wj;
  int i, j;          // cell-ordered indices in the X grid
  int meta;          // image code of point 2 (bits 0-4) | b12 << 8 | t << 20
  int pad;
};
static_assert(sizeof(PairRec) == 48, "pair record layout");

struct EmitArgs {
  int nitems;
  int nt;
  float Ls;       // box side times sigma
  int bins;
  int bpc;
  int ani;
  double R;
  double L;
  double ds_th;
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
  for (int t = threadIdx.x; t <= e.bins; t += kEmitBlock) sT[t] = T[t];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = (int)((blockIdx.x * (unsigned)kEmitBlock + threadIdx.x) >> 6);
  const int nwaves = (int)((gridDim.x * (unsigned)kEmitBlock) >> 6);
  const unsigned ntm1 = (unsigned)(e.nt - 1);
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
    int my_cnt = 0;
    for (int j = j0; j < j0 + nj; ++j) {
      const float4 q = xs[j];
      const float dx = q.x - sx, dy = q.y - sy, dz = q.z - sz;
      const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
      unsigned row = table[min(__float2uint_rz(d2), ntm1)];
      if (masked && j <= my_i) row = bins;
      if (!active) row = bins;
      int b12 = -1, t12 = 0, t21 = 0;
      // Everything emitted gets the exact bin: the pairs are few compared
      // with the triplets they feed.
      if (row < bins || row == amb_row) {
        const float4 qo = xo[j];
        b12 = exact_pair(po, qo, img, e.ani, e.bpc, sT, e.bins, e.R, e.L, e.ds_th, &t12);
        if (e.ani && b12 >= 0) {
          // the same pair seen from point 2 (3ani needs both orders)
          if (exact_pair(qo, po, img_neg(img), 1, e.bpc, sT, e.bins, e.R, e.L, e.ds_th, &t21) < 0)
            t21 = -1;
        }
        if (b12 >= 0) {
          if (PASS == 0) {
            if (!e.ani || t12 >= 0) ++my_cnt;
            if (e.ani && t21 >= 0) atomicAdd(&cnt[j], 1);
          } else {
            const int slot_i = (!e.ani || t12 >= 0) ? off[my_i] + atomicAdd(&cursor[my_i], 1) : 0;
            // Both passes take identical decisions, so slot_i < off[my_i + 1];
            // the bound is kept as a guard against writing past the owner's range.
            if ((!e.ani || t12 >= 0) && slot_i < off[my_i + 1]) {
              const int slot = slot_i;
              PairRec r;
              r.xi = ps.x; r.yi = ps.y; r.zi = ps.z;
              r.xj = q.x + kxL; r.yj = q.y + kyL; r.zj = q.z + kzL;
              r.wi = po.w; r.wj = qo.w;
              r.i = my_i; r.j = j;
              r.meta = img | (b12 << 8) | (t12 << 20);
              r.pad = 0;
              out[slot] = r;
            }
            const int slot_j = (e.ani && t21 >= 0) ? off[j] + atomicAdd(&cursor[j], 1) : 0;
            if (e.ani && t21 >= 0 && slot_j < off[j + 1]) {
              const int slot = slot_j;
              PairRec r;
              r.xi = q.x; r.yi = q.y; r.zi = q.z;
              r.xj = sx; r.yj = sy; r.zj = sz;
              r.wi = qo.w; r.wj = po.w;
              r.i = j; r.j = my_i;
              r.meta = img_neg(img) | (b12 << 8) | (t21 << 20);
              r.pad = 0;
              out[slot] = r;
            }
          }
        }
      }
    }
    if (PASS == 0 && my_cnt) atomicAdd(&cnt[my_i], my_cnt);
  }
}

struct TripArgs {
  int nitems;
  int nt;
  float Ls;
  int bins;
  int auto_;      // XXX: all sides > 0 and the third point is neither 1 nor 2
  int bpc;
  double R, L, ds_th;
  int lo, hi;     // b12 slice (3iso)
  float reach2s;  // (rmax + slack)^2 in scaled units
  float ds_th_f;
};

// Wave-wide min / max (butterfly over all 64 lanes).
__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

constexpr float kCosGuard = 1.0f / 1024.0f;  // cosine-bin edge band of the FP32 path

// ANI = 0: V histogram [b12 - lo][b23][b31] in LDS.  ANI = 1: the ordered 5-D
// histogram in global memory.
template <int ANI, int WEIGHTED>
__global__ __launch_bounds__(ANI ? kAniTripBlock : kTripBlock) void k_triplets(
    const float4* __restrict__ ks_all, const float4* __restrict__ ko_all,
    const float4* __restrict__ xo_all, const PairItem* __restrict__ items,
    int* __restrict__ counter, const PairRec* __restrict__ pairs,
    const uint8_t* __restrict__ table, const double* __restrict__ T,
    unsigned long long* __restrict__ hist, double* __restrict__ whist,
    unsigned long long* __restrict__ stats, TripArgs a) {
  constexpr int BLOCK = ANI ? kAniTripBlock : kTripBlock;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint8_t tab[kMixedTableMax];
  __shared__ double sT[kMaxTripBins + 2];
  const int bins = a.bins;
  const int bins2 = bins * bins;
  const int nrows = ANI ? 0 : (a.hi - a.lo) * bins2 + 1;  // + dummy row
  uint32_t* hc = reinterpret_cast<uint32_t*>(smem);
  double* hw = reinterpret_cast<double*>(smem + ((nrows * 4 + 15) & ~15));
  for (int t = threadIdx.x; t < a.nt; t += BLOCK) tab[t] = table[t];
  for (int t = threadIdx.x; t <= bins; t += BLOCK) sT[t] = T[t];
  if (!ANI) {
    for (int t = threadIdx.x; t < nrows; t += BLOCK) {
      hc[t] = 0u;
      if (WEIGHTED) hw[t] = 0.0;
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & (kWave - 1);
  const unsigned ntm1 = (unsigned)(a.nt - 1);
  const unsigned ubins = (unsigned)bins;
  const unsigned dummy = (unsigned)(nrows - 1);
  unsigned long long evals = 0, nre = 0;
  const int64_t b5 = ANI ? (int64_t)bins2 * bins2 * bins : 0;

  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= a.nitems) break;
    const int k0 = __builtin_amdgcn_readfirstlane(items[it].i0);
    const int nk = __builtin_amdgcn_readfirstlane(items[it].ni);
    const int p0 = __builtin_amdgcn_readfirstlane(items[it].j0);
    const int np = __builtin_amdgcn_readfirstlane(items[it].nj);
    const int img = __builtin_amdgcn_readfirstlane(items[it].code) & 31;
    const bool active = lane < nk;
    const int kidx = k0 + lane;
    float4 kp = make_float4(1e18f, 1e18f, 1e18f, 0.f), ko = kp;
    if (active) {
      kp = ks_all[kidx];
      ko = ko_all[kidx];
    }
    // The tile (third points) is shifted by the item's image: k' = k - s*L.
    const int sxi = img / 9 - 1, syi = (img / 3) % 3 - 1, szi = img % 3 - 1;
    const float kx = kp.x - (float)sxi * a.Ls, ky = kp.y - (float)syi * a.Ls,
                kz = kp.z - (float)szi * a.Ls;
    // Bounding box of the tile, for skipping pairs whose point 2 is out of reach.
    const float bxl = wave_min(active ? kx : 3e38f), bxh = wave_max(active ? kx : -3e38f);
    const float byl = wave_min(active ? ky : 3e38f), byh = wave_max(active ? ky : -3e38f);
    const float bzl = wave_min(active ? kz : 3e38f), bzh = wave_max(active ? kz : -3e38f);
    // 3ani, periodic: the reference's z term of the 1-3 side is |dz| - L*f3,
    // i.e. minus |dz'| on a projected axis (3aniBPC_k.cuh:73).
    const float zsign = (a.bpc && szi != 0) ? -1.f : 1.f;
    const double wk = (double)ko.w;

    for (int p = p0; p < p0 + np; ++p) {
      const PairRec r = pairs[p];
      const int b12 = (r.meta >> 8) & 0xfff;
      if (!ANI && (b12 < a.lo || b12 >= a.hi)) continue;
      {
        const float ex = fmaxf(0.f, fmaxf(bxl - r.xj, r.xj - bxh));
        const float ey = fmaxf(0.f, fmaxf(byl - r.yj, r.yj - byh));
        const float ez = fmaxf(0.f, fmaxf(bzl - r.zj, r.zj - bzh));
        if (ex * ex + ey * ey + ez * ez >= a.reach2s) continue;
      }
      evals += (unsigned long long)nk;
      const float ax = kx - r.xi, ay = ky - r.yi, az = kz - r.zi;
      const float d31 = __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax));
      const float cx = kx - r.xj, cy = ky - r.yj, cz = kz - r.zj;
      const float d23 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx));
      // Table codes: bin, bins = certainly out, bins+1 = ambiguous.  An
      // ambiguous side with the other certainly out is sent to the exact path
      // too (conservative).
      const unsigned c31 = tab[min(__float2uint_rz(d31), ntm1)];
      const unsigned c23 = tab[min(__float2uint_rz(d23), ntm1)];
      const unsigned m = max(c23, c31);
      unsigned row;
      bool amb = m > ubins;
      int pb = 0;
      if (!ANI) {
        row = m < ubins ? (unsigned)(b12 - a.lo) * bins2 + c23 * ubins + c31 : dummy;
      } else {
        // cosine bin of the 1-3 direction in FP32 (the z term is |dz| - L*f
        // when periodic); near a bin edge -> exact path
        const float dzz = a.bpc ? zsign * fabsf(az) : az;
        const float cth = __builtin_fmaf(dzz, __builtin_amdgcn_rsqf(fmaxf(d31, 1e-30f)), 1.f);
        const float pf = cth * a.ds_th_f;
        pb = (int)pf;
        const float fr = pf - (float)pb;
        if (m < ubins && (fr < kCosGuard || fr > 1.f - kCosGuard)) amb = true;
        row = (m < ubins && pb <= bins - 1) ? 0u : 1u;  // 0 = deposit, 1 = none
      }
      if (!active) amb = false;
      const unsigned long long ab = __ballot(amb);
      unsigned b23 = c23, b31 = c31;
      if (ab) {
        nre += __popcll(ab);
        if (amb) {
          bool ok = true;
          if (a.auto_ && (kidx == r.i || kidx == r.j)) ok = false;
          int f2 = 0, f3 = 0;
          float4 io = make_float4(0, 0, 0, 0), jo = io;
          if (ok) {
            io = xo_all[r.i];
            jo = xo_all[r.j];
            const int im2 = r.meta & 31;
            f2 = ((im2 / 9 - 1) != 0 ? 1 : 0) | (((im2 / 3) % 3 - 1) != 0 ? 2 : 0) |
                 ((im2 % 3 - 1) != 0 ? 4 : 0);
            // third point's image relative to point 1 is -s (k' = k - s*L)
            f3 = image_flags(io, ko, img, true);
            ok = f3 >= 0;
          }
          if (ok) {
            const float a1[3] = {io.x, io.y, io.z}, a2[3] = {jo.x, jo.y, jo.z},
                        a3[3] = {ko.x, ko.y, ko.z};
            double t31[3], t23[3];
            for (int q = 0; q < 3; ++q) {
              const double d31q = __dsub_rn((double)a3[q], (double)a1[q]);
              const double d23q = __dsub_rn((double)a3[q], (double)a2[q]);
              if (a.bpc) {
                const int g2 = (f2 >> q) & 1, g3 = (f3 >> q) & 1;
                t31[q] = proj_term(d31q, a.L, g3);
                t23[q] = proj_term(d23q, a.L, g2 != g3);
              } else {
                t31[q] = d31q;
                t23[q] = d23q;
              }
            }
            const double s31 =
                __dadd_rn(__dadd_rn(__dmul_rn(t31[0], t31[0]), __dmul_rn(t31[1], t31[1])),
                          __dmul_rn(t31[2], t31[2]));
            const double s23 =
                __dadd_rn(__dadd_rn(__dmul_rn(t23[0], t23[0]), __dmul_rn(t23[1], t23[1])),
                          __dmul_rn(t23[2], t23[2]));
            if (a.auto_ && !(s23 > 0.0 && s31 > 0.0)) ok = false;
            const int e23 = class_lds(s23, sT, bins, a.R, 0);
            const int e31 = class_lds(s31, sT, bins, a.R, 0);
            if (e23 < 0 || e31 < 0) ok = false;
            if (ok) {
              b23 = (unsigned)e23;
              b31 = (unsigned)e31;
              if (ANI) {
                const double cth = __dadd_rn(1.0, t31[2] / sqrt(s31));
                pb = (int)__dmul_rn(cth, a.ds_th);
                if (pb > bins - 1) ok = false;
              }
            }
          }
          if (!ANI) row = ok ? (unsigned)(b12 - a.lo) * bins2 + b23 * ubins + b31 : dummy;
          else row = ok ? 0u : 1u;
        }
      }
      if (!ANI) {
        atomicAdd(&hc[row], 1u);
        if (WEIGHTED)
          atomicAdd(&hw[row], __dmul_rn(__dmul_rn((double)r.wi, (double)r.wj), wk));
      } else if (row == 0u && active) {
        const int t12 = (r.meta >> 20) & 0xfff;
        const int64_t idx = ((((int64_t)b12 * bins + b31) * bins + b23) * bins + t12) * bins + pb;
        atomicAdd(&hist[idx], 1ull);
        if (WEIGHTED)
          atomicAdd(&whist[idx], __dmul_rn(__dmul_rn((double)r.wi, (double)r.wj), wk));
        (void)b5;
      }
    }
  }
  if (!ANI) {
    __syncthreads();
    const int base = a.lo * bins2;
    for (int t = threadIdx.x; t < nrows - 1; t += BLOCK) {
      if (hc[t]) {
        atomicAdd(&hist[base + t], (unsigned long long)hc[t]);
        if (WEIGHTED) atomicAdd(&whist[base + t], hw[t]);
      }
    }
  }
  if (lane == 0) atomicAdd(&stats[0], evals);
  if (lane == 0 && nre) atomicAdd(&stats[1], nre);
}

// H[a][b][c] = V[a][b][c] + V[a][c][b]
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

// Single-workgroup exclusive scan (counts -> offsets, offsets[n] = total).
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_scan(const int* counts, int64_t m, int* out) {
  __shared__ long long part[kScanBlock];
  const int t = threadIdx.x;
  const int64_t chunk = (m + kScanBlock - 1) / kScanBlock;
  const int64_t b = t * chunk;
  const int64_t e = b + chunk < m ? b + chunk : m;
  long long s = 0;
  for (int64_t i = b; i < e; ++i) s += counts[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    long long v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - s;
  for (int64_t i = b; i < e; ++i) {
    int c = counts[i];
    out[i] = (int)run;
    run += c;
  }
  if (t == kScanBlock - 1) out[m] = (int)part[t];
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
  if (bins <= 0 || bins > kMaxTripBins || (ani && bins > 40))
    return fail(c, PCF_ERR_ARG, "bins out of range");
  if (!(job->dmax > 0) || !std::isfinite(job->dmax)) return fail(c, PCF_ERR_ARG, "dmax must be > 0");
  const int bpc = job->bpc ? 1 : 0;
  const float L = job->box;
  if (bpc) {
    if (!(L > 0)) return fail(c, PCF_ERR_ARG, "box side required for periodic boundaries");
    if (!(2.0 * job->dmax < L))
      return fail(c, PCF_ERR_UNSUPPORTED, "periodic boundaries need 2*dmax < box side");
    for (Catalog* cc : {X, K})
      for (int a = 0; a < 3; ++a)
        if (cc->n && (cc->lo[a] < 0.f || !(cc->hi[a] < L)))
          return fail(c, PCF_ERR_RANGE, "coordinates outside [0, box) with periodic boundaries");
  }
  const int64_t b64 = bins;
  const int64_t nflat = ani ? b64 * b64 * b64 * b64 * b64 : b64 * b64 * b64;
  const bool weighted = !(X->all_w1 && K->all_w1);

  GridGeom g;
  int Kc[3];
  float reach;
  int m = 0;
  if (choose_geometry(X, auto_ ? nullptr : K, job->dmax, bpc, L, job->cells_per_rmax, &g, Kc,
                      &reach, &m) != 0)
    return fail(c, PCF_ERR_UNSUPPORTED, "geometry");
  int rc = PCF_OK;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  DevGrid* gx = get_grid(c, X, g, job->reserved[0] != 0, &rc);
  if (!gx) return rc;
  DevGrid* gk = auto_ ? gx : get_grid(c, K, g, job->reserved[0] != 0, &rc);
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
  if (build_table(pr, A, kMixedTableMax, &ptab, &inv_h, &nt) != 0 ||
      build_table(kr, A, kMixedTableMax, &ktab, &inv_h2, &nt2) != 0 || inv_h != inv_h2 ||
      nt != nt2)
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

  // 1. Pair list over X (unordered pairs i < j).
  WorkListArgs wa;
  wa.x_cell_start = gx->cell_start;
  wa.y_cell_start = gx->cell_start;
  wa.g = g;
  for (int a = 0; a < 3; ++a) wa.K[a] = Kc[a];
  wa.reach = reach;
  wa.auto_ = 1;
  wa.cylinder = 0;
  wa.map = nullptr;
  int npitems = 0;
  HIPCHK(c, build_pair_worklist(wa, c->wl_counts, c->wl_items, &npitems, c->stream));
  const int64_t nx = gx->n;
  HIPCHK(c, c->trip_idx.ensure(sizeof(int) * (size_t)(3 * (nx + 1) + 8)));
  int* d_cnt = c->trip_idx.as<int>();
  int* d_off = d_cnt + (nx + 1);
  int* d_cur = d_off + (nx + 1);
  HIPCHK(c, launch_zero(d_cnt, sizeof(int) * (size_t)(nx + 1), c->stream));
  HIPCHK(c, launch_zero(d_cur, sizeof(int) * (size_t)(nx + 1), c->stream));
  EmitArgs ea;
  ea.nitems = npitems;
  ea.nt = nt;
  ea.Ls = L * sigma;
  ea.bins = bins;
  ea.bpc = bpc;
  ea.ani = ani ? 1 : 0;
  ea.R = pr.R;
  ea.L = (double)L;
  float ds_th_f = (float)bins / 2;  // 3ani_k.cuh:42
  ea.ds_th = (double)ds_th_f;
  const int emit_blocks = std::max(1, c->ncu * 8);
  if (npitems > 0) {
    hipLaunchKernelGGL(k_emit_pairs<0>, dim3(emit_blocks), dim3(kEmitBlock), 0, c->stream, xs,
                       gx->pts, c->wl_items.as<PairItem>(), d_ptab, d_T, d_cnt, d_off, d_cur,
                       (PairRec*)nullptr, ea);
    HIPCHK(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanBlock), 0, c->stream, d_cnt, nx, d_off);
  HIPCHK(c, hipGetLastError());
  int npairs = 0;
  HIPCHK(c, hipMemcpyAsync(&npairs, d_off + nx, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (npairs < 0) return fail(c, PCF_ERR_UNSUPPORTED, "pair list exceeds 2^31 records");
  HIPCHK(c, c->trip_pairs.ensure(sizeof(PairRec) * (size_t)std::max(npairs, 1)));
  PairRec* d_pairs = c->trip_pairs.as<PairRec>();
  if (npitems > 0 && npairs > 0) {
    hipLaunchKernelGGL(k_emit_pairs<1>, dim3(emit_blocks), dim3(kEmitBlock), 0, c->stream, xs,
                       gx->pts, c->wl_items.as<PairItem>(), d_ptab, d_T, d_cnt, d_off, d_cur,
                       d_pairs, ea);
    HIPCHK(c, hipGetLastError());
  }

  // 2. Triplet work list: tiles of third points against pair ranges.
  WorkListArgs ta;
  ta.x_cell_start = gk->cell_start;
  ta.y_cell_start = gx->cell_start;
  ta.g = g;
  for (int a = 0; a < 3; ++a) ta.K[a] = Kc[a];
  ta.reach = reach;
  ta.auto_ = 0;
  ta.cylinder = 0;
  ta.map = d_off;
  int ntitems = 0;
  if (npairs > 0)
    HIPCHK(c, build_pair_worklist(ta, c->wl_counts, c->wl_items2, &ntitems, c->stream));

  // 3. Histograms.
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
  t.nitems = ntitems;
  t.nt = nt;
  t.Ls = L * sigma;
  t.bins = bins;
  t.auto_ = auto_ ? 1 : 0;
  t.bpc = bpc;
  t.R = kr.R;
  t.L = (double)L;
  t.ds_th = (double)ds_th_f;
  t.ds_th_f = ds_th_f;
  t.reach2s = (float)((double)reach * reach * (double)inv_h);
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  if (ntitems > 0) {
    if (!ani) {
      // LDS budget of one 1024-thread workgroup per CU: the slice of b12 rows
      // whose counts (and FP64 weights) fit beside the table.
      const size_t budget = 160 * 1024 - kMixedTableMax - 8 * (kMaxTripBins + 2) - 1024;
      const size_t per_b12 = (size_t)bins * bins * (4 + (weighted ? 8 : 0));
      const int S = (int)std::max<size_t>(1, std::min<size_t>(bins, (budget - 16) / per_b12));
      for (int lo = 0; lo < bins; lo += S) {
        t.lo = lo;
        t.hi = std::min(bins, lo + S);
        const int nrows = (t.hi - t.lo) * bins * bins + 1;
        const size_t lds = ((size_t)nrows * 4 + 15) / 16 * 16 + (weighted ? (size_t)nrows * 8 : 0);
        HIPCHK(c, launch_zero(d_counter, sizeof(int), c->stream));
        if (weighted) {
          HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_triplets<0, 1>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
          hipLaunchKernelGGL((k_triplets<0, 1>), dim3(c->ncu), dim3(kTripBlock), lds, c->stream,
                             ksc, gk->pts, gx->pts, c->wl_items2.as<PairItem>(), d_counter,
                             d_pairs, d_ktab, d_T, d_V, d_VW, d_stats, t);
        } else {
          HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_triplets<0, 0>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
          hipLaunchKernelGGL((k_triplets<0, 0>), dim3(c->ncu), dim3(kTripBlock), lds, c->stream,
                             ksc, gk->pts, gx->pts, c->wl_items2.as<PairItem>(), d_counter,
                             d_pairs, d_ktab, d_T, d_V, d_VW, d_stats, t);
        }
        HIPCHK(c, hipGetLastError());
      }
    } else {
      t.lo = 0;
      t.hi = bins;
      HIPCHK(c, launch_zero(d_counter, sizeof(int), c->stream));
      const int blocks = c->ncu * 4;
      if (weighted)
        hipLaunchKernelGGL((k_triplets<1, 1>), dim3(blocks), dim3(kAniTripBlock), 0, c->stream,
                           ksc, gk->pts, gx->pts, c->wl_items2.as<PairItem>(), d_counter, d_pairs,
                           d_ktab, d_T, d_V, d_VW, d_stats, t);
      else
        hipLaunchKernelGGL((k_triplets<1, 0>), dim3(blocks), dim3(kAniTripBlock), 0, c->stream,
                           ksc, gk->pts, gx->pts, c->wl_items2.as<PairItem>(), d_counter, d_pairs,
                           d_ktab, d_T, d_V, d_VW, d_stats, t);
      HIPCHK(c, hipGetLastError());
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  const unsigned long long* outH = d_V;
  const double* outW = d_VW;
  if (!ani) {
    hipLaunchKernelGGL(k_sym_bc, dim3((unsigned)((nflat + 255) / 256)), dim3(256), 0, c->stream,
                       d_V, d_VW, bins, d_H, d_HW);
    HIPCHK(c, hipGetLastError());
    outH = d_H;
    outW = d_HW;
  }
  std::vector<unsigned long long> hc((size_t)nflat + 4);
  std::vector<double> hw((size_t)nflat);
  HIPCHK(c, hipMemcpyAsync(hc.data(), outH, hbytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hw.data(), outW, wbytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hc.data() + nflat, d_stats, 2 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t b = 0; b < nflat; ++b) {
    if (counts) counts[b] = (uint64_t)hc[b];
    if (weights) weights[b] = weighted ? hw[b] : (double)hc[b];
  }
  pcf_stats st{};
  st.evaluations = hc[nflat];
  st.rechecks = hc[nflat + 1];
  st.work_items = (uint64_t)ntitems;
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
  st.kernel_ms = ms;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
  st.total_ms = ms;
  st.cells_per_rmax = m;
  st.reserved[0] = npairs;
  c->last = st;
  return PCF_OK;
}

}  // namespace pcf
