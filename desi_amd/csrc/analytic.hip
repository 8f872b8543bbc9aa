// analytic.hip -- periodic 3iso with analytic random triplets (PCF 3iso -bpc -a).
//
// cuda/pcf3isoBPC.cuh:177-375 computes, besides the periodic DDD:
//   1. two fine periodic DD histograms of the data (ptt*bn and ptt*bn_ref*bn
//      bins over [0, dmax)), cuda/kernels/3isoBPC_k.cuh:1694;
//   2. the analytic uniform RR of the same binnings,
//      cuda/kernels/2isoBPC_k.cuh:559;
//   3. the bin averages ff_av (bn) and ff_av_ref (bn*bn_ref) of
//      r*(DD/RR - 1), cuda/kernels/3isoBPC_k.cuh:1984-2019;
//   4. per (i, j, k) bin cube the analytic RRR and DDR (DRR := DDR):
//      a closed form when all 8 corners satisfy the triangle inequality, a
//      sum over bn_ref^3 sub-cubes when only some do,
//      cuda/kernels/3isoBPC_k.cuh:2021-2119.
// Here: step 1 runs over the pair work list in FP64 (the canonical distance),
// the coarse fine-histogram in LDS and the finer one with global atomics;
// step 3 sums in index order (one lane per output, deterministic); step 4
// gives one workgroup to each bin cube, which for the partial cubes splits the
// sub-cube (u, v) plane over its lanes and reduces in LDS (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "exact.h"
#include "pcf_ctx.h"

namespace pcf {

namespace {

constexpr int kFineBlock = 256;
constexpr int kCubeBlock = 256;
constexpr int kFineLdsMax = 48 * 1024;  // bytes of the coarse fine histogram kept in LDS

struct FineArgs {
  int nitems;
  int nb1, nb2;
  double ds1, ds2;  // FP32 scales (float)nb/dmax, widened
  double R;         // (double)(float)(dmax*dmax)
  double L;
};

// Fine DD over the unordered pairs of the work list (mask bit: j > i inside a
// shared tile); the host doubles the sums.  Deposits w_i*w_j in double.
template <bool LDS1>
__global__ __launch_bounds__(kFineBlock) void k_fine_dd(const float4* __restrict__ pts,
                                                        const PairItem* __restrict__ items,
                                                        double* __restrict__ h1,
                                                        double* __restrict__ h2, FineArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* s1 = reinterpret_cast<double*>(smem);
  if (LDS1) {
    for (int t = threadIdx.x; t < a.nb1; t += kFineBlock) s1[t] = 0.0;
    __syncthreads();
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = (blockIdx.x * kFineBlock + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * (kFineBlock / kWave);
  for (int it = wave; it < a.nitems; it += nwaves) {
    const PairItem item = items[it];
    const int img = item.code & 31;
    const int kx = img / 9 - 1, ky = (img / 3) % 3 - 1, kz = img % 3 - 1;
    const bool masked = (item.code & kCodeMask) != 0;
    const bool active = lane < item.ni;
    const int my_i = item.i0 + lane;
    const float4 p = active ? pts[my_i] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j = item.j0; j < item.j0 + item.nj; ++j) {
      const float4 q = pts[j];
      double tx, ty, tz;
      bool ok = active && !(masked && j <= my_i);
      ok = ok && axis_sq(p.x, q.x, kx, a.L, &tx);
      ok = ok && axis_sq(p.y, q.y, ky, a.L, &ty);
      ok = ok && axis_sq(p.z, q.z, kz, a.L, &tz);
      if (!ok) continue;
      const double d2 = __dadd_rn(__dadd_rn(tx, ty), tz);
      if (!(d2 < a.R && d2 > 0.0)) continue;
      const double d = sqrt(d2);
      const double w = __dmul_rn((double)p.w, (double)q.w);
      const int b1 = (int)__dmul_rn(d, a.ds1);
      const int b2 = (int)__dmul_rn(d, a.ds2);
      if (b1 <= a.nb1 - 1) {
        if (LDS1) atomicAdd(&s1[b1], w);
        else atomicAdd(&h1[b1], w);
      }
      if (b2 <= a.nb2 - 1) atomicAdd(&h2[b2], w);
    }
  }
  if (LDS1) {
    __syncthreads();
    for (int t = threadIdx.x; t < a.nb1; t += kFineBlock)
      if (s1[t] != 0.0) atomicAdd(&h1[t], s1[t]);
  }
}

// Shell count 3q^2+3q+1 of the fine RR: 64-bit, or (wrap32, under
// pcf_set_ref_bugs) the reference's int, which wraps from q = 26755 on
// (cuda/kernels/2isoBPC_k.cuh:572).
__device__ __forceinline__ double shell_count(long long q, int wrap32) {
  if (wrap32) {
    const unsigned u = (unsigned)q;
    return (double)(int)(3u * u * u + 3u * u + 1u);
  }
  return (double)(3 * q * q + 3 * q + 1);
}

// ff_av[i] = sum_j (ri + rj) (DD/RR - 1) / ptt over the ptt fine bins of coarse
// bin i; RR = alpha*(3q^2+3q+1).  cuda/kernels/3isoBPC_k.cuh:1984.
__global__ void k_ff_av(const double* __restrict__ dd, double alpha, float dmax, int bins,
                        int ptt, double* __restrict__ ff, int wrap32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bins) return;
  const int nb1 = ptt * bins;
  const double ri = (double)((float)i * dmax) / (double)bins;
  double s = 0.0;
  for (int j = 0; j < ptt; ++j) {
    const double rj = __ddiv_rn(__dmul_rn((double)j + 0.5, (double)dmax), (double)nb1);
    const long long q = (long long)i * ptt + j;
    const double rr = __dmul_rn(alpha, shell_count(q, wrap32));
    s = __dadd_rn(s, __ddiv_rn(__dmul_rn(__dadd_rn(ri, rj), __dsub_rn(__ddiv_rn(dd[q], rr), 1.0)),
                               (double)ptt));
  }
  ff[i] = s;
}

// ff_av_ref[i*bn_ref + j] over the ptt finest bins of sub-bin (i, j);
// cuda/kernels/3isoBPC_k.cuh:2001.
__global__ void k_ff_av_ref(const double* __restrict__ dd, double alpha, float dmax, int bins,
                            int bn_ref, int ptt, double* __restrict__ ffr, int wrap32) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bins * bn_ref) return;
  const int i = t / bn_ref, j = t % bn_ref;
  const double dr = __ddiv_rn((double)dmax, (double)bins);
  const double dr_ref = __ddiv_rn(dr, (double)bn_ref);
  const double dr_ptt = __ddiv_rn((double)dmax, (double)((long long)ptt * bn_ref * bins));
  const double ri = __dmul_rn((double)i, dr), rj = __dmul_rn((double)j, dr_ref);
  double s = 0.0;
  for (int k = 0; k < ptt; ++k) {
    const double rk = __dmul_rn((double)k + 0.5, dr_ptt);
    const long long q = ((long long)i * bn_ref) * ptt + (long long)j * ptt + k;
    const double rr = __dmul_rn(alpha, shell_count(q, wrap32));
    s = __dadd_rn(s, __ddiv_rn(__dmul_rn(__dadd_rn(__dadd_rn(ri, rj), rk),
                                         __dsub_rn(__ddiv_rn(dd[q], rr), 1.0)),
                               (double)ptt));
  }
  ffr[t] = s;
}

// Number of the 8 corners (r1 + a h, r2 + b h, r3 + c h) that satisfy the
// triangle inequality (cuda/kernels/3isoBPC_k.cuh:2042-2055).
__device__ __forceinline__ int corners_in(double r1, double r2, double r3, double h) {
  int v = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double s1 = __dadd_rn(r1, __dmul_rn((double)a, h));
        const double s2 = __dadd_rn(r2, __dmul_rn((double)b, h));
        const double s3 = __dadd_rn(r3, __dmul_rn((double)c, h));
        v += (__dadd_rn(s1, s2) >= s3 && __dadd_rn(s1, s3) >= s2 && __dadd_rn(s2, s3) >= s1);
      }
  return v;
}

struct CubeArgs {
  int bins, bn_ref;
  double dr, dr2, dr_ref, dr_ref2, alpha, alpha_ref;
};

// One workgroup per bin cube (i, j, k), cube index = blockIdx.x.
__global__ __launch_bounds__(kCubeBlock) void k_cubes(const double* __restrict__ ff,
                                                      const double* __restrict__ ffr,
                                                      double* __restrict__ ddr,
                                                      double* __restrict__ rrr, CubeArgs a) {
  __shared__ double sS[kCubeBlock], sF[kCubeBlock];
  __shared__ int sCon[kCubeBlock];
  const int t = blockIdx.x;
  const int bins = a.bins;
  const int i = t / (bins * bins), j = (t / bins) % bins, k = t % bins;
  const double ri = __dmul_rn((double)i, a.dr), rj = __dmul_rn((double)j, a.dr),
               rk = __dmul_rn((double)k, a.dr);
  const int vin = corners_in(ri, rj, rk, a.dr);
  if (vin == 8) {
    if (threadIdx.x == 0) {
      const double hi = __dadd_rn(ri, a.dr2), hj = __dadd_rn(rj, a.dr2), hk = __dadd_rn(rk, a.dr2);
      const double s = __dmul_rn(__dmul_rn(__dmul_rn(a.alpha, hi), hj), hk);
      double f = 1.0;
      f = __dadd_rn(f, __ddiv_rn(ff[i], __dmul_rn(3.0, hi)));
      f = __dadd_rn(f, __ddiv_rn(ff[j], __dmul_rn(3.0, hj)));
      f = __dadd_rn(f, __ddiv_rn(ff[k], __dmul_rn(3.0, hk)));
      rrr[t] = s;
      ddr[t] = __dmul_rn(f, s);
    }
    return;
  }
  if (vin == 0) {
    if (threadIdx.x == 0) {
      rrr[t] = 0.0;
      ddr[t] = 0.0;
    }
    return;
  }
  const int nr = a.bn_ref;
  const double h = a.dr_ref, h2 = a.dr_ref2;
  double S = 0.0, F = 0.0;
  int con = 0;
  for (int uv = threadIdx.x; uv < nr * nr; uv += kCubeBlock) {
    const int u = uv / nr, v = uv % nr;
    const double ru = __dadd_rn(ri, __dmul_rn((double)u, h));
    const double rv = __dadd_rn(rj, __dmul_rn((double)v, h));
    const double hu = __dadd_rn(ru, h2), hv = __dadd_rn(rv, h2);
    const double fu = __ddiv_rn(ffr[i * nr + u], __dmul_rn(3.0, hu));
    const double fv = __ddiv_rn(ffr[j * nr + v], __dmul_rn(3.0, hv));
    for (int w = 0; w < nr; ++w) {
      const double rw = __dadd_rn(rk, __dmul_rn((double)w, h));
      if (corners_in(ru, rv, rw, h) != 8) continue;
      const double hw = __dadd_rn(rw, h2);
      const double c = __dmul_rn(__dmul_rn(hu, hv), hw);
      S = __dadd_rn(S, c);
      double f = 1.0;
      f = __dadd_rn(f, fu);
      f = __dadd_rn(f, fv);
      f = __dadd_rn(f, __ddiv_rn(ffr[k * nr + w], __dmul_rn(3.0, hw)));
      F = __dadd_rn(F, __dmul_rn(f, c));
      con = 1;
    }
  }
  sS[threadIdx.x] = S;
  sF[threadIdx.x] = F;
  sCon[threadIdx.x] = con;
  __syncthreads();
  for (int off = kCubeBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sS[threadIdx.x] = __dadd_rn(sS[threadIdx.x], sS[threadIdx.x + off]);
      sF[threadIdx.x] = __dadd_rn(sF[threadIdx.x], sF[threadIdx.x + off]);
      sCon[threadIdx.x] |= sCon[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rrr[t] = sCon[0] ? __dmul_rn(sS[0], a.alpha_ref) : 0.0;
    ddr[t] = sCon[0] ? __dmul_rn(sF[0], a.alpha_ref) : 0.0;
  }
}

// alpha of the fine analytic RR with nb bins (cuda/pcf3isoBPC.cuh:233-238):
// FP32 dr = dmax/nb and beta = np^2/L^3 as the reference forms them, np^2 in
// 64-bit (wrap32: the reference's 32-bit int, pcf_set_ref_bugs).
double fine_alpha(float dmax, int64_t nb, int64_t np, float box, bool wrap32) {
  const double dr = (double)(float)(dmax / (float)nb);
  const float np2 = wrap32 ? (float)(int32_t)(uint32_t)((uint64_t)np * (uint64_t)np)
                           : (float)((double)np * (double)np);
  const float L3 = box * box * box;
  const double beta = (double)(float)(np2 / L3);
  return 8 * dr * dr * dr * std::acos(0.0) * beta / 3;
}

}  // namespace

int run_analytic_3iso(pcf_ctx* c, pcf_catalog data, int bins, float dmax, float box, int64_t np,
                      int ptt, int bn_ref, double* ddr, double* rrr, double* ff_out,
                      double* ffr_out) {
  Catalog* X = get_cat(c, data);
  if (!X) return fail(c, PCF_ERR_ARG, "unknown catalog id");
  if (ptt <= 0) ptt = 100;
  if (bn_ref <= 0) bn_ref = 200;
  if (bins <= 0 || bins > 1024) return fail(c, PCF_ERR_ARG, "bins out of range");
  if (!(dmax > 0) || !std::isfinite(dmax)) return fail(c, PCF_ERR_ARG, "dmax must be > 0");
  if (!(box > 0)) return fail(c, PCF_ERR_ARG, "box side required for periodic boundaries");
  if (!(dmax < box))
    return fail(c, PCF_ERR_UNSUPPORTED, "periodic boundaries need dmax < box side");
  if (np <= 0) np = X->n;
  const int64_t nb1 = (int64_t)ptt * bins, nb2 = (int64_t)ptt * bn_ref * bins;
  if (nb2 > (1ll << 27) || (int64_t)bins * bn_ref > (1 << 24))
    return fail(c, PCF_ERR_ARG, "ptt*bn_ref*bins too large");
  for (int ax = 0; ax < 3; ++ax)
    if (X->n && (X->lo[ax] < 0.f || !(X->hi[ax] < box)))
      return fail(c, PCF_ERR_RANGE, "coordinates outside [0, box) with periodic boundaries");

  GridGeom g;
  int K[3];
  float reach;
  int m = 0;
  if (choose_geometry(X, nullptr, dmax, 1, box, 0, &g, K, &reach, &m) != 0)
    return fail(c, PCF_ERR_UNSUPPORTED, "geometry");
  int rc = PCF_OK;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  GridPins pins(c);
  DevGrid* gx = pins.add(get_grid(c, X, g, false, &rc));
  if (!gx) return rc;
  WorkListArgs wa;
  wa.x_cell_start = gx->cell_start;
  wa.y_cell_start = gx->cell_start;
  wa.g = g;
  for (int a = 0; a < 3; ++a) wa.K[a] = K[a];
  wa.reach = reach;
  wa.auto_ = 1;
  wa.cylinder = 0;
  wa.map = nullptr;
  wa.y_fine_start = nullptr;  // windows from whole z cells
  wa.x_pts = nullptr;
  wa.zq = 1;
  int nitems = 0;
  HIPCHK(c, build_pair_worklist(wa, c->wl_counts, c->wl_items, &nitems, c->stream));

  // Device layout (misc_buf): DD1[nb1] DD2[nb2] ff[bins] ffr[bins*bn_ref]
  // DDR[bins^3] RRR[bins^3].
  const int64_t b3 = (int64_t)bins * bins * bins;
  const size_t nd = (size_t)(nb1 + nb2 + bins + (int64_t)bins * bn_ref + 2 * b3);
  HIPCHK(c, c->misc_buf.ensure(sizeof(double) * nd));
  double* d_dd1 = c->misc_buf.as<double>();
  double* d_dd2 = d_dd1 + nb1;
  double* d_ff = d_dd2 + nb2;
  double* d_ffr = d_ff + bins;
  double* d_ddr = d_ffr + (int64_t)bins * bn_ref;
  double* d_rrr = d_ddr + b3;
  HIPCHK(c, launch_zero(d_dd1, sizeof(double) * (nb1 + nb2), c->stream));

  FineArgs fa;
  fa.nitems = nitems;
  fa.nb1 = (int)nb1;
  fa.nb2 = (int)nb2;
  fa.ds1 = (double)((float)nb1 / dmax);
  fa.ds2 = (double)((float)nb2 / dmax);
  fa.R = (double)(dmax * dmax);
  fa.L = (double)box;
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  if (nitems > 0) {
    const int nblocks = std::max(1, std::min(c->ncu * 8, (nitems + 3) / 4));
    const PairItem* items = c->wl_items.as<PairItem>();
    if (nb1 * 8 <= kFineLdsMax) {
      hipLaunchKernelGGL(k_fine_dd<true>, dim3(nblocks), dim3(kFineBlock), (size_t)nb1 * 8,
                         c->stream, gx->pts, items, d_dd1, d_dd2, fa);
    } else {
      hipLaunchKernelGGL(k_fine_dd<false>, dim3(nblocks), dim3(kFineBlock), 0, c->stream,
                         gx->pts, items, d_dd1, d_dd2, fa);
    }
    HIPCHK(c, hipGetLastError());
  }
  // The work list holds each unordered pair once and the ordered sums are
  // twice that.  Halving alpha instead is exact (a power of two), so
  // DD/(RR/2) equals (2 DD)/RR bit for bit.
  const bool wrap = ref_bugs(c);
  const double a1 = fine_alpha(dmax, nb1, np, box, wrap) * 0.5;
  const double a2 = fine_alpha(dmax, nb2, np, box, wrap) * 0.5;
  hipLaunchKernelGGL(k_ff_av, dim3((bins + 127) / 128), dim3(128), 0, c->stream, d_dd1, a1, dmax,
                     bins, ptt, d_ff, wrap ? 1 : 0);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_ff_av_ref, dim3((bins * bn_ref + 127) / 128), dim3(128), 0, c->stream,
                     d_dd2, a2, dmax, bins, bn_ref, ptt, d_ffr, wrap ? 1 : 0);
  HIPCHK(c, hipGetLastError());

  // Cube constants (cuda/pcf3isoBPC.cuh:240-245).
  CubeArgs ca;
  ca.bins = bins;
  ca.bn_ref = bn_ref;
  ca.dr = (double)dmax / (double)bins;
  ca.dr2 = ca.dr * 0.5;
  const double V = (double)(box * box);
  const double beta3 = (double)np / V;
  double gama = 8 * (4 * std::acos(0.0) * std::acos(0.0)) * ((double)np * beta3 * beta3);
  gama /= V;
  ca.dr_ref = ca.dr / bn_ref;
  ca.dr_ref2 = ca.dr_ref * 0.5;
  ca.alpha = gama * ca.dr * ca.dr * ca.dr;
  ca.alpha_ref = gama * ca.dr_ref * ca.dr_ref * ca.dr_ref;
  hipLaunchKernelGGL(k_cubes, dim3((unsigned)b3), dim3(kCubeBlock), 0, c->stream, d_ff, d_ffr,
                     d_ddr, d_rrr, ca);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  if (ddr) HIPCHK(c, hipMemcpyAsync(ddr, d_ddr, sizeof(double) * b3, hipMemcpyDeviceToHost, c->stream));
  if (rrr) HIPCHK(c, hipMemcpyAsync(rrr, d_rrr, sizeof(double) * b3, hipMemcpyDeviceToHost, c->stream));
  if (ff_out)
    HIPCHK(c, hipMemcpyAsync(ff_out, d_ff, sizeof(double) * bins, hipMemcpyDeviceToHost, c->stream));
  if (ffr_out)
    HIPCHK(c, hipMemcpyAsync(ffr_out, d_ffr, sizeof(double) * bins * bn_ref,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  pcf_stats st{};
  st.work_items = (uint64_t)nitems;
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
  st.kernel_ms = ms;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
  st.total_ms = ms;
  st.cells_per_rmax = m;
  c->last = st;
  return PCF_OK;
}

}  // namespace pcf

extern "C" int pcf_analytic_3iso(pcf_ctx* c, pcf_catalog data, int32_t bins, float dmax,
                                 float box, int64_t np, int32_t ptt, int32_t bn_ref, double* ddr,
                                 double* rrr, double* ff_av, double* ff_av_ref) {
  using namespace pcf;
  if (!c) return PCF_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  return pcf::run_analytic_3iso(c, data, bins, dmax, box, np, ptt, bn_ref, ddr, rrr, ff_av,
                                ff_av_ref);
}
