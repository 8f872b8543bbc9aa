// estimators.hip -- correlation-function estimators from finished histograms
// (pcf_estimator), the step that follows the histograms in the reference's
// analysis scripts:
//
//   2PCF Landy-Szalay   xi = (DD - 2 DR + RR) / RR
//       cpp/2PCFisotropic/normal/graph_2DPF.py:8-14 (estim_LS; every zero of
//       DD, RR and DR is replaced by 1e-11 before the division)
//   3PCF                xi = (DDD - 3 DDR + 3 DRR - RRR) / RRR
//       cpp/3PCFisotropic/normal/graph_3PCF.py:5-6,21-22 (estim_LS; NaN and
//       infinite results are set to 0)
//
// The scripts use the raw counts, which is right when data and random
// catalogs have the same size.  With PCF_EST_NORMALIZE each histogram is
// first divided by its number of ordered pairs (triplets) of distinct points:
// DD by N_D (N_D - 1), DR by N_D N_R, RR by N_R (N_R - 1); DDD by N_D (N_D - 1)
// (N_D - 2), DDR by N_D (N_D - 1) N_R, DRR by N_D N_R (N_R - 1), RRR by N_R
// (N_R - 1) (N_R - 2).  Elementwise FP64 on the device, no contraction
// (the library builds with -ffp-contract=off), correctly rounded division.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "pcf_ctx.h"

namespace pcf {
namespace {

__global__ void k_estimator(int order, int64_t n, const double* __restrict__ h,
                            double s0, double s1, double s2, double s3, int flags,
                            double* __restrict__ xi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool norm = (flags & PCF_EST_NORMALIZE) != 0;
  double a = h[i], b = h[n + i], c = h[2 * n + i];
  if (norm) {
    a /= s0;
    b /= s1;
    c /= s2;
  }
  double r;
  if (order == 2) {
    // a = DD, b = DR, c = RR
    if (a == 0.0) a = 1e-11;
    if (c == 0.0) c = 1e-11;
    if (b == 0.0) b = 1e-11;
    r = (a - 2.0 * b + c) / c;
  } else {
    // a = DDD, b = DDR, c = DRR, d = RRR
    double d = h[3 * n + i];
    if (norm) d /= s3;
    r = (a - 3.0 * b + 3.0 * c - d) / d;
    if (isnan(r) || isinf(r)) r = 0.0;
  }
  xi[i] = r;
}

}  // namespace
}  // namespace pcf

extern "C" int pcf_estimator(pcf_ctx* c, int32_t order, int64_t n, const double* h0,
                             const double* h1, const double* h2, const double* h3,
                             double n_data, double n_rand, int32_t flags, double* xi) {
  using namespace pcf;
  if (!c) return PCF_ERR_ARG;
  if ((order != 2 && order != 3) || n < 0 || !h0 || !h1 || !h2 || (order == 3 && !h3) ||
      (n > 0 && !xi) || (flags & ~PCF_EST_NORMALIZE))
    return fail(c, PCF_ERR_ARG, "bad estimator arguments");
  double s[4] = {1, 1, 1, 1};
  if (flags & PCF_EST_NORMALIZE) {
    const double nd = n_data, nr = n_rand;
    if (!(nd >= order) || !(nr >= order))
      return fail(c, PCF_ERR_ARG, "normalisation needs at least `order` points per catalog");
    if (order == 2) {
      s[0] = nd * (nd - 1);
      s[1] = nd * nr;
      s[2] = nr * (nr - 1);
    } else {
      s[0] = nd * (nd - 1) * (nd - 2);
      s[1] = nd * (nd - 1) * nr;
      s[2] = nd * nr * (nr - 1);
      s[3] = nr * (nr - 1) * (nr - 2);
    }
  }
  c->err.clear();
  if (n == 0) return PCF_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int k = order == 2 ? 3 : 4;
  const size_t bytes = sizeof(double) * (size_t)n;
  // Inputs and result through page-locked staging, one upload and one read-back.
  HIPCHK(c, c->pin.ensure(bytes * (size_t)(k + 1)));
  HIPCHK(c, c->misc_buf.ensure(bytes * (size_t)(k + 1)));
  double* st = c->pin.as<double>();
  const double* hs[4] = {h0, h1, h2, h3};
  for (int q = 0; q < k; ++q) std::memcpy(st + (size_t)q * n, hs[q], bytes);
  double* dev = c->misc_buf.as<double>();
  HIPCHK(c, hipMemcpyAsync(dev, st, bytes * k, hipMemcpyHostToDevice, c->stream));
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_estimator, dim3((unsigned)blocks), dim3(256), 0, c->stream, (int)order, n,
                     dev, s[0], s[1], s[2], s[3], (int)flags, dev + (size_t)k * n);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(st + (size_t)k * n, dev + (size_t)k * n, bytes, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::memcpy(xi, st + (size_t)k * n, bytes);
  return PCF_OK;
}
