// exact.h -- the FP64 arithmetic of the CPU oracle, for the in-kernel exact
// path of the pair and triplet kernels.  The library is compiled with
// -ffp-contract=off (HIP's __dadd_rn/__dmul_rn are plain + and *, which the
// default contraction would fuse into FMAs): every operation rounds as written.
#pragma once
#include "pcf_internal.h"

namespace pcf {

// ---------------------------------------------------------------------------
// Exact FP64 re-evaluation (the oracle's arithmetic, no contraction).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool axis_sq(float ai, float aj, int k, double L, double* out) {
  double d = __dsub_rn((double)aj, (double)ai);
  if (k == 0) {
    *out = __dmul_rn(d, d);
    return true;
  }
  // Only the image that brings the points closer is a reference term
  // (size_box - |delta|, cuda/kernels/2isoBPC_k.cuh:87).
  if ((k > 0 && d >= 0.0) || (k < 0 && d < 0.0)) return false;
  double t = __dsub_rn(L, fabs(d));
  *out = __dmul_rn(t, t);
  return true;
}

// Bin of one exact squared separation against thresholds held in LDS.
__device__ __forceinline__ int class_lds(double d2, const double* T, int bins, double R, int le) {
  if (le ? !(d2 <= R) : !(d2 < R)) return -1;
  if (d2 >= T[bins]) return -1;
  int lo = 0, hi = bins - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T[mid] <= d2) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}


// Periodic triplet term of the reference: |delta| - L*f (f = 1 when the axis is
// projected; for the 2-3 side f = |f2 - f3|), cuda/kernels/3isoBPC_k.cuh:56-75.
__device__ __forceinline__ double proj_term(double d, double L, int f) {
  return __dsub_rn(fabs(d), f ? L : 0.0);
}

// Projection flags (3-bit axis set) of the image `img` (code (kx+1)*9 +
// (ky+1)*3 + (kz+1)) of point q relative to point p, or -1 when img is the far
// image on some axis (the reference only projects toward the close image).
// `neg` flips the image (used when the image code describes p relative to q).
__device__ __forceinline__ int image_flags(const float4 p, const float4 q, int img, bool neg) {
  int k[3] = {img / 9 - 1, (img / 3) % 3 - 1, img % 3 - 1};
  const float a1[3] = {p.x, p.y, p.z}, a2[3] = {q.x, q.y, q.z};
  int f = 0;
  for (int a = 0; a < 3; ++a) {
    const int kk = neg ? -k[a] : k[a];
    if (kk == 0) continue;
    const double d = __dsub_rn((double)a2[a], (double)a1[a]);
    if ((kk > 0 && d >= 0.0) || (kk < 0 && d < 0.0)) return -1;
    f |= 1 << a;
  }
  return f;
}

}  // namespace pcf
