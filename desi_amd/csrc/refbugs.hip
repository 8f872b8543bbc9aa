// refbugs.hip -- the reference's node order and the defects that depend on it,
// reproduced on request (pcf_set_ref_bugs; the command line's --ref-bugs).
//
// The reference copies each catalog into partitions^3 nodes of side
// size_box / partitions (FP32, cuda/main.cu:216): a point goes to node
// ((int)(x / size_node), (int)(y / size_node), (int)(z / size_node))
// (cuda/create_grid.cuh:133-135), add() appends it (file order inside a node,
// create_grid.cuh:81-102), and the nodes are laid out row-major in (x, y, z)
// (cuda/main.cu:276-300).  Two defects read that order:
//   - the triplet kernels' cumulative weight (v *= w3 along a node's points,
//     cuda/kernels/3iso_k.cuh:79), reproduced by k_trip_refbugs (triplets.hip)
//     over the ordered third-point catalog built here;
//   - XY2ani's weight (v = w1 * elementsD[j].w with j an index of the random
//     catalog's ordered array, cuda/kernels/2ani_k.cuh:128), reproduced by
//     replacing the random catalog's weights with the data weights that read
//     gives (ref_alias_catalog).
// The order is a stable radix sort of the node ids (hipCUB), so points of one
// node keep their upload (file) order.  A coordinate whose FP32 quotient
// reaches past the last node (x just below the box side) is clamped into it,
// where the reference indexes out of bounds (cuda/create_grid.cuh:133-135).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "pcf_ctx.h"

namespace pcf {
namespace {

__global__ void k_ref_node(const float4* __restrict__ pts, int64_t n, float size_node, int P,
                           unsigned* __restrict__ key, int* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pts[i];
  auto ax = [&](float v) {
    int c = (int)(v / size_node);  // IEEE FP32 division, as the host reference does
    return c < 0 ? 0 : (c >= P ? P - 1 : c);
  };
  key[i] = (unsigned)(((int64_t)ax(p.x) * P + ax(p.y)) * P + ax(p.z));
  idx[i] = (int)i;
}

__global__ void k_node_counts(const unsigned* __restrict__ key, int64_t n, int* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[key[i]], 1);
}

__global__ void k_gather(const float4* __restrict__ pts, const int* __restrict__ ord, int64_t n,
                         float4* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = pts[ord[q]];
}

// Random point ord_r[q] (slot q of the random catalog's ordered array) takes
// the weight of the data point in slot q of the data catalog's ordered array.
__global__ void k_alias(const float4* __restrict__ r, const int* __restrict__ ord_r,
                        const float4* __restrict__ d, const int* __restrict__ ord_d, int64_t n,
                        float4* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int j = ord_r[q];
  float4 p = r[j];
  p.w = d[ord_d[q]].w;
  out[j] = p;
}

int grid_blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

int ref_partitions(int partitions) { return partitions > 0 ? partitions : 35; }

int build_ref_order(pcf_ctx* c, const Catalog* X, float box, int partitions, DevBuf* order,
                    DevBuf* node_start, DevBuf* sorted_pts) {
  const int P = ref_partitions(partitions);
  if (!(box > 0)) return fail(c, PCF_ERR_ARG, "reference node order: box side required");
  if ((int64_t)P * P * P >= ((int64_t)1 << 31))
    return fail(c, PCF_ERR_ARG, "reference node order: too many partitions");
  const int64_t n = X->n;
  const int64_t nodes = (int64_t)P * P * P;
  const float size_node = box / (float)P;  // FP32, cuda/main.cu:216
  HIPCHK(c, order->ensure(sizeof(int) * (size_t)std::max<int64_t>(n, 1)));
  HIPCHK(c, node_start->ensure(sizeof(int) * (size_t)(nodes + 1) +
                               sizeof(long long) * (size_t)scan_tiles(nodes)));
  if (sorted_pts) HIPCHK(c, sorted_pts->ensure(sizeof(float4) * (size_t)std::max<int64_t>(n, 1)));
  int* d_order = order->as<int>();
  int* d_start = node_start->as<int>();
  long long* d_tsum = reinterpret_cast<long long*>(
      node_start->as<char>() + sizeof(int) * (size_t)(nodes + 1));
  // Temporaries: keys and indices in and out, the sort's workspace, counts.
  ScopedDevBuf tmp;
  const size_t kb = (sizeof(unsigned) * (size_t)std::max<int64_t>(n, 1) + 255) & ~(size_t)255;
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < nodes) ++bits;
  size_t sort_bytes = 0;
  HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (unsigned*)nullptr,
                                               (unsigned*)nullptr, (int*)nullptr, (int*)nullptr,
                                               (int)n, 0, bits, c->stream));
  const size_t cb = (sizeof(int) * (size_t)(nodes + 1) + 255) & ~(size_t)255;
  HIPCHK(c, tmp.ensure(3 * kb + cb + sort_bytes + 256));
  char* t = tmp.as<char>();
  unsigned* key_in = reinterpret_cast<unsigned*>(t);
  unsigned* key_out = reinterpret_cast<unsigned*>(t + kb);
  int* idx_in = reinterpret_cast<int*>(t + 2 * kb);
  int* cnt = reinterpret_cast<int*>(t + 3 * kb);
  void* sort_tmp = t + 3 * kb + cb;
  int rc = PCF_OK;
  if (n > 0) {
    hipLaunchKernelGGL(k_ref_node, dim3(grid_blocks(n)), dim3(256), 0, c->stream, X->raw, n,
                       size_node, P, key_in, idx_in);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, key_in, key_out, idx_in,
                                                 d_order, (int)n, 0, bits, c->stream));
  }
  HIPCHK(c, launch_zero(cnt, sizeof(int) * (size_t)(nodes + 1), c->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_node_counts, dim3(grid_blocks(n)), dim3(256), 0, c->stream, key_in, n,
                       cnt);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, launch_scan(cnt, nodes, d_start, d_tsum, c->stream));
  if (sorted_pts && n > 0) {
    hipLaunchKernelGGL(k_gather, dim3(grid_blocks(n)), dim3(256), 0, c->stream, X->raw, d_order,
                       n, sorted_pts->as<float4>());
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int ref_alias_catalog(pcf_ctx* c, pcf_catalog data, pcf_catalog rand, float box, int partitions,
                      pcf_catalog* out) {
  Catalog* D = get_cat(c, data);
  Catalog* R = get_cat(c, rand);
  if (!D || !R) return fail(c, PCF_ERR_ARG, "unknown catalog id");
  if (R->n > D->n)
    return fail(c, PCF_ERR_UNSUPPORTED,
                "--ref-bugs 2ani DR: the random catalog is longer than the data catalog, and "
                "the reference reads its weights past the data array");
  ScopedDevBuf od, orr, sd, sr, alias;
  int rc = build_ref_order(c, D, box, partitions, &od, &sd, nullptr);
  if (rc != PCF_OK) return rc;
  rc = build_ref_order(c, R, box, partitions, &orr, &sr, nullptr);
  if (rc != PCF_OK) return rc;
  HIPCHK(c, alias.ensure(sizeof(float4) * (size_t)std::max<int64_t>(R->n, 1)));
  if (R->n > 0) {
    hipLaunchKernelGGL(k_alias, dim3(grid_blocks(R->n)), dim3(256), 0, c->stream, R->raw,
                       orr.as<int>(), D->raw, od.as<int>(), R->n, alias.as<float4>());
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return pcf_catalog_upload_device(c, alias.p, R->n, out);
}

}  // namespace pcf

extern "C" int pcf_set_ref_bugs(pcf_ctx* c, int32_t on) {
  if (!c) return PCF_ERR_ARG;
  c->store->ref_bugs.store(on != 0);
  c->err.clear();
  return PCF_OK;
}
