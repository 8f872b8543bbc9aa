// noalloc.hip -- pcf_histogram_noalloc: a pair histogram whose device buffers
// all live in one file-scope device array, so the call allocates nothing.
//
// For callers that cannot allocate device memory (tight serving loops,
// harnesses that forbid hipMalloc).  The inputs are resident catalogs
// (pcf_catalog_upload) and the outputs the caller's host arrays, as for
// pcf_histogram; the grids, scaled copies, work list, rules, bin table and
// device histogram of the call are carved from `scratch`.
#include <mutex>
#include <string>

#include "../../include/pcf_abi.h"
#include "pcf_ctx.h"
#include "pcf_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)PCF_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

}  // namespace

// One copy per GPU (each device that loads this code object holds its own
// array at its own address, until the process ends).  Host code never passes
// `scratch` itself to a kernel: in host code the name is a stand-in, not a
// device address; the address comes from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct DeviceScratch {
  void* ptr = nullptr;     // device address of this GPU's `scratch`
  bool resolved = false;
  std::mutex call;         // calls on one GPU share the array: one at a time
};
std::mutex g_resolve_mu;
DeviceScratch g_scratch[kMaxDevices];

}  // namespace

namespace pcf {

// Resolves the device address of this GPU's scratch array, once per device,
// with `device` current (pcf_create calls it after hipSetDevice).
hipError_t resolve_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_resolve_mu);
  DeviceScratch& d = g_scratch[device];
  if (d.resolved) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return e;
  }
  d.ptr = p;
  d.resolved = true;
  return hipSuccess;
}

}  // namespace pcf

extern "C" int pcf_histogram_noalloc(pcf_ctx* c, const pcf_job* job, uint64_t* counts,
                                     double* weights) {
  using namespace pcf;
  if (!c || !job) return fail(c, PCF_ERR_ARG, "null argument");
  c->err.clear();
  if (job->type != PCF_2ISO && job->type != PCF_2ANI)
    return fail(c, PCF_ERR_UNSUPPORTED, "pcf_histogram_noalloc computes pair histograms only");
  if (job->hist != PCF_XX && job->hist != PCF_XY)
    return fail(c, PCF_ERR_ARG, "pair types take PCF_XX or PCF_XY");
  if (ref_bugs_alias_needed(c, job))
    return fail(c, PCF_ERR_UNSUPPORTED,
                "reference 2ani DR weights (pcf_set_ref_bugs) allocate; use pcf_histogram");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(c, PCF_ERR_ARG, "device index out of range");
  DeviceScratch& d = g_scratch[c->device];
  {
    std::lock_guard<std::mutex> lk(g_resolve_mu);
    if (!d.resolved) return fail(c, PCF_ERR_HIP, "scratch array of this device not resolved");
  }
  std::lock_guard<std::mutex> lk(d.call);
  pcf_job tuned = *job;
  if (get_cat(c, job->cat_x) && (job->hist == PCF_XX || get_cat(c, job->cat_y)))
    apply_tuning(c, &tuned, 1);
  return run_pairs(c, &tuned, counts, weights, d.ptr, kScratchBytes);
}
