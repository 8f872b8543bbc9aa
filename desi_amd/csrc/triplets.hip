// triplets.hip -- 3iso / 3ani triplet histograms (placeholder until the
// triplet kernels land).
#include "pcf_ctx.h"

namespace pcf {

int run_triplets(pcf_ctx* c, const pcf_job* job, uint64_t* counts, double* weights) {
  (void)job;
  (void)counts;
  (void)weights;
  return fail(c, PCF_ERR_UNSUPPORTED, "triplet histograms not available in this build");
}

}  // namespace pcf
