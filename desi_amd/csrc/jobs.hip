// jobs.hip -- asynchronous job engine behind pcf_submit / pcf_wait
// (placeholder until the serving engine lands).
#include "pcf_ctx.h"

namespace pcf {
struct JobEngine {};
void destroy_engine(JobEngine* e) { delete e; }
void engine_drain(JobEngine* e) { (void)e; }
}  // namespace pcf

extern "C" {
int pcf_set_hbm_budget(pcf_ctx* c, int64_t bytes, int32_t max_inflight) {
  (void)bytes; (void)max_inflight;
  return pcf::fail(c, PCF_ERR_UNSUPPORTED, "serving engine not available in this build");
}
int pcf_submit(pcf_ctx* c, const pcf_job* job, pcf_ticket* ticket) {
  (void)job; (void)ticket;
  return pcf::fail(c, PCF_ERR_UNSUPPORTED, "serving engine not available in this build");
}
int pcf_poll(pcf_ctx* c, pcf_ticket ticket, int32_t* done) {
  (void)ticket; (void)done;
  return pcf::fail(c, PCF_ERR_UNSUPPORTED, "serving engine not available in this build");
}
int pcf_wait(pcf_ctx* c, pcf_ticket ticket, uint64_t* counts, double* weights, pcf_stats* st) {
  (void)ticket; (void)counts; (void)weights; (void)st;
  return pcf::fail(c, PCF_ERR_UNSUPPORTED, "serving engine not available in this build");
}
}
