/* pcf_selftest.h -- self-test hooks of libpcf_hip.so (not part of the
 * reference's interface; used by the library's own tests).
 */
#ifndef PCF_SELFTEST_H
#define PCF_SELFTEST_H

#include <stdint.h>

#include "pcf_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the library's device-wide exclusive prefix sum (the scan behind the
 * grid build, the work lists and the triplet pair lists) over n counts copied
 * from host memory and returns in *total the total it writes: the sum, or -1
 * when the sum is 2^31 or more (callers of the scan test for that flag before
 * using its offsets).  No counterpart in the reference. */
int pcf_selftest_scan(pcf_ctx* ctx, const int32_t* counts, int64_t n, int32_t* total);

#ifdef __cplusplus
}
#endif
#endif /* PCF_SELFTEST_H */
