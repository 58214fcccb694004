// ug_fast.hip -- split-BF16 prefilter (placeholder until the MFMA path lands).
#include "ug_internal.h"
namespace ug {
int32_t fast_setup(Pack* h) { (void)h; return fail(UG_EINVAL, "split-BF16 form not built"); }
int32_t launch_fast(Pack* h, uint64_t, int64_t, uint8_t*) { (void)h; return fail(UG_EINVAL, "split-BF16 form not built"); }
int32_t launch_fast_eval(Pack* h, const int64_t*, int64_t, double*) { (void)h; return fail(UG_EINVAL, "split-BF16 form not built"); }
void fast_release(Pack* h) { (void)h; }
}
