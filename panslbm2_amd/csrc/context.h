// panslbm2_amd/csrc/context.h -- what a translation unit other than capi.hip needs of a context:
// its device, its stream, and the thread's error message behind lbm_last_error.
#pragma once
#include <hip/hip_runtime.h>

struct lbm_context;

namespace lbm {
int context_device(const lbm_context *C);
hipStream_t context_stream(const lbm_context *C);
// records the message for lbm_last_error and returns code
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
}  // namespace lbm
