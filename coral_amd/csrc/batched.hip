// placeholder: filled in below
#include "coral_common.hpp"
#include <string.h>
extern "C" {
int coral_sgemm_batched(int32_t, int32_t, int64_t, const int32_t *, const int32_t *, const int32_t *, float, const float *const *, const int32_t *, const float *const *, const int32_t *, float, float *const *, const int32_t *, coral_stream_t) { return CORAL_ERR_UNSUPPORTED; }
int coral_dgemm_batched(int32_t, int32_t, int64_t, const int32_t *, const int32_t *, const int32_t *, double, const double *const *, const int32_t *, const double *const *, const int32_t *, double, double *const *, const int32_t *, coral_stream_t) { return CORAL_ERR_UNSUPPORTED; }
int coral_batched_set_hint(int32_t, int32_t, int32_t) { return CORAL_OK; }
}
namespace coral {
int batched_menu_size(const char *routine) {
  return (!strcmp(routine, "sgemm_batched") || !strcmp(routine, "dgemm_batched")) ? 2 : 0;
}
}
