// Private HIP-side definitions shared by the .hip translation units.
#ifndef RFSYN_RF_HIP_H
#define RFSYN_RF_HIP_H

#include <hip/hip_runtime.h>

struct rf_stream { hipStream_t s; };

static inline hipStream_t rf_raw_stream(rf_stream *s) { return s ? s->s : nullptr; }

#endif
