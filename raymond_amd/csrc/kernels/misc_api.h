// Launchers for the small ahead-of-time kernels (misc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "../device/queues.h"

namespace rm {

hipError_t launch_clear_counters(Counters* c, hipStream_t s);
hipError_t launch_advance_iter(Counters* c, uint32_t n, hipStream_t s);
hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_envmap_reduce(float* level, int res, hipStream_t s);
hipError_t launch_envmap_normalize(const float* total, float* pdfs, uint32_t n, hipStream_t s);
hipError_t launch_resolve(const float* img, float* out, uint32_t n, float exposure, float divisor, int tonemap,
                          hipStream_t s);
hipError_t launch_accum_f64(float* img, double* acc, uint32_t n4, hipStream_t s);

}  // namespace rm
