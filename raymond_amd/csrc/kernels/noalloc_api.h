// Allocation-free closest-hit batches (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct rm_device;
struct rm_scene;

namespace rm {

// Rays per rm_trace_closest_noalloc call; the library's fixed device array
// holds 10 words per ray plus a 256-byte head (10.5 MB per GPU).
constexpr uint32_t kNoAllocMaxRays = 1u << 18;

// Takes the device address of the fixed array on the current device (called
// once per device at set-up, with that device current).
void noalloc_init(rm_device* dev);
void trace_closest_noalloc(rm_scene* sc, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim,
                           float* bary, hipStream_t stream);

}  // namespace rm
