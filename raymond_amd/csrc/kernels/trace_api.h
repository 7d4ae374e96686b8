// Launch interface of the ahead-of-time traversal kernels (trace.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "../device/bvh.h"
#include "../device/queues.h"
#include "../device/scene.h"
#include "../device/batch.h"
#include "../device/bvh2l.h"
#include "../device/camera.h"

namespace rm {

constexpr int kTraceBlock = 64;      // one wave64 per workgroup
constexpr int kMaxTraceStack = 64;   // LDS rows + spill rows; BVH depth is bounded below this
// Stack rows per lane in LDS.  The stack holds at most one entry per tree
// level, so 12 rows cover BVHs up to depth 12 without spilling; 6 KB of LDS
// per wave leaves the VGPRs, not LDS, limiting occupancy (16 rows measured
// 2462 against 3079-3097 Mrays/s on the 5.2M-triangle sphere field).  RM_LDS_STACK
// changes the row count for a test build of the library (1 row: every
// traversal below the root's children runs through the spill rows).
#ifndef RM_LDS_STACK
#define RM_LDS_STACK 12
#endif
static_assert(RM_LDS_STACK >= 1 && RM_LDS_STACK < 64, "LDS stack rows out of range");
constexpr int kLdsStack = RM_LDS_STACK;
struct ClosestArgs {
    const Bvh8Node* nodes;
    const BvhTri* tris;
    uint32_t count;             // used when count_ptr is null
    const uint32_t* count_ptr;  // device queue length
    uint32_t* fetch;            // persistent work counter (zero before launch)
    PathQueue paths;
    HitQueue hits;
    const float* tmin_arr;      // optional per-ray range (raw batches); else eps / inf
    const float* tmax_arr;
    uint2* spill;               // (kMaxTraceStack - kLdsStack) x grid threads
    Counters* counters;
    Camera camera;              // clip range of depth-0 paths
    BatchView batch;            // batched jobs (jobs null: one scene, root 0)
    // captured bounce loops: count and fetch come from counters->it[iteration_of(iter_dev, iter)]
    const uint32_t* iter_dev;
    uint32_t iter;
    // two-level scenes: instance records of the top-level tree (null: flat)
    const TlasInstance* tlas;
    uint32_t root2l;
};

struct RegenArgs {
    Lens lens;                  // num_surfaces 0: pinhole camera
    Counters* counters;
    uint32_t iter;
    const uint32_t* iter_dev;   // captured bounce loops (iteration_of)
    uint32_t capacity;          // queue capacity
    uint32_t total;             // paths in the stream (W H samples)
    PathQueue paths;            // queue of iteration `iter`
    Camera camera;
    uint32_t width, height, random_seed, image_offset;
    // radiance cache: training slot per new path (null when the cache is off)
    uint32_t* slots;
    float* slot_radiance;
    uint32_t train_shift;
    // batched jobs: path g belongs to the slice containing it (null: one job)
    const BatchSlice* slices;
    uint32_t num_slices;
    const JobInfo* jobs;
};

struct AnyArgs {
    const Bvh8Node* nodes;
    const BvhTri* tris;
    uint32_t count;
    const uint32_t* count_ptr;
    uint32_t* fetch;
    ShadowQueue shadows;
    const float* tmin_arr;      // optional; else eps
    float* image;               // accumulate unoccluded weights when non-null
    uint32_t* occluded;         // else write 0/1 per ray
    float* slot_radiance;       // radiance cache: unoccluded weights of training paths
    uint2* spill;
    Counters* counters;
    BatchView batch;
    const uint32_t* iter_dev;   // captured bounce loops, as in ClosestArgs
    uint32_t iter;
    const TlasInstance* tlas;   // two-level scenes, as in ClosestArgs
    uint32_t root2l;
};

hipError_t launch_trace_closest(const ClosestArgs& a, int blocks, bool stats, hipStream_t stream);
hipError_t launch_trace_any(const AnyArgs& a, int blocks, bool stats, hipStream_t stream);
hipError_t launch_regenerate(const RegenArgs& a, int blocks, hipStream_t stream);
int trace_blocks_per_cu(bool any);

}  // namespace rm
