// Batched rendering: many independent jobs (scene, camera, resolution,
// samples, depth) in one regenerating bounce loop.
//
// Not a reference component: the reference renders one scene per Renderer.
// The jobs' scenes are merged into one set of device arrays.  Each job keeps
// its own BVH8 subtree, entered at its root node, its own lights, camera,
// depth and a contiguous range of the batch image, so the job a path belongs
// to follows from its pixel index.
#pragma once
#include "common.h"
#include "scene.h"

namespace rm {

struct JobInfo {
    uint32_t image_offset;  // first pixel of the job in the batch image (ascending over jobs)
    uint32_t width, height;
    uint32_t root;          // BVH8 root node of the job's geometry
    uint32_t max_depth;
    uint32_t seed;          // random seed of the job's sample 0
    uint32_t spp;
    uint32_t pad;
    Camera camera;
};

// A slice is one sample of one job: width x height consecutive paths of the
// bounce loop's sample space.
struct BatchSlice {
    uint32_t first;   // first path index of the slice within its bounce loop
    uint32_t job;
    uint32_t sample;  // sample index within the job
    uint32_t pad;
};

struct BatchView {
    const JobInfo* jobs;
    const LightsView* lights;  // per job
    uint32_t num_jobs;
    uint32_t* done;            // per job: paths that have ended
};

// Job of a batch-image pixel: the last job whose range starts at or before it.
RM_HD uint32_t job_of_pixel(const JobInfo* jobs, uint32_t n, uint32_t pixel) {
    uint32_t lo = 0, hi = n;  // jobs[lo].image_offset <= pixel < jobs[hi].image_offset
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].image_offset <= pixel) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Slice holding path g of a bounce loop over slices [0, n).
RM_HD uint32_t slice_of_path(const BatchSlice* s, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid].first <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

}  // namespace rm
