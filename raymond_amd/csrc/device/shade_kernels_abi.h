// Argument blocks of the run-time compiled kernels (shade_kernels.h), shared
// with the host code that launches them.
#pragma once
#include "common.h"
#include "queues.h"
#include "rcache.h"
#include "batch.h"
#include "scene.h"

namespace rm {

struct ShadeArgs {
    SceneView ctx;
    PathQueue cur;
    HitQueue hits;
    PathQueue next;
    ShadowQueue shadows;
    Counters* counters;
    const uint32_t* order;     // queue entries grouped by material, or null
    CacheView cache;           // radiance cache (enabled = 0: off)
    BatchView batch;           // batched jobs (jobs null: one scene)
    float* image;          // RGBA accumulation
    FrameUniforms u;
    uint32_t iter;             // iteration of the regenerating bounce loop
    const uint32_t* iter_dev;  // captured bounce loops: iteration = *iter_dev + iter
};

struct EnvmapArgs {
    SceneView ctx;
    float* mipmap_last;    // finest mip level (quad layout)
    float* pdfs;
    int resolution;
};

}  // namespace rm
