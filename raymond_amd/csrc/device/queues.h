// Wavefront path state in HBM, structure-of-arrays.
//
// The reference keeps AoS Ray (80 B), ShadowRay (64 B) and Intersection
// (24 B) records (bridge/Ray.hpp:17-46) in ping-pong buffers of 2 x W x H
// (host/Renderer.swift:766-775).  Here every field is its own array so a
// wave's 64 lanes read and write 256 contiguous bytes per field, and only
// the fields a kernel needs are touched.
#pragma once
#include "common.h"

namespace rm {

constexpr int kMaxDepth = 32;
// depth of a camera path that carries no ray: its lens trace failed, and the
// reference commits no ray for it (generate.metal:143-146)
constexpr uint32_t kVoidDepth = 255;

// path meta word: RNG draw index (16) | ray flags (8) | depth (8)
RM_HD uint32_t pack_meta(uint32_t prng_index, uint32_t flags, uint32_t depth) {
    return (prng_index & 0xffffu) | ((flags & 0xffu) << 16) | ((depth & 0xffu) << 24);
}
RM_HD uint32_t meta_index(uint32_t m) { return m & 0xffffu; }
RM_HD uint32_t meta_flags(uint32_t m) { return (m >> 16) & 0xffu; }
RM_HD uint32_t meta_depth(uint32_t m) { return m >> 24; }

struct PathQueue {
    float *ox, *oy, *oz;
    float *dx, *dy, *dz;
    float *wx, *wy, *wz;
    float* pdf;       // pdf of the BSDF sample that produced the ray (inf for camera rays)
    uint32_t* seed;   // RNG key
    uint32_t* meta;   // pack_meta
    uint32_t* pixel;  // accumulation-buffer pixel index
};

struct HitQueue {
    float* t;         // < 0 on miss
    uint32_t* prim;
    uint32_t* inst;
    float* b0;
    float* b1;
};

struct ShadowQueue {
    float *ox, *oy, *oz;
    float *dx, *dy, *dz;
    float* tmax;
    float *wx, *wy, *wz;
    uint32_t* pixel;
    uint32_t* slot;   // training slot of the emitting path (radiance cache), or null
};

// Per-iteration counters of the regenerating bounce loop.  Iteration i:
// regenerate(i) tops the queue up with new camera paths, then trace, shade
// (appends survivors for iteration i+1) and shadow trace run on it.
struct IterCounters {
    uint32_t queue;          // paths traced in iteration i (survivors + new)
    uint32_t appended;       // survivors appended by shade(i-1)
    uint32_t generated;      // camera paths generated before iteration i
    uint32_t shadows;        // shadow rays emitted by shade(i)
    uint32_t fetch_closest;  // persistent-scheduler counters
    uint32_t fetch_any;
    uint32_t voids;          // new camera paths without a ray (counted in queue, not traced)
    uint32_t pad1;
};

constexpr int kMaxIters = 2048;
// Material-sort bins: surface materials 0..kSortBins-3 get their own bin
// (higher ids share bin kSortBins-2) and misses go to bin kSortBins-1.
constexpr int kSortBins = 256;

// Device counters of one render stream (zeroed by a kernel at stream start).
struct Counters {
    IterCounters it[kMaxIters];
    // material-sort bins, double-buffered by iteration parity: scatter(i)
    // clears the set that keys(i+1) fills
    uint32_t sort_bins[2][kSortBins];
    uint32_t sort_cursor[2][kSortBins];
    // radiance cache: lookups and training vertices recorded in this stream
    uint32_t cache_queries, train_records;
    // first iteration of the chunk a captured bounce-loop graph is running
    uint32_t iter_base, pad_iter;
    // traversal statistics (only written by the statistics variants)
    unsigned long long nodes_closest, tris_closest, nodes_any, tris_any;
};

// Iteration a kernel works on: `iter` itself when launched directly, or the
// offset `iter` within a captured chunk whose first iteration the device
// counter holds.
RM_HD uint32_t iteration_of(const uint32_t* iter_dev, uint32_t iter) { return iter_dev ? *iter_dev + iter : iter; }

// Render-time settings (bridge/Uniforms.hpp:29-47 plus this library's knobs).
struct FrameUniforms {
    uint32_t width, height;
    uint32_t samples;          // samples per pixel in this stream
    uint32_t random_seed;      // seed of sample 0 (frame index + 1)
    uint32_t sampling_mode;    // 0 bsdf, 1 nee, 2 mis
    uint32_t rr_depth;         // roulette starts at this depth (0 = reference behaviour)
    uint32_t max_depth;
    uint32_t image_offset;     // first pixel of this job in the accumulation buffer
};

}  // namespace rm
