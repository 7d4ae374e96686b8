/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar restatement of the reference's wavefront path tracer, used by
 * tests/ and by bench.py's cpu_baseline leg as the checker.  The product
 * (raymond_amd/) never links, loads or calls anything here.
 *
 * What it restates on its own (each function cites the reference):
 *   generateRays            device/kernels/shading/generate.metal:72-121
 *   handleIntersections     device/kernels/shading/intersection.metal:14-211
 *   handleShadowRays        device/kernels/shading/shadow.metal:6-29
 *   bounce loop             host/Renderer.swift:480-696
 *   PerInstanceData         host/scene/EntityBuilder.swift:97-141
 *   buildLightDistribution  host/lights/distribution.m:7-39
 *   envmap build/reduce/normalize  device/kernels/envmap/{build,reduce,normalize}.metal
 *   closest / any hit       exhaustive loop and an independent binary BVH
 *                           (median split) -- not the product's BVH8
 *   radiance cache          rcache_ref.cpp: encoding, MLP, loss and gradients
 *                           in double (no reference counterpart)
 * What it shares with the product (pinned separately by property tests in
 * tests/): the device function library under raymond_amd/csrc/device
 * (RNG, warps, BSDF lobes, lights, node library, elementary functions) and
 * the watertight triangle / conservative box tests of isect.h, which define
 * the library's hit semantics (the reference's are Metal's and unpublished).
 */
#ifndef RAYMOND_AMD_ORACLE_H
#define RAYMOND_AMD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/rm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* material_so: shared object built from the generated material source (see
 * oracle/binding.py); must export or_material_surface / or_material_light. */
or_scene* or_scene_create(const rm_scene_desc* desc, const char* material_so, char* err, size_t err_cap);
void or_scene_destroy(or_scene* s);

int or_build_envmap(or_scene* s, int threads);
int or_set_envmap(or_scene* s, const float* pdfs, const float* mipmap);
int or_read_envmap(or_scene* s, float* pdfs, float* mipmap);
int or_read_light_tables(or_scene* s, float* light_faces, float* emissive_areas);

/* rays: n x 8 (ox, oy, oz, tmin, dx, dy, dz, tmax).  brute_force != 0 tests
 * every triangle; otherwise the oracle's own BVH2 is used. */
int or_trace_closest(or_scene* s, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim, float* bary,
                     int brute_force, int threads);
int or_trace_any(or_scene* s, const float* rays, uint32_t n, uint32_t* occluded, int brute_force, int threads);
/* Two-level hits (the product's rm_scene_build_accel_two_level): every
 * instance, the ray mapped into its object space, every triangle of its
 * shape.  on = 0 returns to the flat world-space triangles.  -1 for a
 * singular instance transform. */
int or_scene_set_two_level(or_scene* s, int on);

typedef struct or_render_args {
    uint32_t width, height;
    uint32_t max_depth;
    uint32_t sampling_mode;
    uint32_t rr_depth;
    uint32_t frame_begin;     /* first frame index (seed = frame + 1) */
    uint32_t frames;
    const uint32_t* pixels;   /* optional subset: pixel indices y*W+x */
    uint32_t num_pixels;
    int threads;
    int brute_force;
    const rm_lens_desc* lens;  /* optional: camera rays through this lens (generate.metal:122-163) */
} or_render_args;

typedef struct or_render_result {
    uint64_t rays[33];
    uint64_t shadow_rays[32];
    uint64_t total_rays;
    uint64_t paths;
    double seconds;
    int threads_used;
} or_render_result;

/* Accumulates into image (width*height*4 floats, RGBA sums as the GPU). */
int or_render(or_scene* s, const or_render_args* args, float* image, or_render_result* result);
/* The same paths, each sample's contributions summed in FP32 and the samples
 * accumulated in FP64 (width*height*4 doubles), as the GPU's FP64 image. */
int or_render_f64(or_scene* s, const or_render_args* args, double* image, or_render_result* result);

#ifdef __cplusplus
}
#endif
#endif
