// Internal entry points behind the C ABI (capi.cpp).
#pragma once
#include <string>
#include <vector>

#include "../device/bvh2l.h"
#include "bvh_build.h"
#include "core.h"

namespace rm {

mat3 normal_transform_of(const mat4& m);
float build_light_distribution(const mat3& M, const uint32_t* idx, const float* verts, uint32_t face_count, float* out);
std::vector<BuildTri> world_triangles(const rm_scene* s);

void scene_create(rm_device* dev, const rm_scene_desc* d, rm_scene* s);
void scene_set_camera(rm_scene* s, const rm_camera_desc* c);
void scene_build_accel(rm_scene* s, int threads);
void scene_build_accel_two_level(rm_scene* s, int threads);

// The two-level form's arrays (device/bvh2l.h): the shapes' trees, then the
// top-level tree at `root`; triangles of the shapes; instance records.
struct TwoLevelBuild {
    std::vector<Bvh8Node> nodes;
    std::vector<BvhTri> tris;
    std::vector<TlasInstance> recs;
    uint32_t root = 0;
    int depth = 0;
    uint64_t stored = 0, total = 0;  // triangles stored once per shape, and over all instances
    double seconds = 0;
};
TwoLevelBuild build_two_level(const std::vector<float>& vertices, const std::vector<uint32_t>& indices,
                              const std::vector<rm_instance_desc>& insts, int threads);
void scene_compile_materials(rm_scene* s, const char* src, size_t len, const char* options, std::string& log);
void scene_build_envmap(rm_scene* s);
// Loads the scene's material library with the radiance-cache shading kernel.
void scene_cache_module(rm_scene* s);
void trace_closest_raw(rm_scene* s, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim,
                       float* bary);
void trace_any_raw(rm_scene* s, const float* rays, uint32_t n, uint32_t* occluded);

void renderer_create(rm_scene* s, uint32_t w, uint32_t h, uint32_t max_depth, uint32_t spl, rm_renderer* r);
// A renderer with an explicit path-queue capacity (batches).
void renderer_init(rm_scene* s, uint32_t w, uint32_t h, uint32_t max_depth, size_t capacity, rm_renderer* r);
// One bounce loop over the slices of `br` (rm_batch_render).
void renderer_render_batch_loop(rm_renderer* r, BatchRun* br, rm_render_stats* stats);
void renderer_reset(rm_renderer* r);
void renderer_set_lens(rm_renderer* r, const rm_lens_desc* lens);
void renderer_render(rm_renderer* r, uint32_t samples, rm_render_stats* stats);
void renderer_read_image(rm_renderer* r, float* rgba, size_t count);
void renderer_read_image_f64(rm_renderer* r, double* rgba, size_t count);
void renderer_resolve(rm_renderer* r, float* rgba, size_t count);
void renderer_set_cache(rm_renderer* r, rm_cache* c, uint32_t mode);

}  // namespace rm
