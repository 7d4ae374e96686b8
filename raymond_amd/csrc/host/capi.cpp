// C-ABI entry points (include/rm.h): argument checks, exception-to-status
// translation and the thread-local error message.
#include <cstring>
#include <string>

#include "../kernels/misc_api.h"
#include "../kernels/noalloc_api.h"
#include "../kernels/trace_api.h"
#include "core.h"
#include "batch_api.h"
#include "rcache_host.h"
#include "scene_api.h"

namespace rm {
static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
}  // namespace rm

using rm::HipError;

#define RM_GUARD_BEGIN try {
#define RM_GUARD_END                             \
    }                                            \
    catch (const HipError& e) {                  \
        rm::set_error(e.msg);                    \
        return e.code;                           \
    }                                            \
    catch (const std::bad_alloc&) {              \
        rm::set_error("host allocation failed"); \
        return RM_ERROR_OOM;                     \
    }                                            \
    catch (const std::exception& e) {            \
        rm::set_error(e.what());                 \
        return RM_ERROR_INVALID;                 \
    }                                            \
    return RM_OK;

#define RM_NULL_CHECK(p)                                    \
    do {                                                    \
        if (!(p)) {                                         \
            rm::set_error(std::string("null argument: ") + #p); \
            return RM_ERROR_INVALID;                        \
        }                                                   \
    } while (0)

extern "C" {

const char* rm_last_error(void) { return rm::g_error.c_str(); }
const char* rm_version(void) { return "raymond-amd 0.1 (gfx950)"; }

int rm_device_create(int ordinal, rm_device** out) {
    RM_NULL_CHECK(out);
    *out = nullptr;
    RM_GUARD_BEGIN
    int count = 0;
    RM_HIP(hipGetDeviceCount(&count));
    RM_REQUIRE(ordinal >= 0 && ordinal < count, "device ordinal out of range (" + std::to_string(count) + " devices)");
    RM_HIP(hipSetDevice(ordinal));
    rm_device* d = new rm_device();
    d->ordinal = ordinal;
    try {
        RM_HIP(hipGetDeviceProperties(&d->props, ordinal));
        RM_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        d->trace_blocks_per_cu = rm::trace_blocks_per_cu(false);
        d->any_blocks_per_cu = rm::trace_blocks_per_cu(true);
        if (d->trace_blocks_per_cu < 1) d->trace_blocks_per_cu = 1;
        if (d->any_blocks_per_cu < 1) d->any_blocks_per_cu = 1;
        rm::noalloc_init(d);
    } catch (...) {
        delete d;
        throw;
    }
    *out = d;
    RM_GUARD_END
}

void rm_device_destroy(rm_device* dev) {
    if (!dev) return;
    if (dev->stream) (void)hipStreamDestroy(dev->stream);
    delete dev;
}

int rm_device_get_info(rm_device* dev, rm_device_info* out) {
    RM_NULL_CHECK(dev);
    RM_NULL_CHECK(out);
    std::memset(out, 0, sizeof(*out));
    out->compute_units = dev->props.multiProcessorCount;
    out->wavefront_size = dev->props.warpSize;
    out->global_memory_bytes = dev->props.totalGlobalMem;
    std::strncpy(out->name, dev->props.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, dev->props.gcnArchName, sizeof(out->arch) - 1);
    return RM_OK;
}

int rm_scene_create(rm_device* dev, const rm_scene_desc* desc, rm_scene** out) {
    RM_NULL_CHECK(dev);
    RM_NULL_CHECK(desc);
    RM_NULL_CHECK(out);
    *out = nullptr;
    RM_GUARD_BEGIN
    rm_scene* s = new rm_scene();
    try {
        rm::scene_create(dev, desc, s);
    } catch (...) {
        delete s;
        throw;
    }
    *out = s;
    RM_GUARD_END
}

void rm_scene_destroy(rm_scene* scene) { delete scene; }

int rm_scene_build_accel(rm_scene* scene, int host_threads) {
    RM_NULL_CHECK(scene);
    RM_GUARD_BEGIN
    rm::scene_build_accel(scene, host_threads);
    RM_GUARD_END
}

int rm_scene_build_accel_two_level(rm_scene* scene, int host_threads) {
    RM_NULL_CHECK(scene);
    RM_GUARD_BEGIN
    rm::scene_build_accel_two_level(scene, host_threads);
    RM_GUARD_END
}

int rm_scene_compile_materials(rm_scene* scene, const char* source, size_t length, const char* options, char* log,
                               size_t log_capacity) {
    RM_NULL_CHECK(scene);
    RM_NULL_CHECK(source);
    std::string text_log;
    int status = RM_OK;
    try {
        rm::scene_compile_materials(scene, source, length, options, text_log);
    } catch (const HipError& e) {
        rm::set_error(e.msg);
        status = e.code;
    } catch (const std::exception& e) {
        rm::set_error(e.what());
        status = RM_ERROR_INVALID;
    }
    if (log && log_capacity) {
        std::strncpy(log, text_log.c_str(), log_capacity - 1);
        log[log_capacity - 1] = '\0';
    }
    return status;
}

int rm_scene_build_envmap(rm_scene* scene) {
    RM_NULL_CHECK(scene);
    RM_GUARD_BEGIN
    rm::scene_build_envmap(scene);
    RM_GUARD_END
}

int rm_scene_set_camera(rm_scene* scene, const rm_camera_desc* camera) {
    RM_NULL_CHECK(scene);
    RM_NULL_CHECK(camera);
    RM_GUARD_BEGIN
    rm::scene_set_camera(scene, camera);
    RM_GUARD_END
}

int rm_scene_get_stats(rm_scene* scene, rm_scene_stats* out) {
    RM_NULL_CHECK(scene);
    RM_NULL_CHECK(out);
    *out = scene->stats;
    const rm::DevBuf* bufs[] = {&scene->d_vertices, &scene->d_normals, &scene->d_texcoords, &scene->d_indices,
                                &scene->d_materials, &scene->d_instances, &scene->d_area, &scene->d_point,
                                &scene->d_sun, &scene->d_spot, &scene->d_shape, &scene->d_light_faces,
                                &scene->d_nodes, &scene->d_tris, &scene->d_tlas, &scene->d_env_pdfs,
                                &scene->d_env_mip, &scene->d_spill, &scene->d_tex_table};
    uint64_t bytes = 0;
    for (const rm::DevBuf* b : bufs) bytes += b->bytes;
    for (const auto& t : scene->d_tex) bytes += t ? t->bytes : 0;
    out->device_bytes = bytes;
    return RM_OK;
}

int rm_scene_read_envmap(rm_scene* scene, float* pdfs, size_t pdf_count, float* mipmap, size_t mip_count) {
    RM_NULL_CHECK(scene);
    RM_GUARD_BEGIN
    RM_REQUIRE(scene->envmap_built, "envmap not built");
    hipStream_t st = scene->dev->stream;
    if (pdfs) {
        RM_REQUIRE(pdf_count * 4 >= scene->d_env_pdfs.bytes, "pdf buffer too small");
        RM_HIP(hipMemcpyAsync(pdfs, scene->d_env_pdfs.ptr, scene->d_env_pdfs.bytes, hipMemcpyDeviceToHost, st));
    }
    if (mipmap) {
        RM_REQUIRE(mip_count * 4 >= scene->d_env_mip.bytes, "mipmap buffer too small");
        RM_HIP(hipMemcpyAsync(mipmap, scene->d_env_mip.ptr, scene->d_env_mip.bytes, hipMemcpyDeviceToHost, st));
    }
    RM_HIP(hipStreamSynchronize(st));
    RM_GUARD_END
}

int rm_scene_read_light_tables(rm_scene* scene, float* light_faces, size_t face_count, float* emissive_areas,
                               size_t light_count) {
    RM_NULL_CHECK(scene);
    RM_GUARD_BEGIN
    if (light_faces) {
        RM_REQUIRE(face_count >= scene->light_faces.size(), "face buffer too small");
        std::memcpy(light_faces, scene->light_faces.data(), scene->light_faces.size() * 4);
    }
    if (emissive_areas) {
        RM_REQUIRE(light_count >= scene->shape.size(), "light buffer too small");
        for (size_t i = 0; i < scene->shape.size(); i++) emissive_areas[i] = scene->shape[i].emissive_area;
    }
    RM_GUARD_END
}

int rm_trace_closest(rm_scene* scene, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim,
                     float* bary) {
    RM_NULL_CHECK(scene);
    if (n) RM_NULL_CHECK(rays);
    RM_GUARD_BEGIN
    rm::trace_closest_raw(scene, rays, n, t, inst, prim, bary);
    RM_GUARD_END
}

int rm_trace_closest_noalloc(rm_scene* scene, const float* rays, uint32_t n, float* t, uint32_t* inst,
                             uint32_t* prim, float* bary, void* stream) {
    RM_NULL_CHECK(scene);
    if (n) {
        RM_NULL_CHECK(rays);
        RM_NULL_CHECK(t);
        RM_NULL_CHECK(inst);
        RM_NULL_CHECK(prim);
        RM_NULL_CHECK(bary);
    }
    RM_GUARD_BEGIN
    rm::trace_closest_noalloc(scene, rays, n, t, inst, prim, bary, static_cast<hipStream_t>(stream));
    RM_GUARD_END
}

int rm_trace_any(rm_scene* scene, const float* rays, uint32_t n, uint32_t* occluded) {
    RM_NULL_CHECK(scene);
    if (n) {
        RM_NULL_CHECK(rays);
        RM_NULL_CHECK(occluded);
    }
    RM_GUARD_BEGIN
    rm::trace_any_raw(scene, rays, n, occluded);
    RM_GUARD_END
}

int rm_renderer_create(rm_scene* scene, uint32_t width, uint32_t height, uint32_t max_depth,
                       uint32_t samples_per_launch, rm_renderer** out) {
    RM_NULL_CHECK(scene);
    RM_NULL_CHECK(out);
    *out = nullptr;
    RM_GUARD_BEGIN
    rm_renderer* r = new rm_renderer();
    try {
        rm::renderer_create(scene, width, height, max_depth, samples_per_launch, r);
    } catch (...) {
        delete r;
        throw;
    }
    *out = r;
    RM_GUARD_END
}

void rm_renderer_destroy(rm_renderer* r) {
    if (!r) return;
    for (hipEvent_t e : r->events) (void)hipEventDestroy(e);
    delete r;
}

int rm_renderer_set_uniforms(rm_renderer* r, const rm_uniforms* u) {
    RM_NULL_CHECK(r);
    RM_NULL_CHECK(u);
    if (u->sampling_mode > 2 || u->tonemapping > 2) {
        rm::set_error("invalid sampling mode or tonemapping");
        return RM_ERROR_INVALID;
    }
    r->u = *u;
    return RM_OK;
}

int rm_renderer_set_lens(rm_renderer* r, const rm_lens_desc* lens) {
    RM_NULL_CHECK(r);
    RM_GUARD_BEGIN
    rm::renderer_set_lens(r, lens);
    RM_GUARD_END
}

int rm_renderer_reset(rm_renderer* r) {
    RM_NULL_CHECK(r);
    RM_GUARD_BEGIN
    rm::renderer_reset(r);
    RM_GUARD_END
}

int rm_render(rm_renderer* r, uint32_t samples, rm_render_stats* stats) {
    RM_NULL_CHECK(r);
    RM_GUARD_BEGIN
    rm::renderer_render(r, samples, stats);
    RM_GUARD_END
}

int rm_read_image(rm_renderer* r, float* rgba, size_t count) {
    RM_NULL_CHECK(r);
    RM_NULL_CHECK(rgba);
    RM_GUARD_BEGIN
    rm::renderer_read_image(r, rgba, count);
    RM_GUARD_END
}

int rm_read_image_f64(rm_renderer* r, double* rgba, size_t count) {
    RM_NULL_CHECK(r);
    RM_NULL_CHECK(rgba);
    RM_GUARD_BEGIN
    rm::renderer_read_image_f64(r, rgba, count);
    RM_GUARD_END
}

int rm_resolve(rm_renderer* r, float* rgba, size_t count) {
    RM_NULL_CHECK(r);
    RM_NULL_CHECK(rgba);
    RM_GUARD_BEGIN
    rm::renderer_resolve(r, rgba, count);
    RM_GUARD_END
}

uint32_t rm_renderer_frame_index(rm_renderer* r) { return r ? r->frame_index : 0; }

int rm_device_copy_probe(rm_device* dev, size_t bytes, double* ms) {
    RM_NULL_CHECK(dev);
    if (bytes == 0 || bytes % 16) {
        rm::set_error("copy size must be a positive multiple of 16 bytes");
        return RM_ERROR_INVALID;
    }
    RM_GUARD_BEGIN
    hipStream_t st = dev->stream;
    rm::DevBuf src, dst;
    src.alloc(bytes);
    dst.alloc(bytes);
    RM_HIP(hipMemsetAsync(src.ptr, 0x3c, bytes, st));
    hipEvent_t e0, e1;
    RM_HIP(hipEventCreate(&e0));
    RM_HIP(hipEventCreate(&e1));
    RM_HIP(hipEventRecord(e0, st));
    const hipError_t err = rm::launch_copy_probe(src.ptr, dst.ptr, bytes, st);
    RM_HIP(hipEventRecord(e1, st));
    RM_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    RM_HIP(err);
    if (ms) *ms = t;
    RM_GUARD_END
}

int rm_cache_create(rm_device* dev, const rm_cache_desc* desc, rm_cache** out) {
    RM_NULL_CHECK(dev);
    RM_NULL_CHECK(out);
    *out = nullptr;
    RM_GUARD_BEGIN
    rm_cache* c = new rm_cache();
    try {
        rm::cache_create(dev, desc, c);
    } catch (...) {
        delete c;
        throw;
    }
    *out = c;
    RM_GUARD_END
}

void rm_cache_destroy(rm_cache* c) { delete c; }

int rm_cache_get_stats(rm_cache* c, rm_cache_stats* out) {
    RM_NULL_CHECK(c);
    RM_NULL_CHECK(out);
    RM_GUARD_BEGIN
    // training inside rm_render does not wait for its loss; the last step's
    // sum is still on the device (rm_render has returned, so it is final)
    if (c->steps > 0 && c->loss.ptr) {
        float l = 0.0f;  // a device-to-host copy finds the buffer's device by its address
        RM_HIP(hipMemcpy(&l, c->loss.ptr, sizeof(float), hipMemcpyDeviceToHost));
        c->last_loss = l;
    }
    out->steps = c->steps;
    out->samples = c->samples;
    out->queries = c->queries;
    out->train_records = c->records;
    out->last_loss = c->last_loss;
    out->ms_infer = c->ms_infer;
    out->ms_train = c->ms_train;
    RM_GUARD_END
}

uint64_t rm_cache_table_floats(rm_cache* c) { return c ? (uint64_t)c->table_floats : 0; }

int rm_cache_read_params(rm_cache* c, float* mlp, size_t mlp_count, float* table, size_t table_count) {
    RM_NULL_CHECK(c);
    RM_GUARD_BEGIN
    rm::cache_read_params(c, mlp, mlp_count, table, table_count);
    RM_GUARD_END
}

int rm_cache_write_params(rm_cache* c, const float* mlp, size_t mlp_count, const float* table, size_t table_count) {
    RM_NULL_CHECK(c);
    RM_GUARD_BEGIN
    rm::cache_write_params(c, mlp, mlp_count, table, table_count);
    RM_GUARD_END
}

int rm_cache_infer(rm_cache* c, const float* features, uint32_t n, float* out) {
    RM_NULL_CHECK(c);
    if (n) {
        RM_NULL_CHECK(features);
        RM_NULL_CHECK(out);
    }
    RM_GUARD_BEGIN
    rm::cache_infer_raw(c, features, n, out);
    RM_GUARD_END
}

int rm_cache_gradients(rm_cache* c, const float* features, const float* targets, uint32_t n, float* grad_mlp,
                       float* grad_table, float* loss) {
    RM_NULL_CHECK(c);
    RM_NULL_CHECK(features);
    RM_NULL_CHECK(targets);
    RM_GUARD_BEGIN
    rm::cache_gradients_raw(c, features, targets, n, grad_mlp, grad_table, loss);
    RM_GUARD_END
}

int rm_cache_train(rm_cache* c, const float* features, const float* targets, uint32_t n, uint32_t steps,
                   float* last_loss) {
    RM_NULL_CHECK(c);
    RM_NULL_CHECK(features);
    RM_NULL_CHECK(targets);
    RM_GUARD_BEGIN
    rm::cache_train_raw(c, features, targets, n, steps, last_loss);
    RM_GUARD_END
}

int rm_batch_create(rm_device* dev, const rm_job_desc* jobs, uint32_t n, rm_batch** out) {
    RM_NULL_CHECK(dev);
    RM_NULL_CHECK(jobs);
    RM_NULL_CHECK(out);
    *out = nullptr;
    RM_GUARD_BEGIN
    rm_batch* b = new rm_batch();
    try {
        rm::batch_create(dev, jobs, n, b);
    } catch (...) {
        delete b;
        throw;
    }
    *out = b;
    RM_GUARD_END
}

void rm_batch_destroy(rm_batch* b) {
    if (!b) return;
    if (b->renderer)
        for (hipEvent_t e : b->renderer->events) (void)hipEventDestroy(e);
    delete b;
}

int rm_batch_render(rm_batch* b, const rm_batch_config* config, rm_batch_stats* stats) {
    RM_NULL_CHECK(b);
    RM_GUARD_BEGIN
    rm::batch_render(b, config, stats);
    RM_GUARD_END
}

int rm_batch_job_times(rm_batch* b, double* ms, uint32_t n) {
    RM_NULL_CHECK(b);
    RM_NULL_CHECK(ms);
    for (uint32_t j = 0; j < n && j < b->job_ms.size(); j++) ms[j] = b->job_ms[j];
    return RM_OK;
}

int rm_batch_read_image(rm_batch* b, uint32_t job, float* rgba, size_t count) {
    RM_NULL_CHECK(b);
    RM_NULL_CHECK(rgba);
    RM_GUARD_BEGIN
    rm::batch_read_image(b, job, rgba, count);
    RM_GUARD_END
}

int rm_batch_autotune(rm_batch* b, rm_batch_config* best, rm_batch_tune_entry* entries, uint32_t capacity,
                      uint32_t* count, int* cache_hit) {
    RM_NULL_CHECK(b);
    RM_NULL_CHECK(best);
    RM_GUARD_BEGIN
    rm::batch_autotune(b, best, entries, capacity, count, cache_hit);
    RM_GUARD_END
}

int rm_renderer_set_cache(rm_renderer* r, rm_cache* c, uint32_t mode) {
    RM_NULL_CHECK(r);
    if (c && (mode < 1 || mode > 2)) {
        rm::set_error("cache mode must be 1 (query and train) or 2 (query only)");
        return RM_ERROR_INVALID;
    }
    RM_GUARD_BEGIN
    rm::renderer_set_cache(r, c, c ? mode : 0u);
    RM_GUARD_END
}

}  // extern "C"
