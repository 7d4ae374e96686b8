// Internal definitions of the C-ABI objects (include/rm.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <memory>
#include <vector>

#include "../../../include/rm.h"
#include "../device/bvh.h"
#include "../device/queues.h"
#include "../device/rcache.h"
#include "../device/batch.h"
#include "../device/scene.h"
#include "../device/camera.h"

namespace rm {

void set_error(const std::string& msg);

struct HipError {
    int code;
    std::string msg;
};

#define RM_HIP(expr)                                                                               \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            throw ::rm::HipError{RM_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

#define RM_REQUIRE(cond, msg)                                        \
    do {                                                             \
        if (!(cond)) throw ::rm::HipError{RM_ERROR_INVALID, (msg)}; \
    } while (0)

// Device allocation owned by a C-ABI object.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        RM_HIP(hipMalloc(&ptr, n));
        bytes = n;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
    void upload(const void* src, size_t n, hipStream_t s) {
        alloc(n);
        if (n) RM_HIP(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s));
    }
};

struct JitModule {
    hipModule_t module = nullptr;
    hipFunction_t shade = nullptr;
    hipFunction_t shade_batch = nullptr;  // present in modules compiled with -DRM_BATCH_KERNEL
    hipFunction_t shade_cache = nullptr;  // present in modules compiled with -DRM_CACHE_KERNEL
    hipFunction_t envmap_build = nullptr;
    ~JitModule() {
        if (module) (void)hipModuleUnload(module);
    }
};

// One bounce loop of a batch (rm_batch_render): the slices it covers and the
// per-job completion bookkeeping the host updates at every poll.
struct BatchRun {
    const BatchSlice* slices = nullptr;  // device, firsts relative to this loop
    uint32_t num_slices = 0;
    uint32_t total = 0;                  // paths of this loop
    BatchView view{};
    const std::vector<uint64_t>* job_paths = nullptr;  // paths per job (whole batch)
    std::vector<double>* job_ms = nullptr;             // completion time per job, < 0 while running
    double t0_ms = 0.0;                                // steady-clock origin of the batch
};

// Compiles material source + the shading kernels; caches code objects on disk.
bool jit_compile(const std::string& body, const std::string& options, std::vector<char>& code, std::string& log,
                 bool& cache_hit);
// Full translation unit handed to hiprtc (exposed for inspection/tests).
std::string jit_translation_unit(const std::string& body);

}  // namespace rm

struct rm_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    int trace_blocks_per_cu = 1;
    int any_blocks_per_cu = 1;
    // this device's address of the allocation-free entry point's fixed array
    uint32_t* noalloc_scratch = nullptr;
};

struct rm_scene {
    rm_device* dev = nullptr;
    // host copies
    std::vector<float> vertices, normals, texcoords;
    std::vector<uint32_t> indices;
    std::vector<uint16_t> materials;
    std::vector<rm_instance_desc> inst_desc;
    std::vector<rm::PerInstance> instances;
    std::vector<rm::AreaLight> area;
    std::vector<rm::PointLight> point;
    std::vector<rm::SunLight> sun;
    std::vector<rm::SpotLight> spot;
    std::vector<rm::ShapeLight> shape;
    std::vector<float> light_faces;
    uint32_t world_shader = 0;
    int envmap_res = 2048;
    // device copies
    rm::DevBuf d_vertices, d_normals, d_texcoords, d_indices, d_materials, d_instances;
    rm::DevBuf d_area, d_point, d_sun, d_spot, d_shape, d_light_faces;
    rm::DevBuf d_nodes, d_tris;
    // two-level form: instance records of the top-level tree and its root
    rm::DevBuf d_tlas;
    uint32_t root2l = 0;
    bool two_level = false;
    rm::DevBuf d_env_pdfs, d_env_mip;
    rm::DevBuf d_spill;
    size_t spill_threads = 0;
    std::vector<std::unique_ptr<rm::DevBuf>> d_tex;  // texel arrays of the texture slots
    rm::DevBuf d_tex_table;                          // rm::Texture per slot
    rm::SceneView view{};
    bool accel_built = false;
    bool materials_built = false;
    bool envmap_built = false;
    rm::JitModule jit;
    rm::JitModule jit_cache;      // the library again with the radiance-cache shading kernel (on demand)
    std::string material_source;  // as compiled (batches recompile it with the batch kernel)
    rm_scene_stats stats{};
};

struct rm_renderer {
    rm_scene* scene = nullptr;
    uint32_t width = 0, height = 0, max_depth = 8, spl = 1;
    size_t capacity = 0;
    rm::DevBuf q_buf[2], hit_buf, shadow_buf, counters_buf, image_buf, image64_buf, resolve_buf, sort_buf;
    rm::PathQueue q[2]{};
    rm::HitQueue hits{};
    rm::ShadowQueue shadows{};
    rm_uniforms u{};
    uint32_t frame_index = 0;  // samples accumulated so far
    uint32_t last_samples = 0; // samples of the last rm_render call (the image's content when not accumulating)
    rm::Lens lens{};           // num_surfaces 0: pinhole
    rm::DevBuf lens_buf;
    std::vector<hipEvent_t> events;
    // radiance cache (rm_renderer_set_cache)
    rm_cache* cache = nullptr;
    uint32_t cache_mode = 0;           // 0 off, 1 query + train, 2 query only
    size_t slot_capacity = 0;          // training paths per stream
    size_t query_capacity = 0;         // lookups per stream
    uint32_t cache_frames = 0;         // samples rendered since the cache was attached
    uint32_t cache_parity = 0;         // which of the two queue sets the next stream fills
    rm::DevBuf slot_buf, shadow_slot_buf, target_buf;
    rm::DevBuf slot_rad_buf[2], query_buf[2], train_buf[2];  // per stream parity
    rm::DevBuf qsort_buf;  // lookup and training-record sorts: keys, order, bin starts and cursors
    // Training and lookups of one stream run on their own HIP stream while
    // the next stream traces; done[p] marks the end of the cache work that
    // reads the queue set of parity p.
    struct CacheOverlap {
        hipStream_t stream = nullptr;
        hipEvent_t loop_end = nullptr;
        hipEvent_t done[2] = {nullptr, nullptr};
        bool pending[2] = {false, false};
        ~CacheOverlap() {
            if (stream) (void)hipStreamDestroy(stream);
            if (loop_end) (void)hipEventDestroy(loop_end);
            for (hipEvent_t e : done)
                if (e) (void)hipEventDestroy(e);
        }
    } overlap;
    rm::CacheView cv{};
    rm::vec3 world_lo, world_hi;
    rm::BatchRun* batch = nullptr;     // set while a batch renders through this renderer
    // Completion polls of the bounce loop, one chunk behind the GPU: two
    // pinned slots (the next iteration's counters, then per-job done counts)
    // and an event per slot.
    struct LoopPoll {
        void* host = nullptr;
        size_t slot_bytes = 0;
        hipEvent_t ready[2] = {nullptr, nullptr};
        ~LoopPoll() {
            if (host) (void)hipHostFree(host);
            for (hipEvent_t e : ready)
                if (e) (void)hipEventDestroy(e);
        }
    } poll;
};
