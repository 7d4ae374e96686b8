/*
 * raymond-amd C ABI: the drop-in boundary between a render-loop host and the
 * MI355X (gfx950) wavefront path tracer.
 *
 * The reference has no FFI layer; its Swift host drives Metal directly.  Each
 * entry point below replaces one reference interface (cited per function):
 *   scene build    host/scene/SceneLoader.swift:46-143 (+ Shape/Entity/Light builders)
 *   material JIT   host/scene/MaterialBuilder.swift:253-325,889-903 (Metal JIT -> hiprtc)
 *   render loop    host/Renderer.swift:480-696 (execute), :757-798 (setSize), :729-732 (reset)
 *   readback       host/Renderer.swift:804-807 (saveFrame), normalizeImage.metal:39-59
 *
 * Conventions: opaque handles, int status (RM_OK = 0, negative on error),
 * thread-local message in rm_last_error().  All arrays are caller-owned and
 * copied; no torch or HIP types appear in signatures.  A renderer is
 * single-threaded with at most one rm_render in flight (the reference's
 * DispatchSemaphore(1), host/Renderer.swift:285,481-482).
 */
#ifndef RAYMOND_AMD_RM_H
#define RAYMOND_AMD_RM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RM_OK 0
#define RM_ERROR_INVALID (-1)
#define RM_ERROR_HIP (-2)
#define RM_ERROR_COMPILE (-3)
#define RM_ERROR_STATE (-4)
#define RM_ERROR_OOM (-5)

typedef struct rm_device rm_device;
typedef struct rm_scene rm_scene;
typedef struct rm_renderer rm_renderer;

/* Ray flags (bridge/Ray.hpp:6-15), used for instance visibility masks. */
#define RM_RAY_CAMERA 1u
#define RM_RAY_REFLECTION 2u
#define RM_RAY_TRANSMISSION 4u
#define RM_RAY_SHADOW 8u
#define RM_RAY_VOLUME 16u
#define RM_RAY_DIFFUSE 32u
#define RM_RAY_GLOSSY 64u
#define RM_RAY_SINGULAR 128u
#define RM_RAY_ALL 255u

/* Sampling modes (bridge/Uniforms.hpp:5-9) */
#define RM_SAMPLING_BSDF 0u
#define RM_SAMPLING_NEE 1u
#define RM_SAMPLING_MIS 2u
/* Tonemapping (bridge/Uniforms.hpp:11-15) */
#define RM_TONEMAP_LINEAR 0u
#define RM_TONEMAP_HABLE 1u
#define RM_TONEMAP_ACES 2u

const char* rm_last_error(void);
const char* rm_version(void);

/* ---- device ------------------------------------------------------------ */

typedef struct rm_device_info {
    int compute_units;
    int wavefront_size;
    uint64_t global_memory_bytes;
    char name[64];
    char arch[32];
} rm_device_info;

/* Replaces MTLCreateSystemDefaultDevice (Main.swift:22). */
int rm_device_create(int ordinal, rm_device** out);
void rm_device_destroy(rm_device* dev);
int rm_device_get_info(rm_device* dev, rm_device_info* out);

/* ---- scene -------------------------------------------------------------- */

/* One entity: an instance of a shape (a vertex/face range of the scene
 * arrays) under a transform (EntityBuilder.swift:66-77,97-141). */
typedef struct rm_instance_desc {
    uint32_t vertex_offset;
    uint32_t vertex_count;
    uint32_t face_offset;   /* faces index vertices relative to vertex_offset */
    uint32_t face_count;
    float transform[16];    /* column-major 4x4 (columns c0..c3), object -> world */
    uint32_t visibility;    /* RM_RAY_* mask */
    uint32_t emissive;      /* non-zero: register as a shape light (EntityBuilder.swift:103-114) */
    uint32_t has_bounds;    /* non-zero: bounds_min/bounds_max are the object-space bounds of the
                               "Generated" texture coordinates (PerInstanceData.hpp:13-14, as the
                               reference's PLY reader computes them, PLYReader.m:104-108); zero:
                               the exact bounds of the instance's vertices */
    float bounds_min[3];
    float bounds_max[3];
} rm_instance_desc;

typedef struct rm_light_info {
    uint32_t shader_index;  /* light-material index */
    uint32_t casts_shadows;
    uint32_t uses_mis;      /* the reference forces 0 (LightBuilder.swift:45) */
} rm_light_info;

typedef struct rm_area_light {
    rm_light_info info;
    float transform[12];    /* float3x4 columns t0,t1,t2 (each 4 floats) */
    float color[3];
    uint32_t is_circular;
} rm_area_light;

typedef struct rm_point_light {
    rm_light_info info;
    float location[3];
    float radius;
    float color[3];
} rm_point_light;

typedef struct rm_sun_light {
    rm_light_info info;
    float direction[3];
    float cos_angle;
    float color[3];
} rm_sun_light;

typedef struct rm_spot_light {
    rm_light_info info;
    float location[3];
    float direction[3];
    float radius;
    float color[3];
    float spot_size;        /* cos(size/2) */
    float spot_blend;
} rm_spot_light;

typedef struct rm_camera_desc {
    float transform[16];    /* column-major camera-to-world */
    float near_clip;
    float far_clip;
    float focal_length;     /* f_mm / (film_width / 2) (SceneLoader.swift:135-136) */
    float shift[2];
} rm_camera_desc;

/* An image of the material library (the reference's texture slots,
 * MaterialBuilder.swift:889-903): channels 1 (grey) or 4 (RGBA) floats per
 * texel, width x height texels, row 0 first (texture coordinate v from 0 at
 * row 0).  Slot i of the generated material code reads textures[i]; the
 * slots of a material library are listed by rmh_codegen_textures. */
typedef struct rm_texture_desc {
    uint32_t width, height;
    uint32_t channels;
    const float* pixels;
} rm_texture_desc;

typedef struct rm_scene_desc {
    uint32_t num_vertices;
    const float* vertices;      /* 3 per vertex, object space */
    const float* normals;       /* 3 per vertex */
    const float* texcoords;     /* 2 per vertex, may be NULL */
    uint32_t num_faces;
    const uint32_t* indices;    /* 3 per face, relative to the instance's vertex_offset */
    const uint16_t* face_materials; /* surface-material index per face */
    uint32_t num_instances;
    const rm_instance_desc* instances;
    uint32_t num_area_lights;
    const rm_area_light* area_lights;
    uint32_t num_point_lights;
    const rm_point_light* point_lights;
    uint32_t num_sun_lights;
    const rm_sun_light* sun_lights;
    uint32_t num_spot_lights;
    const rm_spot_light* spot_lights;
    uint32_t world_shader;      /* light-material index of the world (LightBuilder.swift:247-248) */
    uint32_t envmap_resolution; /* 0 = 2048 (LightBuilder.swift:90) */
    rm_camera_desc camera;
    uint32_t num_textures;
    const rm_texture_desc* textures;
} rm_scene_desc;

/* Copies the description, uploads geometry and light tables
 * (ShapeBuilder.build, EntityBuilder.build, LightBuilder.build minus envmap). */
int rm_scene_create(rm_device* dev, const rm_scene_desc* desc, rm_scene** out);
void rm_scene_destroy(rm_scene* scene);
/* Builds the BVH8 over all instances (replaces the Metal primitive and
 * instance acceleration structures, ShapeBuilder.swift:7-94, EntityBuilder.swift:146-165). */
int rm_scene_build_accel(rm_scene* scene, int host_threads);
/* Builds the two-level form instead: one BVH8 per shape (a distinct vertex
 * and face range) over its object-space triangles, and one over the
 * instances' world boxes whose leaves map the ray into the instance's object
 * space (the reference's primitive and instance acceleration structures,
 * ShapeBuilder.swift:7-94, EntityBuilder.swift:146-165).  A shape placed by
 * many instances is stored once.  Closest and any hit then traverse both
 * levels; hit distances are the same ray parameter as in the flat form, up
 * to the rounding of the object-space map. */
int rm_scene_build_accel_two_level(rm_scene* scene, int host_threads);
/* Compiles generated material source (shade_surface / shade_light switch
 * tables) together with the shading kernels via hiprtc.  `log` receives the
 * compiler log (may be NULL). (MaterialBuilder.swift:253-325) */
int rm_scene_compile_materials(rm_scene* scene, const char* source, size_t length, const char* options, char* log,
                               size_t log_capacity);
/* Builds the importance map of the world light (LightBuilder.swift:83-167). */
int rm_scene_build_envmap(rm_scene* scene);
int rm_scene_set_camera(rm_scene* scene, const rm_camera_desc* camera);

typedef struct rm_scene_stats {
    uint64_t triangles;
    uint64_t bvh_nodes;
    uint32_t bvh_depth;
    double bvh_build_seconds;
    double jit_seconds;
    double envmap_seconds;
    uint64_t device_bytes;      /* the scene's device buffers: geometry, trees, lights, envmap, textures, spill rows */
    uint64_t stored_triangles;  /* triangles in the acceleration structure (two-level: each shape once) */
    uint32_t two_level;         /* built by rm_scene_build_accel_two_level */
} rm_scene_stats;
int rm_scene_get_stats(rm_scene* scene, rm_scene_stats* out);

/* Reads back the envmap tables (resolution^2 pdfs, sum_k 4^k mip floats). */
int rm_scene_read_envmap(rm_scene* scene, float* pdfs, size_t pdf_count, float* mipmap, size_t mip_count);
/* Reads back the shape-light face CDFs (distribution.m:7-39) and per-light emissive areas. */
int rm_scene_read_light_tables(rm_scene* scene, float* light_faces, size_t face_count, float* emissive_areas,
                               size_t light_count);

/* Raw batch traversal for parity tests (raytrace.metal:11-85).  rays: n x 8
 * floats (ox, oy, oz, tmin, dx, dy, dz, tmax).  Outputs: t (< 0 on miss),
 * instance, primitive, barycentrics (2 per ray: weights of vertex 0 and 1). */
int rm_trace_closest(rm_scene* scene, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim,
                     float* bary);
int rm_trace_any(rm_scene* scene, const float* rays, uint32_t n, uint32_t* occluded);

/* Allocation-free closest-hit batch.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Results
 * equal rm_trace_closest's; the difference is that rays, t, inst, prim and
 * bary are device arrays in the same layouts (rays 16-byte aligned, bary
 * 8-byte aligned) and the call only enqueues work on `stream` (a
 * hipStream_t; null: the device's stream) without synchronising.  The
 * intermediate state lives in a fixed device array of the library (10.5 MB
 * per GPU, kept until the process ends), so a call takes at most 262144
 * rays (RM_ERROR_INVALID beyond) and calls that overlap on one GPU race.
 * Deep traversal stacks use the scene's spill rows, allocated when the
 * scene was built, as rm_trace_closest does. */
int rm_trace_closest_noalloc(rm_scene* scene, const float* rays, uint32_t n, float* t, uint32_t* inst,
                             uint32_t* prim, float* bary, void* stream);

/* ---- renderer ----------------------------------------------------------- */

typedef struct rm_uniforms {
    uint32_t sampling_mode;   /* RM_SAMPLING_*; default MIS */
    uint32_t accumulate;      /* default 1 */
    float exposure;           /* default 1 */
    uint32_t tonemapping;     /* RM_TONEMAP_*; default linear */
    uint32_t rr_depth;        /* Russian roulette from this depth; 0 = reference behaviour */
    uint32_t accumulate_f64;  /* also accumulate every launch into an FP64 buffer */
    uint32_t timing;          /* record per-stage GPU times (hipEvents) */
    uint32_t traversal_stats; /* count BVH nodes / triangles visited (slower) */
    uint32_t material_sort;   /* group each iteration's hits by material before shading */
    uint32_t use_graph;       /* replay the bounce loop from a captured hipGraph (ignored with timing) */
    uint32_t trace_blocks_per_cu; /* persistent trace workgroups per CU; 0 or more than fit = the most that fit */
    uint32_t sort_key_shift;  /* material-sort granularity: materials m >> shift share a bin (0: one per material) */
} rm_uniforms;

typedef struct rm_render_stats {
    /* Per bounce-loop iteration.  With samples <= samples_per_launch no path
     * regeneration happens and iteration == depth, i.e. the reference's
     * rayCounts / shadowRayCounts (Renderer.swift:190-196). */
    uint64_t rays[33];          /* closest-hit rays traced */
    uint64_t shadow_rays[32];   /* shadow rays traced */
    uint64_t total_rays;        /* closest_total + shadow_total */
    uint64_t closest_total;     /* all closest-hit rays traced */
    uint64_t shadow_total;      /* all shadow rays traced */
    uint64_t samples;           /* paths started */
    double ms_total;            /* wall time of the rm_render call */
    double ms_trace_closest;    /* stage GPU times when timing != 0 */
    double ms_shade;
    double ms_trace_any;
    uint32_t trace_closest_launches;
    uint32_t trace_any_launches;
    uint64_t iterations;        /* bounce-loop iterations */
    uint64_t nodes_closest, tris_closest, nodes_any, tris_any; /* traversal_stats */
    double ms_sort;             /* material sort (keys + scatter) when timing != 0 */
} rm_render_stats;

/* Replaces Renderer.init + setSize: samples_per_launch paths per pixel are
 * traced together in one wavefront.  About 2^24 paths in flight
 * (samples_per_launch = 2^24 / (width * height), at least 1) keeps every
 * bounce-loop iteration large enough to fill an MI355X; the queues then take
 * about 200 bytes per path. */
int rm_renderer_create(rm_scene* scene, uint32_t width, uint32_t height, uint32_t max_depth,
                       uint32_t samples_per_launch, rm_renderer** out);
void rm_renderer_destroy(rm_renderer* r);
int rm_renderer_set_uniforms(rm_renderer* r, const rm_uniforms* u);
/* Clears the accumulation and the frame counter (Renderer.reset). */
int rm_renderer_reset(rm_renderer* r);

/* A lens of the camera (the reference's lore::Surface array, loaded by
 * io/LensLoader.mm:17-42).  Surfaces from the object side to the image plane
 * (the last one): signed radius of the sphere cap (0: plane) with its vertex
 * on the axis, distance to the next surface toward the sensor, aperture
 * radius (mm), and the medium after the surface toward the sensor by n_d and
 * Abbe number V_d (0: no dispersion). */
typedef struct rm_lens_surface {
    float radius, thickness, aperture, ior_d, abbe;
} rm_lens_surface;

/* Lens settings (bridge/Uniforms.hpp:34-41; defaults Renderer.swift:333-343:
 * spectral 1, sensor_scale 1, camera_scale 0.001, focus 0, stop_index 0,
 * relative_stop 1, aperture_blades 7). */
typedef struct rm_lens_desc {
    uint32_t num_surfaces;           /* 0: the pinhole camera */
    const rm_lens_surface* surfaces;
    int32_t stop_index;              /* surface with the aperture blades (0: none) */
    float relative_stop;             /* opening of the stop, 0..1 */
    int32_t aperture_blades;
    float sensor_scale;              /* sensor size relative to 36 mm film */
    float camera_scale;              /* lens millimetres to scene units */
    float focus;                     /* sensor position along the axis (mm) */
    uint32_t spectral;               /* one wavelength per path, weighted to RGB */
} rm_lens_desc;

/* Renderer.setLens (Renderer.swift:738-746): camera rays start on the
 * sensor, aim at the rear element and are traced through the lens
 * (generate.metal:122-163); a ray the lens stops contributes nothing.
 * NULL or num_surfaces 0 restores the pinhole camera.  The surfaces are
 * copied. */
int rm_renderer_set_lens(rm_renderer* r, const rm_lens_desc* lens);
/* Renders `samples` more samples per pixel (Renderer.execute) as one
 * regenerating stream: the path queue holds width*height*samples_per_launch
 * paths and is refilled with new camera paths as paths terminate.  Blocks
 * until the GPU is done. */
int rm_render(rm_renderer* r, uint32_t samples, rm_render_stats* stats);
/* Accumulated RGBA sums (alpha counts contributions) -- width*height*4 floats. */
int rm_read_image(rm_renderer* r, float* rgba, size_t count);
int rm_read_image_f64(rm_renderer* r, double* rgba, size_t count);
/* exposure, / frames, tonemap (normalizeImage.metal:39-59) -- width*height*4 floats. */
int rm_resolve(rm_renderer* r, float* rgba, size_t count);
uint32_t rm_renderer_frame_index(rm_renderer* r);

/* Copies `bytes` (a multiple of 16) from one device buffer to another with a
 * streaming kernel (copy_probe_kernel) and returns its time: the known-size
 * reference that scales hardware byte counters.  No reference counterpart. */
int rm_device_copy_probe(rm_device* dev, size_t bytes, double* ms);

/* ------------------------------------------------------------------------
 * Learned radiance cache.  No reference counterpart: the reference traces
 * every path to termination.  A multiresolution hash-grid encoding (16 levels
 * x 2 FP16 features, resolutions 16..2048) feeds a 48-64-64-3 ReLU MLP with
 * FP16 weights; paths end at their first vertex after a diffuse bounce and add
 * throughput x cache(vertex).  One camera path in 2^train_shift is traced
 * unbiased instead and its vertices train the cache online (Adam) after every
 * bounce loop.
 *
 * Cache inputs, per vertex (RM_CACHE_FEATURES floats): position mapped to
 * [0,1]^3 over the scene bounds, then the outgoing direction and the shading
 * normal as (acos(z)/pi, atan2(y,x)/(2 pi) + 1/2), roughness, albedo RGB and
 * the specular-lobe probability. */
#define RM_CACHE_FEATURES 12
#define RM_CACHE_LEVELS 16
#define RM_CACHE_MLP_PARAMS 9216 /* W1 [48][64], W2 [64][64], W3 [64][32] (columns 3..31 unused) */

typedef struct rm_cache rm_cache;

typedef struct rm_cache_desc {
    uint32_t log2_table_size;  /* entries per level = 2^n; default 19 */
    uint32_t seed;             /* parameter initialisation */
    float learning_rate;       /* Adam; default 1e-2 */
    uint32_t train_shift;      /* one camera path in 2^n trains; default 4 */
    uint32_t batch_size;       /* samples per training step; default 8192 */
    uint32_t steps_per_launch; /* training steps per sample per pixel rendered; default 2 */
} rm_cache_desc;

typedef struct rm_cache_stats {
    uint64_t steps;            /* Adam steps taken */
    uint64_t samples;          /* training samples consumed */
    uint64_t queries;          /* cache lookups made by renders */
    uint64_t train_records;    /* training vertices recorded by renders */
    float last_loss;           /* mean relative L2 loss of the last step */
    double ms_infer, ms_train; /* GPU time of lookups and training (timing uniform) */
} rm_cache_stats;

/* Zeroed fields take the defaults. */
int rm_cache_create(rm_device* dev, const rm_cache_desc* desc, rm_cache** out);
void rm_cache_destroy(rm_cache* c);
int rm_cache_get_stats(rm_cache* c, rm_cache_stats* out);
/* Number of table floats: RM_CACHE_LEVELS * 2^log2_table_size * 2. */
uint64_t rm_cache_table_floats(rm_cache* c);
/* FP32 master parameters (the FP16 copies the kernels read are rounded from
 * these).  Either pointer may be null. */
int rm_cache_read_params(rm_cache* c, float* mlp, size_t mlp_count, float* table, size_t table_count);
int rm_cache_write_params(rm_cache* c, const float* mlp, size_t mlp_count, const float* table, size_t table_count);
/* Raw batches (host arrays, n x RM_CACHE_FEATURES row-major inputs):
 * infer writes n x 3 outputs; gradients computes the mean-loss gradients
 * without updating (grad_table may be null; table gradients are summed in
 * scaled FP16 on the device and returned as FP32); train takes `steps` Adam
 * steps on the batch. */
int rm_cache_infer(rm_cache* c, const float* features, uint32_t n, float* out);
int rm_cache_gradients(rm_cache* c, const float* features, const float* targets, uint32_t n, float* grad_mlp,
                       float* grad_table, float* loss);
int rm_cache_train(rm_cache* c, const float* features, const float* targets, uint32_t n, uint32_t steps,
                   float* last_loss);
/* ------------------------------------------------------------------------
 * Batches.  No reference counterpart: the reference renders one scene per
 * Renderer.  Many jobs (scene, resolution, samples, depth) render together in
 * one regenerating bounce loop: each job's BVH is a subtree of one merged
 * BVH8 entered at the job's root, each job keeps its lights and camera, and
 * the paths of all jobs share the queues, the kernels and (optionally) the
 * material sort.  The jobs' scenes must be fully built, share one material
 * library and outlive the batch, and either all have the flat or all the
 * two-level acceleration structure (each job then enters at its own
 * top-level root). */
typedef struct rm_batch rm_batch;

typedef struct rm_job_desc {
    rm_scene* scene;     /* camera, geometry, lights and envmap of the job */
    uint32_t width, height;
    uint32_t samples;    /* samples per pixel */
    uint32_t max_depth;
    uint32_t seed;       /* random seed of sample 0 (sample s uses seed + s); 0 = 1 */
} rm_job_desc;

typedef struct rm_batch_config {
    uint32_t queue_paths;    /* paths in flight; 0 = 2^24 */
    uint32_t material_sort;  /* group each iteration's hits by material */
    uint32_t rr_depth;       /* Russian roulette from this depth (0 = from the first bounce) */
    uint32_t timing;         /* per-stage GPU times */
    uint32_t use_graph;      /* replay each bounce loop from a captured hipGraph (ignored with timing) */
    uint32_t order;          /* RM_BATCH_SHORTEST_FIRST (default): jobs in order of their path
                              * count, so small jobs complete early and fewer jobs share the
                              * caches at a time; RM_BATCH_SAMPLE_MAJOR: every job progresses
                              * from the start */
    uint32_t trace_blocks_per_cu; /* persistent trace workgroups per CU (0: the most that fit) */
    uint32_t sort_key_shift; /* material-sort granularity (rm_uniforms.sort_key_shift) */
} rm_batch_config;

enum { RM_BATCH_SHORTEST_FIRST = 0, RM_BATCH_SAMPLE_MAJOR = 1 };

typedef struct rm_batch_stats {
    uint64_t total_rays, closest_rays, shadow_rays, samples, iterations;
    uint32_t loops;          /* bounce loops the batch was split into */
    uint32_t queue_paths;
    double ms_total;         /* wall time of the render */
    double makespan_ms;      /* completion time of the last job */
    double ms_trace_closest, ms_shade, ms_trace_any, ms_sort;
} rm_batch_stats;

typedef struct rm_batch_tune_entry {
    uint32_t queue_paths, material_sort, trace_blocks_per_cu, sort_key_shift;
    double mrays_per_s;
} rm_batch_tune_entry;

int rm_batch_create(rm_device* dev, const rm_job_desc* jobs, uint32_t n, rm_batch** out);
void rm_batch_destroy(rm_batch* b);
int rm_batch_render(rm_batch* b, const rm_batch_config* config, rm_batch_stats* stats);
/* Completion time of each job in the last render, milliseconds from its start. */
int rm_batch_job_times(rm_batch* b, double* ms, uint32_t n);
/* Accumulated RGBA sums of one job (width*height*4 floats; divide by samples). */
int rm_batch_read_image(rm_batch* b, uint32_t job, float* rgba, size_t count);
/* Picks queue depth and material sort for this batch by timing a prefix of
 * it (the first samples of every job) under each candidate, then, from the
 * best of those, the trace grid (workgroups per CU) and the sort granularity
 * (materials per bin); or takes the
 * choice recorded for the same device and batch signature in the tuner cache
 * ($RM_TUNE_CACHE, else ~/.cache/raymond_amd/batch_tune.txt).  `best` enters
 * with the other settings, which are kept.  entries/count report the
 * candidates timed (none on a cache hit). */
int rm_batch_autotune(rm_batch* b, rm_batch_config* best, rm_batch_tune_entry* entries, uint32_t capacity,
                      uint32_t* count, int* cache_hit);

/* Attaches a cache to a renderer (null detaches).  mode 1: query and train;
 * mode 2: query only (frozen weights).  The cache must outlive the renderer's
 * use of it. */
int rm_renderer_set_cache(rm_renderer* r, rm_cache* c, uint32_t mode);

#ifdef __cplusplus
}
#endif
#endif
