// The bounce loop (host/Renderer.swift:480-696) as a sequence of persistent
// kernels on one HIP stream.  Queue lengths stay on the device; the host only
// enqueues launches and reads the per-depth counters once per launch.
#include <chrono>
#include <cstring>

#include "../device/shade_kernels_abi.h"
#include "../kernels/misc_api.h"
#include "../kernels/sort_api.h"
#include "../kernels/trace_api.h"
#include "core.h"
#include "rcache_host.h"
#include "scene_api.h"

namespace rm {

static void carve_path_queue(DevBuf& buf, size_t cap, PathQueue& q) {
    buf.alloc(cap * 13 * 4);
    float* f = buf.as<float>();
    float** fields[10] = {&q.ox, &q.oy, &q.oz, &q.dx, &q.dy, &q.dz, &q.wx, &q.wy, &q.wz, &q.pdf};
    for (int k = 0; k < 10; k++) *fields[k] = f + (size_t)k * cap;
    uint32_t* u = reinterpret_cast<uint32_t*>(f + 10 * cap);
    q.seed = u;
    q.meta = u + cap;
    q.pixel = u + 2 * cap;
}

void renderer_create(rm_scene* s, uint32_t w, uint32_t h, uint32_t max_depth, uint32_t spl, rm_renderer* r) {
    RM_REQUIRE(spl >= 1, "samples_per_launch must be >= 1");
    RM_REQUIRE((uint64_t)w * h * spl < (1ull << 31), "wavefront too large");
    renderer_init(s, w, h, max_depth, (size_t)w * h * spl, r);
    r->spl = spl;
}

void renderer_init(rm_scene* s, uint32_t w, uint32_t h, uint32_t max_depth, size_t capacity, rm_renderer* r) {
    RM_REQUIRE(w > 0 && h > 0, "image size must be positive");
    RM_REQUIRE(max_depth >= 1 && max_depth <= 255, "max_depth out of range");
    RM_REQUIRE(capacity >= 1 && capacity < (1ull << 31), "path queue capacity out of range");
    r->scene = s;
    r->width = w;
    r->height = h;
    r->max_depth = max_depth;
    r->spl = 1;
    r->capacity = capacity;
    const size_t cap = r->capacity;
    carve_path_queue(r->q_buf[0], cap, r->q[0]);
    carve_path_queue(r->q_buf[1], cap, r->q[1]);
    r->hit_buf.alloc(cap * 5 * 4);
    float* hf = r->hit_buf.as<float>();
    r->hits.t = hf;
    r->hits.prim = reinterpret_cast<uint32_t*>(hf + cap);
    r->hits.inst = reinterpret_cast<uint32_t*>(hf + 2 * cap);
    r->hits.b0 = hf + 3 * cap;
    r->hits.b1 = hf + 4 * cap;
    r->shadow_buf.alloc(cap * 11 * 4);
    float* sf = r->shadow_buf.as<float>();
    ShadowQueue& sq = r->shadows;
    float** sfields[10] = {&sq.ox, &sq.oy, &sq.oz, &sq.dx, &sq.dy, &sq.dz, &sq.tmax, &sq.wx, &sq.wy, &sq.wz};
    for (int k = 0; k < 10; k++) *sfields[k] = sf + (size_t)k * cap;
    sq.pixel = reinterpret_cast<uint32_t*>(sf + 10 * cap);
    r->counters_buf.alloc(sizeof(Counters));
    r->image_buf.alloc((size_t)w * h * 16);
    r->resolve_buf.alloc((size_t)w * h * 16);
    r->u.sampling_mode = RM_SAMPLING_MIS;
    r->u.accumulate = 1;
    r->u.exposure = 1.0f;
    r->u.tonemapping = RM_TONEMAP_LINEAR;
    r->u.rr_depth = 0;
    renderer_reset(r);
}

static void join_cache_work(rm_renderer* r);

void renderer_set_lens(rm_renderer* r, const rm_lens_desc* d) {
    if (!d || d->num_surfaces == 0) {
        r->lens = Lens{};
        return;
    }
    RM_REQUIRE(d->surfaces != nullptr, "lens surfaces missing");
    RM_REQUIRE(d->num_surfaces >= 2, "a lens needs a rear element and the image plane");
    RM_REQUIRE(d->stop_index >= 0 && d->stop_index < (int32_t)d->num_surfaces, "lens stop index out of range");
    RM_REQUIRE(d->aperture_blades >= 3, "an aperture needs at least 3 blades");
    RM_REQUIRE(d->relative_stop > 0.0f && d->relative_stop <= 1.0f, "relative stop must be in (0, 1]");
    std::vector<LensSurface> s(d->num_surfaces);
    for (uint32_t i = 0; i < d->num_surfaces; i++) {
        const rm_lens_surface& a = d->surfaces[i];
        RM_REQUIRE(a.aperture > 0.0f && a.ior_d > 0.0f, "lens surface " + std::to_string(i) +
                                                           " needs a positive aperture and index");
        s[i] = LensSurface{a.radius, a.thickness, a.aperture, a.ior_d, a.abbe};
    }
    hipStream_t st = r->scene->dev->stream;
    r->lens_buf.upload(s.data(), s.size() * sizeof(LensSurface), st);
    RM_HIP(hipStreamSynchronize(st));
    Lens& L = r->lens;
    L.surfaces = r->lens_buf.as<LensSurface>();
    L.num_surfaces = (int)d->num_surfaces;
    L.stop_index = d->stop_index;
    L.relative_stop = d->relative_stop;
    L.aperture_blades = d->aperture_blades;
    L.sensor_scale = d->sensor_scale;
    L.camera_scale = d->camera_scale;
    L.focus = d->focus;
    L.spectral = d->spectral ? 1 : 0;
}

void renderer_reset(rm_renderer* r) {
    hipStream_t st = r->scene->dev->stream;
    // pending cache lookups of the last stream add into the image
    join_cache_work(r);
    RM_HIP(hipMemsetAsync(r->image_buf.ptr, 0, r->image_buf.bytes, st));
    if (r->image64_buf.ptr) RM_HIP(hipMemsetAsync(r->image64_buf.ptr, 0, r->image64_buf.bytes, st));
    RM_HIP(hipStreamSynchronize(st));
    r->frame_index = 0;
    r->last_samples = 0;
}

namespace {

struct StageTimer {
    bool on;
    hipStream_t st;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> closest, sort, shade, any;
    std::vector<hipEvent_t>& pool;
    size_t used = 0;
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            RM_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    hipEvent_t begin() {
        if (!on) return nullptr;
        hipEvent_t e = get();
        RM_HIP(hipEventRecord(e, st));
        return e;
    }
    void end(hipEvent_t b, std::vector<std::pair<hipEvent_t, hipEvent_t>>& into) {
        if (!on) return;
        hipEvent_t e = get();
        RM_HIP(hipEventRecord(e, st));
        into.push_back({b, e});
    }
    static double sum(const std::vector<std::pair<hipEvent_t, hipEvent_t>>& v) {
        double ms = 0;
        for (auto& p : v) {
            float t = 0;
            RM_HIP(hipEventElapsedTime(&t, p.first, p.second));
            ms += t;
        }
        return ms;
    }
};

}  // namespace

// One regenerating stream of `samples` samples per pixel.  Iteration i runs
// regenerate(i) -> closest(i) -> shade(i) -> shadow(i); queue lengths live in
// device counters.  The host enqueues chunks of kIterChunk iterations and
// checks for completion one chunk behind the GPU, so the GPU never waits on
// the host; a chunk enqueued after the loop drained finds empty queues.
static constexpr uint32_t kIterChunk = 8;
static constexpr uint32_t kCacheStreamMax = 8;

void renderer_set_cache(rm_renderer* r, rm_cache* c, uint32_t mode) {
    if (c) scene_cache_module(r->scene);
    r->cache = c;
    r->cache_mode = c ? mode : 0u;
    r->cache_frames = 0;
    if (!c) return;
    const std::vector<BuildTri> tris = world_triangles(r->scene);
    vec3 lo(kInf), hi(-kInf);
    for (const BuildTri& t : tris)
        for (const vec3& p : {t.v0, t.v1, t.v2}) {
            lo = vec3(min_(lo.x, p.x), min_(lo.y, p.y), min_(lo.z, p.z));
            hi = vec3(max_(hi.x, p.x), max_(hi.y, p.y), max_(hi.z, p.z));
        }
    RM_REQUIRE(!tris.empty(), "the radiance cache needs scene geometry");
    r->world_lo = lo;
    r->world_hi = hi;
}

// Waits for the cache work of both queue sets (before buffers grow, and
// before a render call returns).
static void join_cache_work(rm_renderer* r) {
    if (r->overlap.stream) RM_HIP(hipStreamSynchronize(r->overlap.stream));
    r->overlap.pending[0] = r->overlap.pending[1] = false;
}

// Buffers of the radiance cache for a stream of `samples` samples per pixel,
// filling the queue set of parity `par`.  They are sized for the longest
// stream at once, so that growing streams allocate nothing on the way.
// The spatial sorts of a stream's lookups and of its training records share
// qsort_buf (the records are sorted and consumed before the lookups): keys
// and order (one word per query each), then the bin starts and cursors.
static size_t qsort_bytes(size_t queries) { return queries * 2 * 4 + 2 * kRcSortBins * 4; }
static RcSortArgs qsort_layout(rm_renderer* r) {
    uint32_t* qs = r->qsort_buf.as<uint32_t>();
    const size_t qc = r->query_capacity;
    RM_REQUIRE(r->qsort_buf.bytes >= qsort_bytes(qc), "lookup sort buffer smaller than its layout");
    RcSortArgs so{};
    so.keys = qs;
    so.order = qs + qc;
    so.hist = qs + 2 * qc;
    so.cursor = so.hist + kRcSortBins;
    return so;
}

static void prepare_cache(rm_renderer* r, uint32_t samples, uint32_t par) {
    rm_cache* c = r->cache;
    CacheView& cv = r->cv;
    const size_t cap = r->capacity;
    const bool train = r->cache_mode == 1u;
    const uint32_t shift = c->desc.train_shift;
    const size_t total = (size_t)r->width * r->height * std::max(samples, kCacheStreamMax);
    const size_t slots = train ? ((total + (1ull << shift) - 1) >> shift) + 1 : 0;
    // every path of the stream looks up at most once
    const size_t queries = std::max<size_t>(cap, total);
    if (queries > r->query_capacity || (train && slots > r->slot_capacity)) join_cache_work(r);
    if (queries > r->query_capacity) {
        r->query_capacity = queries;
        for (DevBuf& b : r->query_buf) b.alloc(queries * kRcQueryFloats * 4);
        r->qsort_buf.alloc(qsort_bytes(queries));
        RM_HIP(hipMemsetAsync(r->qsort_buf.ptr, 0, r->qsort_buf.bytes, r->scene->dev->stream));
    }
    if (train && slots > r->slot_capacity) {
        r->slot_capacity = slots;
        for (DevBuf& b : r->slot_rad_buf) b.alloc(slots * 3 * 4);
        for (DevBuf& b : r->train_buf) b.alloc(slots * (7 + kRcFeatureFloats) * 4);
        r->target_buf.alloc(slots * 3 * 4);
    }
    if (train && !r->slot_buf.ptr) {
        r->slot_buf.alloc(cap * 2 * 4);
        r->shadow_slot_buf.alloc(cap * 4);
    }
    cv = CacheView{};
    cv.enabled = 1u;
    cv.train_shift = shift;
    cv.query_capacity = (uint32_t)r->query_capacity;
    cv.train_capacity = (uint32_t)r->slot_capacity;
    cv.slot_capacity = (uint32_t)r->slot_capacity;
    cv.bounds_min = r->world_lo;
    const vec3 ext = r->world_hi - r->world_lo;
    cv.bounds_inv_extent = vec3(1.0f / max_(ext.x, 1e-6f), 1.0f / max_(ext.y, 1e-6f), 1.0f / max_(ext.z, 1e-6f));
    cv.queries.rows = r->query_buf[par].as<float>();
    if (train) {
        const size_t sc = r->slot_capacity;
        float* tf = r->train_buf[par].as<float>();
        cv.train.slot = r->train_buf[par].as<uint32_t>();
        cv.train.tx = tf + sc;
        cv.train.ty = tf + 2 * sc;
        cv.train.tz = tf + 3 * sc;
        cv.train.rx = tf + 4 * sc;
        cv.train.ry = tf + 5 * sc;
        cv.train.rz = tf + 6 * sc;
        cv.train.feat = tf + 7 * sc;
        cv.slot_radiance = r->slot_rad_buf[par].as<float>();
    }
}

static void init_cache_overlap(rm_renderer* r) {
    auto& o = r->overlap;
    if (o.stream) return;
    RM_HIP(hipStreamCreateWithFlags(&o.stream, hipStreamNonBlocking));
    RM_HIP(hipEventCreateWithFlags(&o.loop_end, hipEventDisableTiming));
    for (hipEvent_t& e : o.done) RM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// Records the completion time of every job whose paths have all ended,
// from per-job done counts already copied to the host.
static void batch_poll_counts(BatchRun* br, const uint32_t* done) {
    const uint32_t n = br->view.num_jobs;
    const double now =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    for (uint32_t j = 0; j < n; j++)
        if ((*br->job_ms)[j] < 0.0 && done[j] >= (*br->job_paths)[j]) (*br->job_ms)[j] = now - br->t0_ms;
}
static void batch_poll(BatchRun* br, hipStream_t st) {
    std::vector<uint32_t> done(br->view.num_jobs);
    RM_HIP(hipMemcpyAsync(done.data(), br->view.done, done.size() * 4, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    batch_poll_counts(br, done.data());
}

// The renderer's pinned poll slots, grown to `slot_bytes` each when needed.
struct LoopPollRef {
    char* host;
    size_t slot_bytes;
    hipEvent_t* ready;
    IterCounters* probe(int slot) const { return reinterpret_cast<IterCounters*>(host + slot * slot_bytes); }
    uint32_t* done(int slot) const {
        return reinterpret_cast<uint32_t*>(host + slot * slot_bytes + sizeof(IterCounters));
    }
};
static LoopPollRef loop_poll(rm_renderer* r, size_t slot_bytes) {
    auto& p = r->poll;
    if (!p.ready[0]) {
        for (hipEvent_t& e : p.ready) RM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (p.slot_bytes < slot_bytes) {
        if (p.host) RM_HIP(hipHostFree(p.host));
        p.host = nullptr;
        p.slot_bytes = 0;
        RM_HIP(hipHostMalloc(&p.host, 2 * slot_bytes, hipHostMallocDefault));
        p.slot_bytes = slot_bytes;
    }
    return LoopPollRef{static_cast<char*>(p.host), p.slot_bytes, p.ready};
}

static void render_stream(rm_renderer* r, uint32_t samples, rm_render_stats* stats) {
    rm_scene* s = r->scene;
    rm_device* dev = s->dev;
    hipStream_t st = dev->stream;
    const uint32_t W = r->width, H = r->height;
    const bool tstats = r->u.traversal_stats != 0;
    Counters* cnt = r->counters_buf.as<Counters>();
    float* image = r->image_buf.as<float>();
    const int cus = dev->props.multiProcessorCount;
    // the persistent grids: at most what fits (the spill rows are sized for it)
    const uint32_t want = r->u.trace_blocks_per_cu;
    const int tbpc = want ? std::min((int)want, dev->trace_blocks_per_cu) : dev->trace_blocks_per_cu;
    const int abpc = want ? std::min((int)want, dev->any_blocks_per_cu) : dev->any_blocks_per_cu;
    const int tblocks = cus * tbpc;
    const int ablocks = cus * abpc;
    BatchRun* br = r->batch;
    hipFunction_t shade_fn = br ? s->jit.shade_batch : (r->cache_mode ? s->jit_cache.shade_cache : s->jit.shade);
    RM_REQUIRE(shade_fn != nullptr, "material module lacks the batch shading kernel");
    int shade_bpc = 0;
    RM_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&shade_bpc, shade_fn, 256, 0));
    if (shade_bpc < 1) shade_bpc = 1;
    const int sblocks = cus * shade_bpc;
    const uint32_t total = br ? br->total : W * H * samples;

    StageTimer tm{r->u.timing != 0, st, {}, {}, {}, {}, r->events};
    RM_HIP(launch_clear_counters(cnt, st));

    FrameUniforms fu{};
    fu.width = W;
    fu.height = H;
    fu.samples = samples;
    fu.random_seed = r->frame_index + 1;
    fu.sampling_mode = r->u.sampling_mode;
    fu.rr_depth = r->u.rr_depth;
    fu.max_depth = r->max_depth;
    fu.image_offset = 0;

    RegenArgs ga{};
    ga.lens = br ? Lens{} : r->lens;
    ga.counters = cnt;
    ga.capacity = (uint32_t)r->capacity;
    ga.total = total;
    ga.camera = s->view.camera;
    ga.width = W;
    ga.height = H;
    ga.random_seed = fu.random_seed;
    ga.image_offset = 0;
    if (br) {
        ga.slices = br->slices;
        ga.num_slices = br->num_slices;
        ga.jobs = br->view.jobs;
    }

    ClosestArgs ca{};
    ca.nodes = s->d_nodes.as<Bvh8Node>();
    ca.tris = s->d_tris.as<BvhTri>();
    ca.hits = r->hits;
    ca.spill = s->d_spill.as<uint2>();
    ca.counters = cnt;
    ca.camera = s->view.camera;
    ca.tlas = s->two_level ? s->d_tlas.as<TlasInstance>() : nullptr;
    ca.root2l = s->root2l;
    if (br) ca.batch = br->view;

    const bool sorting = r->u.material_sort != 0;
    SortArgs so{};
    if (sorting) {
        if (!r->sort_buf.ptr) r->sort_buf.alloc(r->capacity * 2 * sizeof(uint32_t));
        so.hits = r->hits;
        so.instances = s->view.instances;
        so.materials = s->view.materials;
        so.keys = r->sort_buf.as<uint32_t>();
        so.order = so.keys + r->capacity;
        so.counters = cnt;
        so.key_shift = std::min(r->u.sort_key_shift, 15u);
    }
    const int sort_blocks = cus * 4;

    const bool cached = r->cache_mode != 0u;
    const bool cache_train = r->cache_mode == 1u;
    // the cache work overlaps the next stream unless stages are being timed
    // or the stream's image is accumulated right after it
    const bool overlap = cached && !tm.on && !r->u.accumulate_f64;
    const uint32_t par = r->cache_parity;
    if (cached) {
        prepare_cache(r, samples, par);
        if (r->overlap.pending[par]) RM_HIP(hipStreamWaitEvent(st, r->overlap.done[par], 0));
    }
    uint32_t* slot_arr[2] = {nullptr, nullptr};
    if (cache_train) {
        slot_arr[0] = r->slot_buf.as<uint32_t>();
        slot_arr[1] = slot_arr[0] + r->capacity;
        ga.slot_radiance = r->cv.slot_radiance;
        ga.train_shift = r->cv.train_shift;
    }
    std::vector<std::pair<hipEvent_t, hipEvent_t>> t_infer;

    AnyArgs aa{};
    aa.nodes = ca.nodes;
    aa.tris = ca.tris;
    aa.tlas = ca.tlas;
    aa.root2l = ca.root2l;
    aa.shadows = r->shadows;
    aa.image = image;
    aa.spill = ca.spill;
    aa.counters = cnt;
    if (br) aa.batch = br->view;

    // One iteration: regenerate, closest hit, material sort, shade, shadow
    // rays.  Directly launched, `it` is the iteration; inside a captured chunk
    // it is the offset from the device counter iter_base (chunks start at even
    // iterations, so the queue parity is the offset's).
    const uint32_t* iter_dev = nullptr;
    auto enqueue = [&](uint32_t it) {
        PathQueue& cur = r->q[it & 1];
        PathQueue& next = r->q[(it + 1) & 1];
        ga.iter = it;
        ga.iter_dev = iter_dev;
        ga.paths = cur;
        ga.slots = slot_arr[it & 1];
        RM_HIP(launch_regenerate(ga, cus * 4, st));
        ca.iter = it;
        ca.iter_dev = iter_dev;
        ca.count_ptr = iter_dev ? nullptr : &cnt->it[it].queue;
        ca.fetch = iter_dev ? nullptr : &cnt->it[it].fetch_closest;
        ca.paths = cur;
        hipEvent_t e = tm.begin();
        RM_HIP(launch_trace_closest(ca, tblocks, tstats, st));
        tm.end(e, tm.closest);
        if (sorting) {
            so.iter = it;
            so.iter_dev = iter_dev;
            hipEvent_t eo = tm.begin();
            RM_HIP(launch_sort_keys(so, sort_blocks, st));
            RM_HIP(launch_sort_scatter(so, sort_blocks, st));
            tm.end(eo, tm.sort);
        }
        ShadeArgs sa;
        sa.ctx = s->view;
        sa.cur = cur;
        sa.hits = r->hits;
        sa.next = next;
        sa.shadows = r->shadows;
        sa.counters = cnt;
        sa.order = sorting ? so.order : nullptr;
        sa.cache = CacheView{};
        sa.batch = br ? br->view : BatchView{};
        if (cached) {
            sa.cache = r->cv;
            sa.cache.slots = slot_arr[it & 1];
            sa.cache.slots_next = slot_arr[(it + 1) & 1];
            sa.shadows.slot = cache_train ? r->shadow_slot_buf.as<uint32_t>() : nullptr;
        }
        sa.image = image;
        sa.u = fu;
        sa.iter = it;
        sa.iter_dev = iter_dev;
        void* args[] = {&sa};
        hipEvent_t es = tm.begin();
        RM_HIP(hipModuleLaunchKernel(shade_fn, sblocks, 1, 1, 256, 1, 1, 0, st, args, nullptr));
        tm.end(es, tm.shade);
        if (cached) {
            aa.shadows.slot = cache_train ? r->shadow_slot_buf.as<uint32_t>() : nullptr;
            aa.slot_radiance = cache_train ? r->cv.slot_radiance : nullptr;
        }
        aa.iter = it;
        aa.iter_dev = iter_dev;
        aa.count_ptr = iter_dev ? nullptr : &cnt->it[it].shadows;
        aa.fetch = iter_dev ? nullptr : &cnt->it[it].fetch_any;
        hipEvent_t ea = tm.begin();
        RM_HIP(launch_trace_any(aa, ablocks, tstats, st));
        tm.end(ea, tm.any);
    };

    // Captured bounce loop: a chunk of kIterChunk iterations and the advance
    // of iter_base, recorded once per stream and replayed until it drains.
    // The graph holds kernels only (no memset nodes).
    hipGraphExec_t chunk_exec = nullptr;
    if (r->u.use_graph && !tm.on) {
        iter_dev = &cnt->iter_base;
        hipGraph_t g = nullptr;
        RM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            for (uint32_t c = 0; c < kIterChunk; c++) enqueue(c);
            RM_HIP(launch_advance_iter(cnt, kIterChunk, st));
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        RM_HIP(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&chunk_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RM_HIP(e);
    }
    struct ExecGuard {
        hipGraphExec_t& e;
        ~ExecGuard() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } exec_guard{chunk_exec};

    // pinned poll slots: the counters of the iteration after a chunk, then
    // the per-job done counts of a batch
    const uint32_t njobs = br ? br->view.num_jobs : 0u;
    const size_t slot_bytes = (sizeof(IterCounters) + (size_t)njobs * 4 + 255) & ~(size_t)255;
    LoopPollRef lp = loop_poll(r, slot_bytes);
    uint32_t iter = 0;
    uint32_t chunk_end[2] = {0, 0};
    auto enqueue_chunk = [&](int slot) {
        RM_REQUIRE(iter + kIterChunk + 1 < (uint32_t)kMaxIters, "render stream exceeded the iteration limit");
        if (chunk_exec) {
            RM_HIP(hipGraphLaunch(chunk_exec, st));
            iter += kIterChunk;
        } else {
            for (uint32_t c = 0; c < kIterChunk; c++, iter++) enqueue(iter);
        }
        RM_HIP(hipMemcpyAsync(lp.probe(slot), &cnt->it[iter], sizeof(IterCounters), hipMemcpyDeviceToHost, st));
        if (br) RM_HIP(hipMemcpyAsync(lp.done(slot), br->view.done, (size_t)njobs * 4, hipMemcpyDeviceToHost, st));
        RM_HIP(hipEventRecord(lp.ready[slot], st));
        chunk_end[slot] = iter;
    };
    std::vector<IterCounters> host_it;
    int slot = 0;
    enqueue_chunk(0);
    for (;;) {
        enqueue_chunk(slot ^ 1);  // runs while the host reads the previous chunk's counters
        RM_HIP(hipEventSynchronize(lp.ready[slot]));
        if (br) batch_poll_counts(br, lp.done(slot));
        const IterCounters& probe = *lp.probe(slot);
        if (probe.generated >= total && probe.appended == 0) break;
        slot ^= 1;
    }
    // the chunk after the last one held no work: the stream's statistics and
    // stage timers end with the chunk that drained it
    iter = chunk_end[slot];
    for (auto* v : {&tm.closest, &tm.sort, &tm.shade, &tm.any})
        if (v->size() > iter) v->resize(iter);
    uint32_t cache_ctr[2] = {0, 0};  // lookups, training vertices
    if (cached) {
        RM_HIP(hipMemcpyAsync(cache_ctr, &cnt->cache_queries, 8, hipMemcpyDeviceToHost, st));
        RM_HIP(hipStreamSynchronize(st));
        rm_cache* c = r->cache;
        c->queries += cache_ctr[0];
        hipStream_t cs = st;
        if (overlap) {
            init_cache_overlap(r);
            cs = r->overlap.stream;
            RM_HIP(hipEventRecord(r->overlap.loop_end, st));
            RM_HIP(hipStreamWaitEvent(cs, r->overlap.loop_end, 0));
        }
        const uint32_t records = std::min<uint32_t>(cache_ctr[1], r->cv.train_capacity);
        if (cache_train && records > 0) {
            // targets of this stream's training vertices, then Adam steps on
            // them (steps_per_launch per sample of the stream), before the
            // stream's lookups so that they see what the stream taught
            std::vector<std::pair<hipEvent_t, hipEvent_t>> t_train;
            hipEvent_t et = tm.begin();
            RM_HIP(launch_rc_targets(r->cv.train, r->cv.slot_radiance, records, r->cv.train_capacity,
                                     r->target_buf.as<float>(), cs));
            // the records in spatial order (the lookup sort's buffers are
            // free until the lookups), sampled systematically per step
            RcSortArgs so = qsort_layout(r);
            so.pos = r->cv.train.feat;
            so.pos_elem = 1;
            so.pos_comp = r->cv.train_capacity;
            so.count = records;
            RM_HIP(launch_rc_sort_queries(so, cus * 4, cs));
            RcSamples smp{};
            smp.feat = r->cv.train.feat;
            smp.target = r->target_buf.as<float>();
            smp.stride = r->cv.train_capacity;
            smp.count = records;
            smp.order = so.order;
            const uint32_t batch = std::min<uint32_t>(c->desc.batch_size, records);
            const uint32_t steps = c->desc.steps_per_launch * samples;
            for (uint32_t k = 0; k < steps; k++) {
                // golden-ratio sequence of offsets
                const double u = (double)(c->steps + 1) * 0.6180339887498949;
                smp.offset = (float)(u - std::floor(u));
                cache_step(c, smp, batch, false, cs);
            }
            tm.end(et, t_train);
            c->records += records;
            if (tm.on) {
                RM_HIP(hipStreamSynchronize(st));
                c->ms_train += StageTimer::sum(t_train);
            }
        }
        // the lookups of the whole stream in one batch (counts by value: the
        // next stream clears the counters): sorted by position, then
        // predicted and added to their pixels in that order
        const uint32_t nq = std::min<uint32_t>(cache_ctr[0], r->cv.query_capacity);
        hipEvent_t ei = tm.begin();
        if (nq > 0) {
            RcSortArgs so = qsort_layout(r);
            so.pos = r->cv.queries.rows + 4;
            so.pos_elem = kRcQueryFloats;
            so.pos_comp = 1;
            so.count = nq;
            RM_HIP(launch_rc_sort_queries(so, cus * 4, cs));
            RcInferArgs ia{};
            ia.p = cache_params(r->cache);
            ia.rows = r->cv.queries.rows;
            ia.order = so.order;
            ia.count = nq;
            ia.image = image;
            RM_HIP(launch_rc_infer(ia, cus * 8, cs));
        }
        tm.end(ei, t_infer);
        if (overlap) {
            RM_HIP(hipEventRecord(r->overlap.done[par], cs));
            r->overlap.pending[par] = true;
        }
        r->cache_frames += samples;
        r->cache_parity = par ^ 1u;
    }
    if (r->u.accumulate_f64) {
        if (!r->image64_buf.ptr) {
            r->image64_buf.alloc((size_t)W * H * 4 * sizeof(double));
            RM_HIP(hipMemsetAsync(r->image64_buf.ptr, 0, r->image64_buf.bytes, st));
        }
        RM_HIP(launch_accum_f64(image, r->image64_buf.as<double>(), W * H * 4, st));
    }
    host_it.resize(iter);
    RM_HIP(hipMemcpyAsync(host_it.data(), cnt->it, iter * sizeof(IterCounters), hipMemcpyDeviceToHost, st));
    unsigned long long tstat[4] = {0, 0, 0, 0};
    RM_HIP(hipMemcpyAsync(tstat, &cnt->nodes_closest, sizeof(tstat), hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    if (cached && tm.on) r->cache->ms_infer += StageTimer::sum(t_infer);
    r->frame_index += samples;
    if (stats) {
        for (uint32_t i = 0; i < iter; i++) {
            // paths the lens stopped hold a queue entry but no ray
            const uint32_t traced = host_it[i].queue - host_it[i].voids;
            if (i < 33) stats->rays[i] += traced;
            if (i < 32) stats->shadow_rays[i] += host_it[i].shadows;
            stats->total_rays += (uint64_t)traced + host_it[i].shadows;
            stats->closest_total += traced;
            stats->shadow_total += host_it[i].shadows;
        }
        stats->samples += (uint64_t)total;
        stats->iterations += iter;
        stats->trace_closest_launches += (uint32_t)tm.closest.size();
        stats->trace_any_launches += (uint32_t)tm.any.size();
        if (tm.on) {
            stats->ms_trace_closest += StageTimer::sum(tm.closest);
            stats->ms_shade += StageTimer::sum(tm.shade);
            stats->ms_sort += StageTimer::sum(tm.sort);
            stats->ms_trace_any += StageTimer::sum(tm.any);
        }
        stats->nodes_closest += tstat[0];
        stats->tris_closest += tstat[1];
        stats->nodes_any += tstat[2];
        stats->tris_any += tstat[3];
    }
}

void renderer_render_batch_loop(rm_renderer* r, BatchRun* br, rm_render_stats* stats) {
    r->batch = br;
    try {
        render_stream(r, 1, stats);
    } catch (...) {
        r->batch = nullptr;
        throw;
    }
    r->batch = nullptr;
    batch_poll(br, r->scene->dev->stream);
}

void renderer_render(rm_renderer* r, uint32_t samples, rm_render_stats* stats) {
    rm_scene* s = r->scene;
    RM_REQUIRE(s->accel_built, "acceleration structure not built");
    RM_REQUIRE(s->materials_built, "materials not compiled");
    RM_REQUIRE(s->envmap_built, "envmap not built");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const auto t0 = std::chrono::steady_clock::now();
    if (!r->u.accumulate) {
        // the image holds this call's samples only.  The cache work of earlier
        // streams adds into it on the overlap stream, so it is joined before
        // the one clear of the call; the streams of this call then all add.
        join_cache_work(r);
        RM_HIP(hipMemsetAsync(r->image_buf.ptr, 0, r->image_buf.bytes, s->dev->stream));
    }
    r->last_samples = samples;
    // bound a stream so its iteration count stays within kMaxIters.  With a
    // radiance cache, streams double in length from one sample to
    // kCacheStreamMax as the cache sees more samples: the cache learns
    // between streams, and a longer stream spreads the drain of its last
    // (training) paths over more samples.
    uint32_t left = samples;
    while (left > 0) {
        uint32_t max_stream = 512;
        if (r->cache_mode != 0u) {
            max_stream = 1;
            while (max_stream * 2 <= r->cache_frames && max_stream * 2 <= kCacheStreamMax) max_stream *= 2;
        }
        const uint32_t n = left < max_stream ? left : max_stream;
        try {
            render_stream(r, n, stats);
        } catch (...) {
            join_cache_work(r);
            throw;
        }
        left -= n;
    }
    join_cache_work(r);
    if (stats) stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void renderer_read_image(rm_renderer* r, float* rgba, size_t count) {
    RM_REQUIRE(count >= (size_t)r->width * r->height * 4, "output buffer too small");
    hipStream_t st = r->scene->dev->stream;
    RM_HIP(hipMemcpyAsync(rgba, r->image_buf.ptr, (size_t)r->width * r->height * 16, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
}

void renderer_read_image_f64(rm_renderer* r, double* rgba, size_t count) {
    RM_REQUIRE(count >= (size_t)r->width * r->height * 4, "output buffer too small");
    RM_REQUIRE(r->image64_buf.ptr != nullptr, "FP64 accumulation not enabled");
    hipStream_t st = r->scene->dev->stream;
    RM_HIP(hipMemcpyAsync(rgba, r->image64_buf.ptr, (size_t)r->width * r->height * 32, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
}

void renderer_resolve(rm_renderer* r, float* rgba, size_t count) {
    RM_REQUIRE(count >= (size_t)r->width * r->height * 4, "output buffer too small");
    hipStream_t st = r->scene->dev->stream;
    // accumulating: the samples so far (normalizeImage.metal:45-46 divides by
    // frameIndex + 1, one sample per frame); otherwise the last call's samples,
    // which is 1 for the reference's one-sample frames
    const uint32_t n = r->u.accumulate ? r->frame_index : r->last_samples;
    const float divisor = (float)(n > 0 ? n : 1);
    RM_HIP(launch_resolve(r->image_buf.as<float>(), r->resolve_buf.as<float>(), r->width * r->height, r->u.exposure,
                          divisor, (int)r->u.tonemapping, st));
    RM_HIP(hipMemcpyAsync(rgba, r->resolve_buf.ptr, (size_t)r->width * r->height * 16, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
}

}  // namespace rm
