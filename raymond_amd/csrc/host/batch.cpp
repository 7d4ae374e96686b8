// Batched rendering of many jobs in one regenerating bounce loop.
//
// rm_batch_create merges the jobs' scenes into one device scene:
//   * BVH8 node and triangle arrays are concatenated; node child and triangle
//     bases and the triangles' instance ids are rebased, and each job keeps
//     its own subtree, entered at its root;
//   * vertices, normals, texcoords, indices, face materials and per-instance
//     data are concatenated (instance vertex/face offsets rebased);
//   * each job gets a LightsView: its scene's area/point/sun/spot lights,
//     light-face tables and world light as they are, and its shape lights
//     copied with their instance ids rebased;
//   * the batch image is the jobs' images one after another.
// The jobs' scenes must share one material library; the batch compiles it
// again with the batch shading kernel.  Scenes must outlive the batch (the
// world lights keep pointing at their envmaps).
//
// rm_batch_render walks the samples of all jobs in slices (one sample of one
// job), by default job after job in order of path count (shortest first: small
// jobs complete early, and fewer jobs share the caches at a time), or
// sample-major so that every job progresses from the start, and splits them
// into bounce loops small enough for the loop's counters.
// The renderer behind the batch runs those loops; the shading kernel counts
// ended paths per job, from which the host records each job's completion.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include <sys/stat.h>

#include "../kernels/trace_api.h"
#include "batch_api.h"
#include "scene_api.h"

namespace rm {

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
std::vector<T> download(const DevBuf& b, size_t count, hipStream_t st) {
    std::vector<T> out(count);
    if (count) RM_HIP(hipMemcpyAsync(out.data(), b.ptr, count * sizeof(T), hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    return out;
}

}  // namespace

void batch_create(rm_device* dev, const rm_job_desc* jobs, uint32_t n, rm_batch* b) {
    RM_REQUIRE(n > 0, "a batch needs at least one job");
    b->dev = dev;
    hipStream_t st = dev->stream;
    const std::string& src = jobs[0].scene->material_source;
    uint64_t pixels = 0;
    for (uint32_t j = 0; j < n; j++) {
        const rm_job_desc& d = jobs[j];
        RM_REQUIRE(d.scene != nullptr, "null job scene");
        RM_REQUIRE(d.scene->accel_built && d.scene->materials_built && d.scene->envmap_built,
                   "job scenes must have their acceleration structure, materials and envmap built");
        RM_REQUIRE(d.scene->material_source == src, "the jobs of a batch must share one material library");
        RM_REQUIRE(d.scene->two_level == jobs[0].scene->two_level,
                   "the jobs of a batch must all have the flat or all the two-level acceleration structure");
        RM_REQUIRE(d.width > 0 && d.height > 0 && d.samples > 0, "job size and samples must be positive");
        RM_REQUIRE(d.max_depth >= 1 && d.max_depth <= 255, "job max_depth out of range");
        pixels += (uint64_t)d.width * d.height;
        b->desc.push_back(d);
    }
    RM_REQUIRE(pixels < (1ull << 31), "batch image too large");

    rm_scene* m = &b->merged;
    m->dev = dev;
    std::vector<PerInstance> instances;
    std::vector<ShapeLight> shapes;
    std::vector<Bvh8Node> nodes;
    std::vector<BvhTri> tris;
    std::vector<TlasInstance> recs;  // two-level jobs: the instance records of every job's top-level tree
    const bool two_level = jobs[0].scene->two_level;
    std::vector<uint32_t> shape_off(n);
    uint32_t image_offset = 0;
    for (uint32_t j = 0; j < n; j++) {
        const rm_scene* s = jobs[j].scene;
        const uint32_t v_off = (uint32_t)(m->vertices.size() / 3);
        const uint32_t f_off = (uint32_t)(m->indices.size() / 3);
        const uint32_t i_off = (uint32_t)instances.size();
        const uint32_t node_off = (uint32_t)nodes.size();
        const uint32_t tri_off = (uint32_t)tris.size();
        m->vertices.insert(m->vertices.end(), s->vertices.begin(), s->vertices.end());
        m->normals.insert(m->normals.end(), s->normals.begin(), s->normals.end());
        m->texcoords.insert(m->texcoords.end(), s->texcoords.begin(), s->texcoords.end());
        m->indices.insert(m->indices.end(), s->indices.begin(), s->indices.end());
        m->materials.insert(m->materials.end(), s->materials.begin(), s->materials.end());
        for (PerInstance pi : s->instances) {
            pi.vertex_offset += v_off;
            pi.face_offset += f_off;
            instances.push_back(pi);
        }
        shape_off[j] = (uint32_t)shapes.size();
        for (ShapeLight L : s->shape) {
            L.instance_index += i_off;
            shapes.push_back(L);
        }
        std::vector<Bvh8Node> sn = download<Bvh8Node>(s->d_nodes, s->stats.bvh_nodes, st);
        std::vector<BvhTri> stt = download<BvhTri>(s->d_tris, s->d_tris.bytes / sizeof(BvhTri), st);
        // a two-level scene stores its shapes' trees first and the top-level
        // tree from root2l on; the top-level leaves index instance records
        const uint32_t rec_off = (uint32_t)recs.size();
        for (size_t k = 0; k < sn.size(); k++) {
            Bvh8Node& node = sn[k];
            node.child_base += node_off;
            node.tri_base += (two_level && k >= s->root2l) ? rec_off : tri_off;
        }
        for (BvhTri& t : stt) t.inst += i_off;
        nodes.insert(nodes.end(), sn.begin(), sn.end());
        tris.insert(tris.end(), stt.begin(), stt.end());
        if (two_level) {
            std::vector<TlasInstance> sr = download<TlasInstance>(s->d_tlas, s->d_tlas.bytes / sizeof(TlasInstance), st);
            for (TlasInstance& rec : sr) {
                rec.blas_root += node_off;
                rec.inst += i_off;
            }
            recs.insert(recs.end(), sr.begin(), sr.end());
        }

        JobInfo ji{};
        ji.image_offset = image_offset;
        ji.width = jobs[j].width;
        ji.height = jobs[j].height;
        ji.root = node_off + (two_level ? s->root2l : 0u);
        ji.max_depth = jobs[j].max_depth;
        ji.seed = jobs[j].seed ? jobs[j].seed : 1u;
        ji.spp = jobs[j].samples;
        ji.camera = s->view.camera;
        b->jobs.push_back(ji);
        b->job_paths.push_back((uint64_t)ji.width * ji.height * ji.spp);
        image_offset += ji.width * ji.height;
    }
    RM_REQUIRE(nodes.size() < (1ull << 31) && tris.size() < (1ull << 31), "batch geometry too large");

    m->d_vertices.upload(m->vertices.data(), m->vertices.size() * 4, st);
    m->d_normals.upload(m->normals.data(), m->normals.size() * 4, st);
    m->d_texcoords.upload(m->texcoords.data(), m->texcoords.size() * 4, st);
    m->d_indices.upload(m->indices.data(), m->indices.size() * 4, st);
    m->d_materials.upload(m->materials.data(), m->materials.size() * 2, st);
    m->d_instances.upload(instances.data(), instances.size() * sizeof(PerInstance), st);
    m->d_shape.upload(shapes.data(), shapes.size() * sizeof(ShapeLight), st);
    m->d_nodes.upload(nodes.data(), nodes.size() * sizeof(Bvh8Node), st);
    m->d_tris.upload(tris.data(), tris.size() * sizeof(BvhTri), st);
    if (two_level) {
        m->d_tlas.upload(recs.data(), recs.size() * sizeof(TlasInstance), st);
        m->two_level = true;
        m->stats.two_level = 1;
    }
    m->instances = instances;
    const int bpc = std::max(dev->trace_blocks_per_cu, dev->any_blocks_per_cu);
    m->spill_threads = (size_t)dev->props.multiProcessorCount * (size_t)bpc * kTraceBlock;
    m->d_spill.alloc(m->spill_threads * (size_t)(kMaxTraceStack - kLdsStack) * sizeof(uint2));
    m->stats.triangles = tris.size();
    m->stats.bvh_nodes = nodes.size();

    SceneView& v = m->view;
    v.vertices = m->d_vertices.as<float>();
    v.indices = m->d_indices.as<uint32_t>();
    v.normals = m->d_normals.as<float>();
    v.texcoords = m->d_texcoords.as<float>();
    v.instances = m->d_instances.as<PerInstance>();
    v.materials = m->d_materials.as<uint16_t>();
    v.camera = jobs[0].scene->view.camera;
    // the jobs share one material library, whose texture slots job 0's scene holds
    v.textures = jobs[0].scene->view.textures;
    v.num_textures = jobs[0].scene->view.num_textures;
    std::vector<LightsView> lights(n);
    for (uint32_t j = 0; j < n; j++) {
        lights[j] = jobs[j].scene->view.lights;
        lights[j].shape = m->d_shape.as<ShapeLight>() + shape_off[j];
    }
    v.lights = lights[0];
    m->envmap_res = jobs[0].scene->envmap_res;

    // the shared material library with the batch shading kernel
    std::string log;
    scene_compile_materials(m, src.data(), src.size(), "-DRM_BATCH_KERNEL", log);
    m->accel_built = true;
    m->envmap_built = true;

    b->d_jobs.upload(b->jobs.data(), b->jobs.size() * sizeof(JobInfo), st);
    b->d_lights.upload(lights.data(), lights.size() * sizeof(LightsView), st);
    b->d_done.alloc(n * 4);
    b->pixels = (uint32_t)pixels;
    uint32_t max_depth = 1;
    for (const JobInfo& ji : b->jobs) max_depth = std::max(max_depth, ji.max_depth);
    b->max_depth = max_depth;

    // slices, sample-major
    uint32_t max_spp = 0;
    for (const JobInfo& ji : b->jobs) max_spp = std::max(max_spp, ji.spp);
    for (uint32_t s = 0; s < max_spp; s++)
        for (uint32_t j = 0; j < n; j++)
            if (s < b->jobs[j].spp) b->slices.push_back(BatchSlice{0u, j, s, 0u});
    // shortest job first: by path count, ties by job index
    std::vector<uint32_t> by_size(n);
    for (uint32_t j = 0; j < n; j++) by_size[j] = j;
    std::stable_sort(by_size.begin(), by_size.end(), [&](uint32_t x, uint32_t y) {
        return b->job_paths[x] < b->job_paths[y];
    });
    for (uint32_t j : by_size)
        for (uint32_t s = 0; s < b->jobs[j].spp; s++) b->slices_sjf.push_back(BatchSlice{0u, j, s, 0u});
    RM_HIP(hipStreamSynchronize(st));
}

}  // namespace rm

namespace rm {

void batch_render(rm_batch* b, const rm_batch_config* cfg, rm_batch_stats* stats, size_t max_slices) {
    rm_batch_config c = cfg ? *cfg : rm_batch_config{};
    if (!c.queue_paths) c.queue_paths = 1u << 24;
    const size_t capacity = std::min<size_t>(c.queue_paths, 1u << 26);
    hipStream_t st = b->dev->stream;
    if (!b->renderer || b->renderer_capacity != capacity) {
        if (b->renderer)
            for (hipEvent_t e : b->renderer->events) (void)hipEventDestroy(e);
        b->renderer.reset(new rm_renderer());
        renderer_init(&b->merged, b->pixels, 1, b->max_depth, capacity, b->renderer.get());
        b->renderer_capacity = capacity;
    }
    rm_renderer* r = b->renderer.get();
    r->u.sampling_mode = RM_SAMPLING_MIS;
    r->u.accumulate = 1;
    r->u.rr_depth = c.rr_depth;
    r->u.material_sort = c.material_sort;
    r->u.timing = c.timing;
    r->u.use_graph = c.use_graph;
    r->u.trace_blocks_per_cu = c.trace_blocks_per_cu;
    r->u.sort_key_shift = c.sort_key_shift;
    r->u.traversal_stats = 0;
    renderer_reset(r);
    const uint32_t n = (uint32_t)b->jobs.size();
    RM_HIP(hipMemsetAsync(b->d_done.ptr, 0, n * 4, st));
    b->job_ms.assign(n, -1.0);

    RM_REQUIRE(c.order == RM_BATCH_SHORTEST_FIRST || c.order == RM_BATCH_SAMPLE_MAJOR, "unknown batch order");
    const std::vector<BatchSlice>& order = c.order == RM_BATCH_SHORTEST_FIRST ? b->slices_sjf : b->slices;
    const size_t slices = max_slices ? std::min(max_slices, order.size()) : order.size();
    // bounce loops of at most this many paths: within the 32-bit counters and,
    // at a few paths per queue slot and iteration, within the iteration limit
    const uint64_t loop_limit = std::min<uint64_t>((uint64_t)capacity * 192, (1ull << 31) - 1);
    if (!b->d_slices.ptr || b->d_slices.bytes < b->slices.size() * sizeof(BatchSlice))
        b->d_slices.alloc(b->slices.size() * sizeof(BatchSlice));
    std::vector<uint64_t> partial_paths(n, 0);
    rm_render_stats rs{};
    rm_batch_stats out{};
    const double t0 = now_ms();
    size_t first = 0;
    while (first < slices) {
        // slices [first, last) with stream-local firsts
        std::vector<BatchSlice> loop;
        uint64_t paths = 0;
        size_t last = first;
        while (last < slices) {
            const JobInfo& ji = b->jobs[order[last].job];
            const uint64_t sz = (uint64_t)ji.width * ji.height;
            if (!loop.empty() && paths + sz > loop_limit) break;
            BatchSlice sl = order[last];
            sl.first = (uint32_t)paths;
            loop.push_back(sl);
            paths += sz;
            partial_paths[sl.job] += sz;
            last++;
        }
        RM_HIP(hipMemcpyAsync(b->d_slices.ptr, loop.data(), loop.size() * sizeof(BatchSlice), hipMemcpyHostToDevice,
                              st));
        BatchRun br;
        br.slices = b->d_slices.as<BatchSlice>();
        br.num_slices = (uint32_t)loop.size();
        br.total = (uint32_t)paths;
        br.view.jobs = b->d_jobs.as<JobInfo>();
        br.view.lights = b->d_lights.as<LightsView>();
        br.view.num_jobs = n;
        br.view.done = b->d_done.as<uint32_t>();
        // a partial render completes a job when its rendered share is done
        br.job_paths = max_slices ? &partial_paths : &b->job_paths;
        br.job_ms = &b->job_ms;
        br.t0_ms = t0;
        rm_render_stats ls{};
        renderer_render_batch_loop(r, &br, &ls);
        rs.total_rays += ls.total_rays;
        rs.closest_total += ls.closest_total;
        rs.shadow_total += ls.shadow_total;
        rs.samples += ls.samples;
        rs.iterations += ls.iterations;
        rs.ms_trace_closest += ls.ms_trace_closest;
        rs.ms_shade += ls.ms_shade;
        rs.ms_trace_any += ls.ms_trace_any;
        rs.ms_sort += ls.ms_sort;
        out.loops++;
        first = last;
    }
    out.ms_total = now_ms() - t0;
    out.total_rays = rs.total_rays;
    out.closest_rays = rs.closest_total;
    out.shadow_rays = rs.shadow_total;
    out.samples = rs.samples;
    out.iterations = rs.iterations;
    out.ms_trace_closest = rs.ms_trace_closest;
    out.ms_shade = rs.ms_shade;
    out.ms_trace_any = rs.ms_trace_any;
    out.ms_sort = rs.ms_sort;
    double makespan = 0.0;
    for (double t : b->job_ms) makespan = std::max(makespan, t);
    out.makespan_ms = makespan;
    out.queue_paths = (uint32_t)capacity;
    if (stats) *stats = out;
}

void batch_read_image(rm_batch* b, uint32_t job, float* rgba, size_t count) {
    RM_REQUIRE(job < b->jobs.size(), "job index out of range");
    RM_REQUIRE(b->renderer != nullptr, "batch not rendered");
    const JobInfo& ji = b->jobs[job];
    const size_t px = (size_t)ji.width * ji.height;
    RM_REQUIRE(count >= px * 4, "output buffer too small");
    hipStream_t st = b->dev->stream;
    const float* img = b->renderer->image_buf.as<float>() + (size_t)ji.image_offset * 4;
    RM_HIP(hipMemcpyAsync(rgba, img, px * 16, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
}

namespace {

// Tuner cache: one line per (device, batch signature): queue_paths material_sort
// trace_blocks_per_cu sort_key_shift.
std::string tune_cache_path() {
    if (const char* e = getenv("RM_TUNE_CACHE")) return e;
    if (const char* h = getenv("HOME")) return std::string(h) + "/.cache/raymond_amd/batch_tune.txt";
    return "/tmp/raymond_amd_batch_tune.txt";
}

std::string batch_signature(const rm_batch* b) {
    uint64_t paths = 0, tris = b->merged.stats.triangles;
    for (uint64_t p : b->job_paths) paths += p;
    std::ostringstream os;
    os << b->dev->props.name << "/" << b->dev->props.multiProcessorCount << "cu/" << b->jobs.size() << "jobs/"
       << (uint64_t)std::log2((double)paths + 1.0) << "lp/" << (uint64_t)std::log2((double)tris + 1.0) << "lt";
    return os.str();
}

}  // namespace

void batch_autotune(rm_batch* b, rm_batch_config* best, rm_batch_tune_entry* entries, uint32_t capacity,
                    uint32_t* count, int* cache_hit) {
    RM_REQUIRE(best != nullptr, "null config");
    const std::string key = batch_signature(b);
    const std::string path = tune_cache_path();
    if (cache_hit) *cache_hit = 0;
    if (count) *count = 0;
    {
        std::ifstream f(path);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream is(line);
            std::string k;
            uint32_t q = 0, srt = 0, tb = 0, sh = 0;
            if (is >> k >> q >> srt && k == key && q > 0) {
                is >> tb >> sh;  // absent in lines of older versions: defaults
                best->queue_paths = q;
                best->material_sort = srt;
                best->trace_blocks_per_cu = tb;
                best->sort_key_shift = sh;
                if (cache_hit) *cache_hit = 1;
                return;
            }
        }
    }
    // calibration: the first samples of every job (a prefix of the
    // sample-major slices) rendered with each candidate; the best Mrays/s wins
    uint64_t pixels = 0;
    for (const JobInfo& ji : b->jobs) pixels += (uint64_t)ji.width * ji.height;
    size_t prefix = 0;
    uint64_t prefix_paths = 0;
    const uint64_t want = std::max<uint64_t>(pixels * 2, 1ull << 25);
    while (prefix < b->slices.size() && prefix_paths < want) {
        const JobInfo& ji = b->jobs[b->slices[prefix].job];
        prefix_paths += (uint64_t)ji.width * ji.height;
        prefix++;
    }
    uint32_t used = 0;
    double best_rate = -1.0;
    const rm_batch_config base = *best;
    auto trial = [&](rm_batch_config c) {
        c.timing = 0;
        c.order = RM_BATCH_SAMPLE_MAJOR;  // the calibration prefix mixes all jobs
        rm_batch_stats s{};
        batch_render(b, &c, &s, prefix);
        const double rate = s.total_rays / (s.ms_total * 1e-3) / 1e6;
        if (entries && used < capacity)
            entries[used] = rm_batch_tune_entry{c.queue_paths, c.material_sort, c.trace_blocks_per_cu, c.sort_key_shift,
                                                rate};
        used++;
        if (rate > best_rate) {
            best_rate = rate;
            *best = c;
        }
    };
    // stage 1: queue depth x material sort, at the default grid and bins
    const uint32_t queues[4] = {1u << 21, 1u << 22, 1u << 23, 1u << 24};
    for (uint32_t qi = 0; qi < 4; qi++)
        for (uint32_t srt = 0; srt < 2; srt++) {
            rm_batch_config c = base;
            c.queue_paths = queues[qi];
            c.material_sort = srt;
            c.trace_blocks_per_cu = 0;
            c.sort_key_shift = 0;
            trial(c);
        }
    // stage 2, from the best of stage 1: fewer persistent trace workgroups
    // per CU (less contention for the L2 among the waves of a CU), and
    // coarser sort bins (cheaper keys and scatter, less coherent shading)
    const rm_batch_config stage1 = *best;
    const int most = std::max(b->dev->trace_blocks_per_cu, b->dev->any_blocks_per_cu);
    std::vector<int> grids;
    for (int tb : {most - 1, (most + 1) / 2})
        if (tb >= 1 && tb < most && std::find(grids.begin(), grids.end(), tb) == grids.end()) grids.push_back(tb);
    for (int tb : grids) {
        rm_batch_config c = stage1;
        c.trace_blocks_per_cu = (uint32_t)tb;
        trial(c);
    }
    if (stage1.material_sort) {
        for (uint32_t sh : {1u, 3u}) {
            rm_batch_config c = *best;
            c.sort_key_shift = sh;
            trial(c);
        }
    }
    if (count) *count = std::min(used, capacity);
    best->timing = base.timing;
    best->order = base.order;
    // drop the calibration renders from the batch image and the job clocks
    renderer_reset(b->renderer.get());
    b->job_ms.assign(b->jobs.size(), -1.0);
    // create the cache directory and its parent if needed
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos && slash > 0) {
        const std::string dir = path.substr(0, slash);
        const size_t up = dir.rfind('/');
        if (up != std::string::npos && up > 0) mkdir(dir.substr(0, up).c_str(), 0755);
        mkdir(dir.c_str(), 0755);
    }
    std::ofstream f(path, std::ios::app);
    if (f)
        f << key << " " << best->queue_paths << " " << best->material_sort << " " << best->trace_blocks_per_cu << " "
          << best->sort_key_shift << "\n";
}

}  // namespace rm
