// Kernels compiled at run time together with a scene's generated material
// code (hiprtc).  Included after the material functions, so shade_surface /
// shade_light resolve to the scene's switch tables.
//
//  rm_shade_kernel         handleIntersections (device/kernels/shading/intersection.metal:22-211)
//  rm_envmap_build_kernel  buildEnvironmentMap (device/kernels/envmap/build.metal:4-51)
#pragma once
#include "common.h"
#include "rng.h"
#include "warp.h"
#include "bsdf.h"
#include "scene.h"
#include "shading.h"
#include "lights.h"
#include "queues.h"
#include "shade_kernels_abi.h"

namespace rm {

// power heuristic (device/kernels/shading/intersection.metal:14-20)
__device__ inline float mis_weight(float pdf, float other) {
    pdf = pdf * pdf;
    if (isinf_(pdf)) return 1.0f;
    other = other * other;
    return pdf / (pdf + other);
}

// Cycles' shading-normal fix (device/utils/math.hpp:73-119)
__device__ inline vec3 ensure_valid_reflection(vec3 Ng, vec3 I, vec3 N) {
    const vec3 R = (2.0f * dot(N, I)) * N - I;
    const float threshold = min_(0.9f * dot(Ng, I), 0.01f);
    if (dot(Ng, R) >= threshold) return N;
    const float NdotNg = dot(N, Ng);
    const vec3 X = normalize(N - NdotNg * Ng);
    const float Ix = dot(I, X), Iz = dot(I, Ng);
    const float Ix2 = sqr(Ix), Iz2 = sqr(Iz);
    const float a = Ix2 + Iz2;
    const float b = safe_sqrtf(Ix2 * (a - sqr(threshold)));
    const float c = Iz * threshold + a;
    const float fac = 0.5f / a;
    const float N1_z2 = fac * (b + c), N2_z2 = fac * (-b + c);
    bool valid1 = (N1_z2 > 1e-5f) && (N1_z2 <= (1.0f + 1e-5f));
    bool valid2 = (N2_z2 > 1e-5f) && (N2_z2 <= (1.0f + 1e-5f));
    vec2 N_new;
    if (valid1 && valid2) {
        const vec2 N1(safe_sqrtf(1.0f - N1_z2), safe_sqrtf(N1_z2));
        const vec2 N2(safe_sqrtf(1.0f - N2_z2), safe_sqrtf(N2_z2));
        const float R1 = (2.0f * (N1.x * Ix + N1.y * Iz)) * N1.y - Iz;
        const float R2 = (2.0f * (N2.x * Ix + N2.y * Iz)) * N2.y - Iz;
        valid1 = (R1 >= 1e-5f);
        valid2 = (R2 >= 1e-5f);
        if (valid1 && valid2) N_new = (R1 < R2) ? N1 : N2;
        else N_new = (R1 > R2) ? N1 : N2;
    } else if (valid1 || valid2) {
        const float Nz2 = valid1 ? N1_z2 : N2_z2;
        N_new = vec2(safe_sqrtf(1.0f - Nz2), safe_sqrtf(Nz2));
    } else {
        return Ng;
    }
    return N_new.x * X + N_new.y * Ng;
}

__device__ inline void image_add(float* image, uint32_t pixel, vec3 c) {
    float* p = image + 4u * (size_t)pixel;
    atomicAdd(p + 0, c.x);
    atomicAdd(p + 1, c.y);
    atomicAdd(p + 2, c.z);
    atomicAdd(p + 3, 1.0f);
}

constexpr uint32_t kShadeBlock = 256;
constexpr uint32_t kShadeWaves = kShadeBlock / 64;

// Workgroup-level compaction of N append streams: ballots give each lane's
// rank inside its wave, LDS prefix sums give the wave offsets, and one atomic
// per stream and workgroup reserves the slots (slots in lane order).  Must be
// reached by every thread of the workgroup.
template <int N>
__device__ inline void block_append(uint32_t* const (&ctr)[N], const bool (&want)[N], uint32_t (&slot)[N]) {
    __shared__ uint32_t wave_cnt[N][kShadeWaves];
    __shared__ uint32_t base[N];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64u;
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long m[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        m[k] = __ballot(want[k]);
        if (lane == 0) wave_cnt[k][wave] = (uint32_t)__popcll(m[k]);
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)N) {
        const int k = (int)threadIdx.x;
        uint32_t t = 0;
        for (uint32_t w = 0; w < kShadeWaves; w++) {
            const uint32_t c = wave_cnt[k][w];
            wave_cnt[k][w] = t;
            t += c;
        }
        base[k] = t ? atomicAdd(ctr[k], t) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) slot[k] = base[k] + wave_cnt[k][wave] + (uint32_t)__popcll(m[k] & lt);
    __syncthreads();
}

// Radiance-cache inputs of a shaded vertex.
__device__ inline void cache_features(const CacheView& cv, const ShadingContext& sh, vec3 n, float* f) {
    const vec3 p = (sh.position - cv.bounds_min) * cv.bounds_inv_extent;
    f[0] = saturate(p.x);
    f[1] = saturate(p.y);
    f[2] = saturate(p.z);
    const UberShader& m = sh.material;
    const float spec = m.lobe_prob[1];
    const float rough = spec > 0.0f ? sqrt_(sqrt_(m.specular.alpha_x * m.specular.alpha_y)) : 1.0f;
    rc_vertex_features(sh.wo, n, rough, m.albedo(), spec, f + 3);
}

// One path vertex (device/kernels/shading/intersection.metal:43-208).
// Image contributions are added here; the shadow and continuation rays are
// returned for compacted append by the caller.
struct ShadeOut {
    bool shadow = false;
    vec3 s_origin, s_dir, s_weight;
    float s_tmax = 0.0f;
    bool next = false;
    vec3 n_origin, n_dir, n_weight;
    float n_pdf = 0.0f;
    uint32_t n_meta = 0;
    uint32_t seed = 0;
    uint32_t pixel = 0;
    // radiance cache: 1 = lookup (the path ends here), 2 = training vertex
    uint32_t cache_kind = 0;
    uint32_t slot = kRcNoSlot;
    uint32_t job = 0;
    vec3 c_t, c_r;
    float feat[kRcFeatureFloats];
};

template <bool BATCHED, bool CACHED>
__device__ inline __attribute__((always_inline)) void shade_path(const ShadeArgs& a, uint32_t i, ShadeOut& out) {
    // batched jobs: the path's job supplies the lights and the depth limit
    SceneView job_ctx = a.ctx;
    uint32_t max_depth = a.u.max_depth;
    if (BATCHED) {
        out.job = job_of_pixel(a.batch.jobs, a.batch.num_jobs, a.cur.pixel[i]);
        job_ctx.lights = a.batch.lights[out.job];
        max_depth = a.batch.jobs[out.job].max_depth;
    }
    const SceneView& ctx = job_ctx;
    const uint32_t mode = a.u.sampling_mode;
    const vec3 dir(a.cur.dx[i], a.cur.dy[i], a.cur.dz[i]);
    const vec3 ray_weight(a.cur.wx[i], a.cur.wy[i], a.cur.wz[i]);
    const float ray_pdf = a.cur.pdf[i];
    const uint32_t meta = a.cur.meta[i];
    const uint32_t flags = meta_flags(meta);
    const uint32_t depth = meta_depth(meta);
    const uint32_t pixel = a.cur.pixel[i];
    Prng prng = Prng::from_state(a.cur.seed[i], (uint16_t)meta_index(meta));
    out.pixel = pixel;
    const bool collect_emission = isinf_(ray_pdf) || mode != 1u;
    const float hit_t = a.hits.t[i];

    // a training path of the radiance cache also sums its own contributions
    const CacheView& cv = a.cache;
    const uint32_t slot = (CACHED && cv.slots) ? cv.slots[i] : kRcNoSlot;
    const bool training = slot != kRcNoSlot;
    float* slot_rad = training ? cv.slot_radiance + 3 * (size_t)(slot & ~kRcRecorded) : nullptr;
    vec3 rad(0.0f);
    if (training) rad = vec3(slot_rad[0], slot_rad[1], slot_rad[2]);
    out.slot = slot;
    // a camera path whose lens trace failed carries no ray (generate.metal:143-146)
    if (depth == kVoidDepth) return;
    auto contribute = [&](vec3 c) {
        image_add(a.image, pixel, c);
        if (training) rad += c;
    };
    auto store_radiance = [&]() {
        if (training) {
            slot_rad[0] = rad.x;
            slot_rad[1] = rad.y;
            slot_rad[2] = rad.z;
        }
    };

    ShadingContext shading;
    shading.ray_flags = flags;
    shading.rnd = prng.sample3d();
    shading.wo = -dir;

    if (hit_t <= 0.0f) {
        if (collect_emission) {
            const float mis = (mode == 0u || isinf_(ray_pdf)) ? 1.0f : mis_weight(ray_pdf, lights_envmap_pdf(ctx, dir));
            evaluate_environment(ctx, shading);
            contribute((mis * ray_weight) * shading.material.emission);
            store_radiance();
        }
        return;
    }

    const uint32_t inst_idx = a.hits.inst[i];
    const PerInstance& inst = ctx.instances[inst_idx];
    uint32_t shader_index;
    shading.build(ctx, inst, a.hits.prim[i], vec2(a.hits.b0[i], a.hits.b1[i]), hit_t, shader_index);
    if (inst.visibility & flags) shade_surface(shader_index, ctx, shading);
    else shading.material.alpha = 0.0f;

    if (collect_emission && mean(shading.material.emission) != 0.0f) {
        const float mis = (mode == 0u || isinf_(ray_pdf)) ? 1.0f
                                                           : mis_weight(ray_pdf, lights_shape_pdf(ctx, inst, shading));
        contribute((mis * ray_weight) * shading.material.emission);
    }

    vec3 sh_normal;
    if (all_zero(shading.material.normal)) {
        sh_normal = shading.true_normal;
    } else {
        vec3 geo = shading.true_normal;
        if (dot(geo, shading.wo) < 0.0f) geo = geo * -1.0f;
        sh_normal = ensure_valid_reflection(geo, shading.wo, shading.material.normal);
    }

    // Radiance cache: the first vertex after a diffuse bounce ends an
    // ordinary path with a lookup; a training path records that vertex (its
    // target is the radiance the rest of the path gathers) and goes on.
    if (CACHED && cv.enabled && depth >= 1u && (flags & RayDiffuse) && shading.material.alpha >= 1.0f &&
        !shading.material.is_delta()) {
        if (!training) {
            out.cache_kind = 1u;
            out.c_t = ray_weight;
            cache_features(cv, shading, sh_normal, out.feat);
            return;
        }
        if (!(slot & kRcRecorded)) {
            out.cache_kind = 2u;
            out.c_t = ray_weight;
            out.c_r = rad;
            cache_features(cv, shading, sh_normal, out.feat);
            out.slot = slot | kRcRecorded;
        }
    }

    // the last depth's shadow and continuation rays are never traced
    // (Renderer.swift:623-625), so they are not emitted
    const bool last = depth + 1u >= max_depth;
    if (mode != 0u) {
        const LightSample nee = lights_sample(ctx, shading, prng);
        float bsdf_pdf;
        const vec3 bsdf = shading.material.evaluate(shading.wo, nee.direction, sh_normal, shading.true_normal, bsdf_pdf);
        const vec3 contribution = (nee.weight * bsdf) * ray_weight;
        if (nee.casts_shadows) {
            const float mis = (mode == 1u || !nee.can_be_hit) ? 1.0f : mis_weight(nee.pdf, bsdf_pdf);
            const vec3 nee_weight = mis * contribution;
            if (all_finite(nee_weight) && any_nonzero(nee_weight) && !last) {
                out.shadow = true;
                out.s_origin = shading.position;
                out.s_dir = nee.direction;
                out.s_tmax = nee.distance - kEps;
                out.s_weight = nee_weight;
            }
        } else {
            // the reference applies the path weight twice here (intersection.metal:155,176)
            contribute(ray_weight * contribution);
        }
    }
    store_radiance();

    const BsdfSample s = shading.material.sample(shading.rnd, -dir, sh_normal, shading.true_normal, flags);
    const vec3 w = ray_weight * s.weight;
    const float mean_w = mean(w);
    if (!isfinite_(mean_w)) return;
    float survival;
    bool survive;
    if (depth < a.u.rr_depth) {
        survival = 1.0f;
        survive = mean_w > 0.0f;
    } else {
        survival = min_(mean_w, 1.0f);
        survive = prng.sample() < survival;
    }
    if (survive && !last) {
        out.next = true;
        out.n_origin = shading.position;
        out.n_dir = s.wi;
        out.n_weight = w * (1.0f / survival);
        out.n_pdf = s.pdf;
        out.n_meta = pack_meta(prng.index, s.flags, depth + 1);
        out.seed = prng.seed;
    }
}

// Batched jobs: ended paths per job, counted in LDS and added to the job
// counters once per workgroup (jobs beyond the LDS table count directly).
constexpr uint32_t kJobsLds = 256;

template <bool BATCHED, bool CACHED>
__device__ inline __attribute__((always_inline)) void shade_queue(const ShadeArgs& a) {
    __shared__ uint32_t job_ended[kJobsLds];
    constexpr bool batched = BATCHED;
    if (batched) {
        for (uint32_t j = threadIdx.x; j < kJobsLds; j += kShadeBlock) job_ended[j] = 0u;
        __syncthreads();
    }
    const uint32_t iter = iteration_of(a.iter_dev, a.iter);
    const uint32_t count = a.counters->it[iter].queue;
    const uint32_t stride = gridDim.x * kShadeBlock;
    for (uint32_t base = blockIdx.x * kShadeBlock; base < count; base += stride) {
        const uint32_t k = base + threadIdx.x;
        ShadeOut out;
        if (k < count) shade_path<BATCHED, CACHED>(a, a.order ? a.order[k] : k, out);
        if (batched && k < count && !out.next) {
            if (out.job < kJobsLds) atomicAdd(&job_ended[out.job], 1u);
            else atomicAdd(&a.batch.done[out.job], 1u);
        }
        constexpr int kStreams = CACHED ? 4 : 2;
        uint32_t* const ctr[4] = {&a.counters->it[iter].shadows, &a.counters->it[iter + 1].appended,
                                  &a.counters->cache_queries, &a.counters->train_records};
        const bool want[4] = {out.shadow, out.next, out.cache_kind == 1u, out.cache_kind == 2u};
        uint32_t at[4] = {0u, 0u, 0u, 0u};
        block_append<kStreams>(reinterpret_cast<uint32_t* const(&)[kStreams]>(ctr),
                               reinterpret_cast<const bool(&)[kStreams]>(want),
                               reinterpret_cast<uint32_t(&)[kStreams]>(at));
        const uint32_t sslot = at[0], nslot = at[1];
        if (CACHED && out.cache_kind == 1u && at[2] < a.cache.query_capacity) {
            typedef float row4 __attribute__((ext_vector_type(4)));
            row4* row = reinterpret_cast<row4*>(a.cache.queries.rows + (size_t)at[2] * kRcQueryFloats);
            row4 r0;
            r0[0] = __builtin_bit_cast(float, out.pixel);
            r0[1] = out.c_t.x;
            r0[2] = out.c_t.y;
            r0[3] = out.c_t.z;
            row[0] = r0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                row4 v;
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = out.feat[4 * c + k];
                row[1 + c] = v;
            }
        } else if (CACHED && out.cache_kind == 2u && at[3] < a.cache.train_capacity) {
            const RcTrainQueue& q = a.cache.train;
            const uint32_t j = at[3];
            q.slot[j] = out.slot & ~kRcRecorded;
            q.tx[j] = out.c_t.x;
            q.ty[j] = out.c_t.y;
            q.tz[j] = out.c_t.z;
            q.rx[j] = out.c_r.x;
            q.ry[j] = out.c_r.y;
            q.rz[j] = out.c_r.z;
#pragma unroll
            for (int f = 0; f < kRcFeatureFloats; f++) q.feat[(size_t)f * a.cache.train_capacity + j] = out.feat[f];
        }
        if (out.shadow) {
            a.shadows.ox[sslot] = out.s_origin.x;
            a.shadows.oy[sslot] = out.s_origin.y;
            a.shadows.oz[sslot] = out.s_origin.z;
            a.shadows.dx[sslot] = out.s_dir.x;
            a.shadows.dy[sslot] = out.s_dir.y;
            a.shadows.dz[sslot] = out.s_dir.z;
            a.shadows.tmax[sslot] = out.s_tmax;
            a.shadows.wx[sslot] = out.s_weight.x;
            a.shadows.wy[sslot] = out.s_weight.y;
            a.shadows.wz[sslot] = out.s_weight.z;
            a.shadows.pixel[sslot] = out.pixel;
            if (CACHED && a.shadows.slot) a.shadows.slot[sslot] = out.slot;
        }
        if (out.next) {
            a.next.ox[nslot] = out.n_origin.x;
            a.next.oy[nslot] = out.n_origin.y;
            a.next.oz[nslot] = out.n_origin.z;
            a.next.dx[nslot] = out.n_dir.x;
            a.next.dy[nslot] = out.n_dir.y;
            a.next.dz[nslot] = out.n_dir.z;
            a.next.wx[nslot] = out.n_weight.x;
            a.next.wy[nslot] = out.n_weight.y;
            a.next.wz[nslot] = out.n_weight.z;
            a.next.pdf[nslot] = out.n_pdf;
            a.next.seed[nslot] = out.seed;
            a.next.meta[nslot] = out.n_meta;
            a.next.pixel[nslot] = out.pixel;
            if (CACHED && a.cache.slots_next) a.cache.slots_next[nslot] = out.slot;
        }
    }
    if (batched) {
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < kJobsLds && j < a.batch.num_jobs; j += kShadeBlock)
            if (job_ended[j]) atomicAdd(&a.batch.done[j], job_ended[j]);
    }
}

// Waves per SIMD the shading kernels are held to.  Unbounded, large inlined
// graph sets take ~250 VGPRs (one or two waves); the spills a bound forces
// cost less than the latency it hides.  The plain kernel runs best at 5
// (32 random 20-60 node graphs at 1080p: 1803 -> 1826 Mrays/s against 4;
// sphere field 2772 -> 2798; Cornell box 4719 -> 4827; 6 waves lose on all
// three, e.g. 2831 -> 2783 on the sphere field), the batch kernel,
// which carries per-job state, at 4 (3605 against 3355 Mrays/s at 5); the
// cache kernel keeps 4 with it (at 5 the cache comparison's equal-time ratio
// went from 0.575-0.583 to 0.604-0.607).
constexpr int kShadeMinWaves = 5;
constexpr int kShadeWideMinWaves = 4;

extern "C" __global__ __launch_bounds__(kShadeBlock, kShadeMinWaves) void rm_shade_kernel(ShadeArgs a) {
    shade_queue<false, false>(a);
}

#if defined(RM_CACHE_KERNEL)
// Renderers with a radiance cache: lookups, training paths and their
// radiance sums.  Compiled only into the modules of scenes a cache renders.
extern "C" __global__ __launch_bounds__(kShadeBlock, kShadeWideMinWaves) void rm_shade_cache_kernel(ShadeArgs a) {
    shade_queue<false, true>(a);
}
#endif

#if defined(RM_BATCH_KERNEL)
// Batched jobs (a.batch.jobs non-null): per-job lights, depth and path counts.
// Compiled only into the modules of batches.
extern "C" __global__ __launch_bounds__(kShadeBlock, kShadeWideMinWaves) void rm_shade_batch_kernel(ShadeArgs a) {
    shade_queue<true, false>(a);
}
#endif

// 64 world-shader samples per texel; writes the finest mip level in 2x2-quad
// order and the unnormalised pdf (device/kernels/envmap/build.metal:4-51).
extern "C" __global__ __launch_bounds__(256) void rm_envmap_build_kernel(EnvmapArgs a) {
    const int res = a.resolution;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (uint32_t)(res * res)) return;
    const uint32_t x = tid % (uint32_t)res, y = tid / (uint32_t)res;
    const uint32_t ray_index = y * (uint32_t)res + x;
    float value = 0.0f;
    for (int s = 0; s < 64; s++) {
        Prng prng((uint32_t)s, ray_index);
        const vec2 j = prng.sample2d();
        const vec2 projected((( float)x + j.x) / (float)res, ((float)y + j.y) / (float)res);
        const vec3 wo = warp::uniform_square_to_sphere(projected);
        ShadingContext shading;
        shading.ray_flags = 0;
        shading.rnd = prng.sample3d();
        shading.wo = -wo;
        evaluate_environment(a.ctx, shading);
        value += mean(shading.material.emission);
    }
    value /= 64.0f;
    value += 1e-8f;
    const uint32_t quad = (y / 2u) * ((uint32_t)res / 2u) + (x / 2u);
    a.mipmap_last[4u * quad + (x & 1u) + 2u * (y & 1u)] = value;
    a.pdfs[y * (uint32_t)res + x] = value;
}

}  // namespace rm
