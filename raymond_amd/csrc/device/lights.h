// Light sampling for next-event estimation: device/lights/Lights.hpp:33-151,
// LightSample.hpp:6-34, WorldLight.hpp:14-62, AreaLight.metal:7-38,
// PointLight.metal:9-40, SunLight.metal:9-37, SpotLight.metal:9-54,
// ShapeLight.metal:10-66, Lights.metal:6-15.
#pragma once
#include "common.h"
#include "rng.h"
#include "warp.h"
#include "shading.h"

namespace rm {

struct LightSample {
    bool is_light = false;
    uint32_t shader_index = 0;
    bool can_be_hit = false;
    bool casts_shadows = false;
    vec3 weight;
    float pdf = 0.0f;
    vec3 direction;
    float distance = 0.0f;

    RM_HD static LightSample from_info(const LightInfo& info) {
        LightSample s;
        s.is_light = true;
        s.shader_index = info.shader_index;
        s.can_be_hit = info.uses_mis != 0;
        s.casts_shadows = info.casts_shadows != 0;
        return s;
    }
    RM_HD static LightSample invalid() {
        LightSample s;
        s.pdf = 0.0f;
        s.weight = vec3(0.0f);
        return s;
    }
};

// ---------------------------------------------------------------------------
// World light (device/lights/WorldLight.hpp:14-62)

RM_HD float world_pdf(const WorldLight& w, vec3 wo) {
    const vec2 sq = warp::uniform_sphere_to_square(wo);
    const uint32_t px = ((uint32_t)((float)w.resolution * sq.x)) % (uint32_t)w.resolution;
    const uint32_t py = ((uint32_t)((float)w.resolution * sq.y)) % (uint32_t)w.resolution;
    return w.pdfs[py * (uint32_t)w.resolution + px];
}

// Descends the quad-tree of normalised child probabilities.
RM_HD vec3 world_sample(const WorldLight& w, vec2 uv, float& pdf) {
    int cur_res = 1;
    int sx = 0, sy = 0;
    const float* level = w.mipmap;
    while (cur_res < w.resolution) {
        const int off = 4 * (sy * cur_res + sx);
        level += cur_res * cur_res;
        sx *= 2;
        sy *= 2;
        cur_res *= 2;
        const float top_left = level[off + 0];
        const float top_right = level[off + 1];
        const float bottom_left = level[off + 2];
        const float left_prob = top_left + bottom_left;
        float top_prob;
        if (uv.x < left_prob) {
            const float inv = 1.0f / left_prob;
            uv.x = uv.x * inv;
            top_prob = top_left * inv;
        } else {
            const float inv = 1.0f / (1.0f - left_prob);
            uv.x = (uv.x - left_prob) * inv;
            top_prob = top_right * inv;
            sx += 1;
        }
        if (uv.y < top_prob) {
            uv.y = uv.y / top_prob;
        } else {
            uv.y = (uv.y - top_prob) / (1.0f - top_prob);
            sy += 1;
        }
    }
    pdf = w.pdfs[sy * w.resolution + sx];
    uv = (vec2((float)sx, (float)sy) + uv) / (float)w.resolution;
    return warp::uniform_square_to_sphere(uv);
}

// ---------------------------------------------------------------------------
// Per-type samplers.  `shading` is the light's own shading context: on entry
// it holds the receiving point, on exit the light point's attributes.

RM_HD LightSample sample_area(const AreaLight& L, ShadingContext& shading, Prng& prng) {
    vec2 uv = prng.sample2d();
    if (L.is_circular) uv = warp::uniform_square_to_disk(uv) / 2.0f + 0.5f;
    const vec4 p4(uv.x - 0.5f, uv.y - 0.5f, 0.0f, 1.0f);
    auto dot4 = [](vec4 a, vec4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; };
    const vec3 point(dot4(p4, L.t0), dot4(p4, L.t1), dot4(p4, L.t2));
    const vec4 n4(0.0f, 0.0f, 1.0f, 0.0f);
    const vec3 normal = normalize(vec3(dot4(n4, L.t0), dot4(n4, L.t1), dot4(n4, L.t2)));

    LightSample s = LightSample::from_info(L.info);
    s.direction = point - shading.position;
    const float lensqr = length_squared(s.direction);
    s.distance = sqrt_(lensqr);
    shading.normal = -normal;
    shading.true_normal = -normal;
    shading.position = point;
    shading.generated = point;
    shading.uv = vec3(uv, 0.0f);
    const float G = saturate(dot(normal, s.direction)) / lensqr;
    s.direction = s.direction / s.distance;
    s.weight = (L.color * G) * 0.25f;
    s.pdf = 1.0f;
    return s;
}

RM_HD LightSample sample_point(const PointLight& L, ShadingContext& shading, Prng& prng) {
    const vec3 lightN = normalize(shading.position - L.location);
    vec3 point = L.location;
    if (L.radius > 0.0f) {
        const mat3 basis = build_onb(lightN);
        point = point + L.radius * (basis * vec3(warp::uniform_square_to_disk(prng.sample2d()), 0.0f));
    }
    LightSample s = LightSample::from_info(L.info);
    s.direction = point - shading.position;
    const float lensqr = length_squared(s.direction);
    s.distance = sqrt_(lensqr);
    s.direction = s.direction / s.distance;
    shading.normal = -s.direction;
    shading.true_normal = -s.direction;
    shading.position = point;
    shading.generated = point;
    shading.uv = vec3(warp::uniform_sphere_to_square(-s.direction), 0.0f);
    const float G = 1.0f / lensqr;
    s.weight = (L.color * G) * (kInvPi * 0.25f);
    s.pdf = 1.0f;
    return s;
}

RM_HD LightSample sample_sun(const SunLight& L, ShadingContext& shading, Prng& prng) {
    const vec2 rnd = prng.sample2d();
    const float cos_theta = 1.0f - rnd.y * (1.0f - L.cos_angle);
    const float sin_theta = sqrt_(saturate(1.0f - cos_theta * cos_theta));
    float cos_phi;
    const float sin_phi = sincos_(2.0f * kPi * rnd.x, cos_phi);
    const mat3 fr = build_onb(L.direction);
    const vec3 point = fr * vec3(sin_theta * sin_phi, sin_theta * cos_phi, cos_theta);
    LightSample s = LightSample::from_info(L.info);
    s.direction = point;
    s.distance = kInf;
    shading.normal = -s.direction;
    shading.true_normal = -s.direction;
    shading.position = -s.direction;
    shading.generated = -s.direction;
    shading.uv = vec3(warp::uniform_sphere_to_square(-s.direction), 0.0f);
    s.weight = L.color;
    s.pdf = 1.0f;
    return s;
}

RM_HD float smoothstep_(float e0, float e1, float x) {
    const float t = saturate((x - e0) / (e1 - e0));
    return t * t * (3.0f - 2.0f * t);
}

// (device/lights/SpotLight.metal:9-21)
RM_HD float spot_attenuation(vec3 dir, float spot_angle, float spot_smooth, vec3 N) {
    float att = dot(dir, N);
    if (att <= spot_angle) {
        att = 0.0f;
    } else {
        const float t = att - spot_angle;
        if (t < spot_smooth && spot_smooth != 0.0f) att = att * smoothstep_(0.0f, spot_smooth, t);
    }
    return att;
}

RM_HD LightSample sample_spot(const SpotLight& L, ShadingContext& shading, Prng& prng) {
    const vec3 lightN = normalize(shading.position - L.location);
    vec3 point = L.location;
    if (L.radius > 0.0f) {
        const mat3 basis = build_onb(lightN);
        point = point + L.radius * (basis * vec3(warp::uniform_square_to_disk(prng.sample2d()), 0.0f));
    }
    LightSample s = LightSample::from_info(L.info);
    s.direction = point - shading.position;
    const float lensqr = length_squared(s.direction);
    s.distance = sqrt_(lensqr);
    s.direction = s.direction / s.distance;
    shading.normal = -s.direction;
    shading.true_normal = -s.direction;
    shading.position = point;
    shading.generated = point;
    shading.uv = vec3(warp::uniform_sphere_to_square(-s.direction), 0.0f);
    const float G = 1.0f / lensqr;
    const float att = spot_attenuation(L.direction, L.spot_size, L.spot_blend, s.direction);
    s.weight = ((att * L.color) * G) * (kInvPi * 0.25f);
    s.pdf = 1.0f;
    return s;
}

// Binary search for the first CDF entry above `value` (ShapeLight.metal:14-28)
RM_HD int upper_bound_cdf(const float* buf, int count, float value) {
    int left = 0, right = count;
    while (left < right) {
        const int cand = (left + right) / 2;
        if (value >= buf[cand]) left = cand + 1;
        else right = cand;
    }
    return left;
}

// Solid-angle pdf of hitting an emissive instance (ShapeLight.metal:10-12)
RM_HD float shape_pdf_single(const ShapeLight& L, const ShadingContext& shading) {
    return 1.0f / (shading.geometry_term() * L.emissive_area);
}

RM_HD LightSample sample_shape(const SceneView& ctx, const ShapeLight& L, ShadingContext& shading, Prng& prng) {
    const PerInstance& inst = ctx.instances[L.instance_index];
    const uint32_t prim = (uint32_t)upper_bound_cdf(ctx.lights.light_faces + inst.light_face_offset,
                                                    (int)inst.light_face_count, prng.sample());
    const vec2 bary = warp::uniform_square_to_triangle_barycentric(prng.sample2d());
    LightSample s;
    s.can_be_hit = true;
    s.casts_shadows = true;
    const vec3 origin = shading.position;
    shading.build(ctx, inst, prim, bary, 0.0f, s.shader_index);
    s.direction = shading.position - origin;
    const float lensqr = length_squared(s.direction);
    s.distance = sqrt_(lensqr);
    s.direction = s.direction / s.distance;
    shading.wo = -s.direction;
    shading.distance = s.distance;
    const float G = shading.geometry_term();
    s.weight = vec3(G * L.emissive_area);
    s.pdf = 1.0f / (G * L.emissive_area);
    return s;
}

// ---------------------------------------------------------------------------
// Light selection (device/lights/Lights.hpp:67-151)

RM_HD float lights_shape_pdf(const SceneView& ctx, const PerInstance& inst, const ShadingContext& shading) {
    const float sel = 1.0f / (float)ctx.lights.num_total;
    return sel * shape_pdf_single(ctx.lights.shape[inst.light_index], shading);
}

RM_HD float lights_envmap_pdf(const SceneView& ctx, vec3 wo) {
    const float sel = 1.0f / (float)ctx.lights.num_total;
    return sel * world_pdf(ctx.lights.world, wo);
}

RM_HD LightSample sample_envmap(const SceneView& ctx, ShadingContext& shading, Prng& prng) {
    LightSample s;
    s.direction = world_sample(ctx.lights.world, prng.sample2d(), s.pdf);
    s.distance = kInf;
    s.shader_index = ctx.lights.world.shader_index;
    s.casts_shadows = true;
    s.can_be_hit = true;
    s.is_light = true;
    shading.position = -s.direction;
    shading.normal = -s.direction;
    shading.true_normal = -s.direction;
    shading.generated = s.direction;
    shading.object = s.direction;
    shading.uv = vec3(0.0f);
    s.weight = vec3(1.0f / s.pdf);
    return s;
}

// Uniform light pick, per-type sample, emission through the light's material
// and a throughput-based light roulette (device/lights/Lights.hpp:77-127).
RM_HD LightSample lights_sample(const SceneView& ctx, const ShadingContext& shading, Prng& prng) {
    const LightsView& Ls = ctx.lights;
    int k = prng.sample_int(Ls.num_total);
    LightSample s;
    ShadingContext ls;
    ls.ray_flags = shading.ray_flags;
    ls.position = shading.position;
    if (k == 0) {
        s = sample_envmap(ctx, ls, prng);
    } else if ((k -= 1) < Ls.num_area) {
        s = sample_area(Ls.area[k], ls, prng);
    } else if ((k -= Ls.num_area) < Ls.num_point) {
        s = sample_point(Ls.point[k], ls, prng);
    } else if ((k -= Ls.num_point) < Ls.num_sun) {
        s = sample_sun(Ls.sun[k], ls, prng);
    } else if ((k -= Ls.num_sun) < Ls.num_spot) {
        s = sample_spot(Ls.spot[k], ls, prng);
    } else if ((k -= Ls.num_spot) < Ls.num_shape) {
        s = sample_shape(ctx, Ls.shape[k], ls, prng);
    } else {
        return LightSample::invalid();
    }
    ls.wo = -s.direction;
    if (any_nonzero(s.weight)) {
        if (s.is_light) shade_light(s.shader_index, ctx, ls);
        else shade_surface(s.shader_index, ctx, ls);
        s.weight *= ls.material.emission;
    }
    s.weight = s.weight * (float)Ls.num_total;
    s.pdf = s.pdf / (float)Ls.num_total;
    const float survival = saturate(4.0f * mean(s.weight));
    if (survival < 1.0f) {
        if (prng.sample() < survival) s.weight = s.weight / survival;
        else s.weight = vec3(0.0f);
    }
    return s;
}

// World shader on a miss (device/lights/Lights.metal:6-15)
RM_HD void evaluate_environment(const SceneView& ctx, ShadingContext& shading) {
    shading.position = shading.wo;
    shading.normal = shading.wo;
    shading.true_normal = shading.wo;
    shading.generated = -shading.wo;
    shading.object = -shading.wo;
    shading.uv = vec3(0.0f);
    shade_light(ctx.lights.world.shader_index, ctx, shading);
}

}  // namespace rm
