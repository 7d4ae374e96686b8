// Hit-point reconstruction (device/ShadingContext.hpp:7-32,
// device/ShadingContext.metal:4-47) and the material hook declarations
// (device/shading.hpp:5-12).
#pragma once
#include "common.h"
#include "bsdf.h"
#include "scene.h"

namespace rm {

struct ShadingContext {
    vec3 uv;
    vec3 position;
    vec3 generated;
    vec3 object;
    vec3 normal;
    vec3 true_normal;
    vec3 tu, tv;
    vec3 rnd;
    vec3 wo;  // pointing away from the hit point
    float distance = 0.0f;
    uint32_t ray_flags = 0;
    UberShader material;

    RM_HD float geometry_term() const { return abs_(dot(wo, true_normal)) / square(distance); }

    // interpolate(a, b, c, (x, y)) = a x + b y + c (1 - x - y) (device/utils/math.hpp:121-128)
    RM_HD static vec3 interpolate(vec3 a, vec3 b, vec3 c, vec2 bary) {
        const float u = bary.x, v = bary.y;
        const float w = (1.0f - u) - v;
        return (a * u + b * v) + c * w;
    }

    // Rebuilds position, normals, uv, tangents and the material index of a
    // hit (device/ShadingContext.metal:4-47).
    RM_HD void build(const SceneView& ctx, const PerInstance& inst, uint32_t prim, vec2 bary, float dist,
                     uint32_t& shader_index) {
        const uint32_t face = inst.face_offset + prim;
        const uint32_t i0 = inst.vertex_offset + ctx.indices[3 * face + 0];
        const uint32_t i1 = inst.vertex_offset + ctx.indices[3 * face + 1];
        const uint32_t i2 = inst.vertex_offset + ctx.indices[3 * face + 2];

        const vec2 Tc = ctx.texcoord(i2);
        const vec2 T0 = ctx.texcoord(i0) - Tc;
        const vec2 T1 = ctx.texcoord(i1) - Tc;
        uv = vec3((T0 * bary.x + T1 * bary.y) + Tc, 0.0f);

        const vec3 Pc = ctx.vertex(i2);
        const vec3 P0 = ctx.vertex(i0) - Pc;
        const vec3 P1 = ctx.vertex(i1) - Pc;
        true_normal = normalize(inst.normal_transform * cross(P0, P1));

        const vec3 local_p = (P0 * bary.x + P1 * bary.y) + Pc;
        object = local_p;
        generated = safe_divide(local_p - inst.bounds_min, inst.bounds_size, 0.5f);
        position = inst.point_transform.point(local_p);

        normal = inst.normal_transform * interpolate(ctx.normal(i0), ctx.normal(i1), ctx.normal(i2), bary);
        normal = normalize(normal);

        // T[1][1] = T1.y, T[0][1] = T0.y, T[1][0] = T1.x, T[0][0] = T0.x
        tu = normalize(inst.normal_transform * (P0 * T1.y + P1 * (-T0.y)));
        tv = normalize(inst.normal_transform * (P0 * (-T1.x) + P1 * T0.x));

        distance = dist;
        shader_index = ctx.materials[face];
    }
};

// Material entry points, defined by the generated material module
// (device/shading.hpp:11-12).
RM_HD void shade_surface(uint32_t shader_index, const SceneView& ctx, ShadingContext& shading);
RM_HD void shade_light(uint32_t shader_index, const SceneView& ctx, ShadingContext& shading);

}  // namespace rm
