// Shader-node library for generated material code (device/nodes/nodes.hpp).
// Node fields double as inputs and outputs: the generated code assigns the
// inputs, calls compute(), and downstream nodes read output fields
// (device/nodes/nodes.hpp, e.g. Math :1150-1157; host/scene/MaterialBuilder.swift:124-128).
// Field names follow the reference's camelCase socket identifiers.
#pragma once
#include "common.h"
#include "elem.h"
#include "noise.h"
#include "shading.h"
#include "spectral.h"
#include "texture.h"
#include "warp.h"

namespace rm {

RM_HD float luminance(vec3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }

// (device/utils/color.hpp:3-73)
RM_HD vec3 rgb2hsv(vec3 rgb) {
    const float cmax = max_(rgb.x, max_(rgb.y, rgb.z));
    const float cmin = min_(rgb.x, min_(rgb.y, rgb.z));
    const float cdelta = cmax - cmin;
    float h = 0.0f, s = 0.0f;
    const float v = cmax;
    if (cmax != 0.0f) s = cdelta / cmax;
    if (s != 0.0f) {
        const vec3 c = (cmax - rgb) / cdelta;
        if (rgb.x == cmax) h = (0.0f + c.z) - c.y;
        else if (rgb.y == cmax) h = (2.0f + c.x) - c.z;
        else h = (4.0f + c.y) - c.x;
        h /= 6.0f;
        if (h < 0.0f) h += 1.0f;
    }
    return vec3(h, s, v);
}
RM_HD vec3 hsv2rgb(vec3 hsv) {
    float h = hsv.x;
    const float s = hsv.y, v = hsv.z;
    if (s == 0.0f) return vec3(v, v, v);
    if (h == 1.0f) h = 0.0f;
    h *= 6.0f;
    const float i = floor_(h);
    const float f = h - i;
    const float p = v * (1.0f - s);
    const float q = v * (1.0f - (s * f));
    const float t = v * (1.0f - (s * (1.0f - f)));
    switch ((int)i) {
    case 0: return vec3(v, t, p);
    case 1: return vec3(q, v, p);
    case 2: return vec3(p, v, t);
    case 3: return vec3(p, q, v);
    case 4: return vec3(t, p, v);
    default: return vec3(v, p, q);
    }
}

// ---------------------------------------------------------------------------
// Socket type coercions (device/nodes/nodes.hpp:2007-2031)
RM_HD vec3 VECTOR(float v) { return vec3(v); }
RM_HD vec3 VECTOR(vec2 v) { return vec3(v, 0.0f); }
RM_HD vec3 VECTOR(vec3 v) { return v; }
RM_HD vec3 VECTOR(vec4 v) { return v.xyz(); }
RM_HD vec3 VECTOR(const UberShader&) { return vec3(0.0f); }
RM_HD vec4 RGBA(float v) { return vec4(v, v, v, 1.0f); }
RM_HD vec4 RGBA(vec2 v) { return vec4(v.x, v.y, 0.0f, 1.0f); }
RM_HD vec4 RGBA(vec3 v) { return vec4(v, 1.0f); }
RM_HD vec4 RGBA(vec4 v) { return v; }
RM_HD vec4 RGBA(const UberShader& v) { return vec4(v.emission, 1.0f); }
RM_HD float VALUE(float v) { return v; }
RM_HD float VALUE(vec3 v) { return ((v.x + v.y) + v.z) / 3.0f; }
RM_HD float VALUE(vec4 v) { return (v.w * ((v.x + v.y) + v.z)) / 3.0f; }
RM_HD UberShader SHADER(const UberShader& v) { return v; }
RM_HD UberShader SHADER(vec3 v) { UberShader m; m.emission = v; return m; }
RM_HD UberShader SHADER(float v) { return SHADER(vec3(v)); }
RM_HD UberShader SHADER(vec2 v) { return SHADER(vec3(v, 0.0f)); }
RM_HD UberShader SHADER(vec4 v) { return SHADER(v.xyz()); }

// ---------------------------------------------------------------------------
// Inputs / geometry

struct Value {
    float value = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};
struct RGB {
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

struct TextureCoordinate {
    vec3 generated, uv, object, normal, reflection;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        uv = s.uv;
        generated = s.generated;
        object = s.object;
        normal = s.normal;
        reflection = s.normal;
    }
};

struct NewGeometry {
    vec3 normal, trueNormal, tangent, position, parametric, incoming, randomPerIsland;
    bool backfacing = false;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        normal = s.normal;
        trueNormal = s.true_normal;
        tangent = s.tu;
        position = s.position;
        parametric = s.uv;
        incoming = s.wo;
        randomPerIsland = vec3(0.0f);
        backfacing = dot(s.wo, s.normal) < 0.0f;
    }
};

// ---------------------------------------------------------------------------
// Textures

struct kTexNoise { enum Dimension { DIMENSION_1D, DIMENSION_2D, DIMENSION_3D, DIMENSION_4D }; };

template <int Dimension>
struct TexNoise {
    float detail = 0.0f, distortion = 0.0f, roughness = 0.0f, scale = 0.0f, w = 0.0f;
    vec3 vector;
    float fac = 0.0f;
    vec4 color;

    RM_HD static float off1(float seed) { return 100.0f + hash_float_to_float(seed) * 100.0f; }
    RM_HD static vec2 off2(float seed) {
        return vec2(100.0f + hash_float2_to_float(vec2(seed, 0.0f)) * 100.0f,
                    100.0f + hash_float2_to_float(vec2(seed, 1.0f)) * 100.0f);
    }
    RM_HD static vec3 off3(float seed) {
        return vec3(100.0f + hash_float2_to_float(vec2(seed, 0.0f)) * 100.0f,
                    100.0f + hash_float2_to_float(vec2(seed, 1.0f)) * 100.0f,
                    100.0f + hash_float2_to_float(vec2(seed, 2.0f)) * 100.0f);
    }

    RM_HD void tex1(float p) {
        if (distortion != 0.0f) p += snoise1(p + off1(0.0f)) * distortion;
        color = vec4(fractal_noise(p, detail, roughness), fractal_noise(p + off1(1.0f), detail, roughness),
                     fractal_noise(p + off1(2.0f), detail, roughness), 1.0f);
        fac = color.x;
    }
    RM_HD void tex2(vec2 p) {
        if (distortion != 0.0f) p = p + distortion * vec2(snoise2(p + off2(0.0f)), snoise2(p + off2(1.0f)));
        color = vec4(fractal_noise(p, detail, roughness), fractal_noise(p + off2(2.0f), detail, roughness),
                     fractal_noise(p + off2(3.0f), detail, roughness), 1.0f);
        fac = color.x;
    }
    RM_HD static vec4 off4(float seed) {
        return vec4(100.0f + hash_float2_to_float(vec2(seed, 0.0f)) * 100.0f,
                    100.0f + hash_float2_to_float(vec2(seed, 1.0f)) * 100.0f,
                    100.0f + hash_float2_to_float(vec2(seed, 2.0f)) * 100.0f,
                    100.0f + hash_float2_to_float(vec2(seed, 3.0f)) * 100.0f);
    }
    RM_HD void tex3(vec3 p) {
        if (distortion != 0.0f)
            p = p + distortion * vec3(snoise3(p + off3(0.0f)), snoise3(p + off3(1.0f)), snoise3(p + off3(2.0f)));
        color = vec4(fractal_noise(p, detail, roughness), fractal_noise(p + off3(3.0f), detail, roughness),
                     fractal_noise(p + off3(4.0f), detail, roughness), 1.0f);
        fac = color.x;
    }

    RM_HD void tex4(vec4 p) {
        if (distortion != 0.0f)
            p = p + distortion * vec4(snoise4(p + off4(0.0f)), snoise4(p + off4(1.0f)), snoise4(p + off4(2.0f)),
                                      snoise4(p + off4(3.0f)));
        color = vec4(fractal_noise(p, detail, roughness), fractal_noise(p + off4(4.0f), detail, roughness),
                     fractal_noise(p + off4(5.0f), detail, roughness), 1.0f);
        fac = color.x;
    }

    // (device/nodes/nodes.hpp:881-1011)
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const vec3 p = vector * scale;
        const float ws = w * scale;
        switch (Dimension) {
        case kTexNoise::DIMENSION_1D: tex1(ws); break;
        case kTexNoise::DIMENSION_2D: tex2(p.xy()); break;
        case kTexNoise::DIMENSION_4D: tex4(vec4(p.x, p.y, p.z, ws)); break;
        default: tex3(p); break;
        }
    }
};

// Stub, as in the reference: colour 1, distance 0 (device/nodes/nodes.hpp:590-606)
struct TexVoronoi {
    float exponent = 0.0f, randomness = 0.0f, scale = 0.0f, smoothness = 0.0f;
    vec3 vector;
    float w = 0.0f;
    vec4 color;
    float distance = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        color = vec4(1.0f);
        distance = 0.0f;
    }
};

// Analytic sky: a zenith-horizon gradient plus an exponential sun lobe.
// Not a reference node (the reference's Nishita sky is a precomputed
// texture); `vector` defaults to the direction into the sky.
struct TexSkyAnalytic {
    vec3 vector;
    vec3 sunDirection = vec3(0.0f, 0.0f, 1.0f);
    float sunStrength = 0.0f;
    float sunSharpness = 1.0f;
    vec4 zenithColor, horizonColor, groundColor;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const vec3 d = normalize(vector);
        vec3 c;
        if (d.z >= 0.0f) {
            const float t = sqrt_(d.z);
            c = (1.0f - t) * horizonColor.xyz() + t * zenithColor.xyz();
        } else {
            const float t = min_(1.0f, -4.0f * d.z);
            c = (1.0f - t) * horizonColor.xyz() + t * groundColor.xyz();
        }
        const float cs = dot(d, normalize(sunDirection));
        const float sun = sunStrength * exp_(sunSharpness * (cs - 1.0f));
        color = vec4(c + vec3(sun), 1.0f);
    }
};

// ---------------------------------------------------------------------------
// Converters

template <int NumElements>
struct ColorRamp {
    struct Element {
        float position;
        vec4 color;
    };
    float fac = 0.0f;
    vec4 color;
    Element elements[NumElements];

    // Linear interpolation between stops (device/nodes/nodes.hpp:1013-1046).
    // The first stop at or above fac is found by an unrolled select chain, so
    // the stop array is never indexed at run time and stays in registers.
    RM_HD void compute(const SceneView&, ShadingContext&) {
        if (fac < elements[0].position) { color = elements[0].color; return; }
        if (fac > elements[NumElements - 1].position) { color = elements[NumElements - 1].color; return; }
        // Selects are taken on values: a conditional between two lvalues
        // selects their addresses, which would pin the array in memory.
        bool found = false;
        float pa = 0.0f, pb = 0.0f;
        vec4 ca, cb;
#if defined(RM_GPU)
#pragma unroll
#endif
        for (int i = 1; i < NumElements; ++i) {
            const bool take = !found && elements[i].position >= fac;
            pa = select(take, elements[i - 1].position, pa);
            pb = select(take, elements[i].position, pb);
            ca = select(take, elements[i - 1].color, ca);
            cb = select(take, elements[i].color, cb);
            found = found || take;
        }
        if (found) {
            const float v = (fac - pa) / (pb - pa);
            color = (1.0f - v) * ca + v * cb;
        }
    }
};

struct kMath {
    enum Operation {
        OPERATION_ADD, OPERATION_SUBTRACT, OPERATION_MULTIPLY, OPERATION_DIVIDE, OPERATION_MULTIPLY_ADD,
        OPERATION_POWER, OPERATION_MINIMUM, OPERATION_MAXIMUM, OPERATION_TANGENT, OPERATION_LESS_THAN,
        OPERATION_GREATER_THAN, OPERATION_MODULO
    };
};

RM_HD float fmod_(float a, float b) {
#if defined(RM_GPU)
    return __builtin_fmodf(a, b);
#else
    return fmodf(a, b);
#endif
}

// (device/nodes/nodes.hpp:1145-1193)
template <int Operation, bool Clamp>
struct Math {
    float value = 0.0f, value_001 = 0.0f, value_002 = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        switch (Operation) {
        case kMath::OPERATION_ADD: value = value + value_001; break;
        case kMath::OPERATION_SUBTRACT: value = value - value_001; break;
        case kMath::OPERATION_MULTIPLY: value = value * value_001; break;
        case kMath::OPERATION_DIVIDE: value = safe_divide(value, value_001, 0.0f); break;
        case kMath::OPERATION_MULTIPLY_ADD: value = value * value_001 + value_002; break;
        case kMath::OPERATION_POWER: value = pow_(value, value_001); break;
        case kMath::OPERATION_MINIMUM: value = select(value < value_001, value, value_001); break;
        case kMath::OPERATION_MAXIMUM: value = select(value > value_001, value, value_001); break;
        case kMath::OPERATION_TANGENT: value = tan_(value); break;
        case kMath::OPERATION_LESS_THAN: value = value < value_001 ? 1.0f : 0.0f; break;
        case kMath::OPERATION_GREATER_THAN: value = value > value_001 ? 1.0f : 0.0f; break;
        case kMath::OPERATION_MODULO: value = fmod_(value, value_001); break;
        }
        if (Clamp) value = saturate(value);
    }
};

struct kColorMix {
    enum BlendType {
        BLEND_TYPE_MIX, BLEND_TYPE_ADD, BLEND_TYPE_MULTIPLY, BLEND_TYPE_SCREEN, BLEND_TYPE_OVERLAY,
        BLEND_TYPE_SUB, BLEND_TYPE_COLOR, BLEND_TYPE_LIGHTEN, BLEND_TYPE_DARKEN, BLEND_TYPE_VALUE
    };
};

RM_HD vec4 vmin4(vec4 a, vec4 b) { return vec4(min_(a.x, b.x), min_(a.y, b.y), min_(a.z, b.z), min_(a.w, b.w)); }
RM_HD vec4 vmax4(vec4 a, vec4 b) { return vec4(max_(a.x, b.x), max_(a.y, b.y), max_(a.z, b.z), max_(a.w, b.w)); }

// (device/nodes/nodes.hpp:1319-1385); SCREEN uses color1 twice as in the reference
template <int BlendType, bool Clamp>
struct ColorMix {
    vec4 color1, color2;
    float fac = 0.0f;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        switch (BlendType) {
        case kColorMix::BLEND_TYPE_MIX: color = lerp(color1, color2, fac); break;
        case kColorMix::BLEND_TYPE_ADD: color = color1 + fac * color2; break;
        case kColorMix::BLEND_TYPE_SUB: color = color1 - fac * color2; break;
        case kColorMix::BLEND_TYPE_MULTIPLY: color = color1 * lerp(vec4(1.0f), color2, fac); break;
        case kColorMix::BLEND_TYPE_SCREEN: color = 1.0f - (1.0f - fac * color1) * (1.0f - color1); break;
        case kColorMix::BLEND_TYPE_OVERLAY: {
            color = color1;
            float* c = &color.x;
            const float* c2 = &color2.x;
            for (int dim = 0; dim < 3; dim++) {
                if (c[dim] < 0.5f) c[dim] *= (1.0f - fac) + (2.0f * fac) * c2[dim];
                else c[dim] = 1.0f - ((1.0f - fac) + (2.0f * fac) * (1.0f - c2[dim])) * (1.0f - c[dim]);
            }
            break;
        }
        case kColorMix::BLEND_TYPE_COLOR: {
            color = color1;
            const vec3 hsv2 = rgb2hsv(color2.xyz());
            if (hsv2.y == 0.0f) {
                color = color1;
            } else {
                vec3 hsv = rgb2hsv(color1.xyz());
                hsv.x = hsv2.x;
                hsv.y = hsv2.y;
                color = lerp(color1, vec4(hsv2rgb(hsv), color2.w), fac);
            }
            break;
        }
        case kColorMix::BLEND_TYPE_LIGHTEN: color = lerp(color1, vmax4(color1, color2), fac); break;
        case kColorMix::BLEND_TYPE_DARKEN: color = lerp(color1, vmin4(color1, color2), fac); break;
        case kColorMix::BLEND_TYPE_VALUE: {
            vec3 hsv = rgb2hsv(color1.xyz());
            hsv.z = rgb2hsv(color2.xyz()).z;
            color = lerp(color1, vec4(hsv2rgb(hsv), color2.w), fac);
            break;
        }
        }
        if (Clamp) color = saturate(color);
    }
};

// No-op, as in the reference: `normal` passes its input through (device/nodes/nodes.hpp:294-302)
struct Bump {
    float distance = 0.0f, height = 0.0f;
    vec3 normal;
    float strength = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

// ---------------------------------------------------------------------------
// Shaders

struct kBsdfPrincipled {
    enum Distribution { DISTRIBUTION_GGX };
    enum SubsurfaceMethod { SUBSURFACE_METHOD_BURLEY, SUBSURFACE_METHOD_RANDOM_WALK, SUBSURFACE_METHOD_RANDOM_WALK_FIXED_RADIUS };
};

// Disney parameters to lobe weights (device/nodes/nodes.hpp:1581-1687)
template <int Distribution, int SubsurfaceMethod>
struct BsdfPrincipled {
    float weight = 0.0f;
    vec4 emission;
    float sheenTint = 0.0f, emissionStrength = 0.0f, transmission = 0.0f;
    vec3 clearcoatNormal;
    float alpha = 0.0f, specularTint = 0.0f;
    vec3 tangent;
    float roughness = 0.0f, subsurfaceIor = 0.0f, anisotropic = 0.0f, sheen = 0.0f;
    vec3 subsurfaceRadius, normal;
    float subsurfaceAnisotropy = 0.0f;
    vec4 baseColor;
    float transmissionRoughness = 0.0f, metallic = 0.0f, specular = 0.0f, clearcoatRoughness = 0.0f;
    vec4 subsurfaceColor;
    float subsurface = 0.0f, ior = 0.0f, anisotropicRotation = 0.0f, clearcoat = 0.0f;
    UberShader bsdf;

    RM_HD void compute(const SceneView&, ShadingContext&) {
        clearcoatRoughness = max_(clearcoatRoughness, 1e-4f);
        roughness = max_(roughness, 1e-4f);
        const float diffuse_w = (1.0f - saturate(transmission)) * (1.0f - saturate(metallic));
        const float transmission_w = saturate(transmission) * (1.0f - saturate(metallic));
        const float specular_w = 1.0f - transmission_w;
        const vec3 base = baseColor.xyz();
        const float lum = luminance(base);
        const vec3 tint = lum > 0.0f ? base * (1.0f / lum) : vec3(1.0f);
        const vec3 sheen_color = lerp(vec3(1.0f), tint, sheenTint);
        const vec3 specular_color = lerp(vec3(1.0f), tint, specularTint);
        const vec3 cspec0 = lerp((specular * 0.08f) * specular_color, base, metallic);
        const float aspect = sqrt_(1.0f - 0.9f * anisotropic);
        const float r2 = square(roughness);

        bsdf.diffuse.diffuse_weight = diffuse_w * base;
        bsdf.diffuse.sheen_weight = (diffuse_w * sheen) * sheen_color;
        bsdf.diffuse.roughness = roughness;
        bsdf.diffuse.translucent = false;

        bsdf.specular.alpha_x = r2 / aspect;
        bsdf.specular.alpha_y = r2 * aspect;
        bsdf.specular.cspec0 = cspec0;
        bsdf.specular.ior = (2.0f / (1.0f - sqrt_(0.08f * specular))) - 1.0f;
        bsdf.specular.weight = specular_w;

        bsdf.transmission.reflection_alpha = r2;
        bsdf.transmission.transmission_alpha = square(1.0f - (1.0f - roughness) * (1.0f - transmissionRoughness));
        bsdf.transmission.base_color = base;
        bsdf.transmission.cspec0 = lerp(vec3(1.0f), base, specularTint);
        bsdf.transmission.ior = ior;
        bsdf.transmission.weight = transmission_w;
        bsdf.transmission.only_refract = false;

        bsdf.clearcoat.alpha = square(clearcoatRoughness);
        bsdf.clearcoat.weight = clearcoat;

        bsdf.lobe_prob[0] = diffuse_w;
        bsdf.lobe_prob[1] = specular_w;
        bsdf.lobe_prob[2] = transmission_w;
        bsdf.lobe_prob[3] = clearcoat * 0.25f;
        float sum = 0.0f;
        for (int i = 0; i < 4; i++) sum += bsdf.lobe_prob[i];
        for (int i = 0; i < 4; i++) bsdf.lobe_prob[i] /= sum;

        bsdf.alpha = alpha;
        bsdf.normal = normal;
        bsdf.emission = (alpha * emission.xyz()) * emissionStrength;
    }
};

// Burley diffuse (device/nodes/nodes.hpp:1809-1827)
struct BsdfDiffuse {
    vec4 color;
    vec3 normal;
    float roughness = 0.0f, weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        bsdf.diffuse.diffuse_weight = color.xyz();
        bsdf.diffuse.sheen_weight = vec3(0.0f);
        bsdf.diffuse.roughness = roughness;
        bsdf.diffuse.translucent = false;
        bsdf.normal = normal;
        bsdf.lobe_prob[0] = 1.0f;
    }
};

// (device/nodes/nodes.hpp:1396-1407)
struct Emission {
    vec4 color;
    float strength = 0.0f, weight = 0.0f;
    UberShader emission;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        emission.lobe_prob[0] = 1.0f;
        emission.emission = color.xyz() * strength;
    }
};

// (device/nodes/nodes.hpp:1409-1419)
struct Background {
    vec4 color;
    float strength = 0.0f, weight = 0.0f;
    UberShader background;
    RM_HD void compute(const SceneView&, ShadingContext&) { background.emission = color.xyz() * strength; }
};

// Stochastic shader mix on shading.rnd.x (device/nodes/nodes.hpp:1916-1931)
struct MixShader {
    float fac = 0.0f;
    UberShader shader, shader_001;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        const vec3 em = fac * shader_001.emission + (1.0f - fac) * shader.emission;
        if (s.rnd.x < fac) {
            s.rnd.x = s.rnd.x / fac;
            shader = shader_001;
        } else {
            s.rnd.x = (s.rnd.x - fac) / (1.0f - fac);
        }
        shader.emission = em;
    }
};

// (device/nodes/nodes.hpp:1899-1913); the weight doubling is the reference's
struct AddShader {
    UberShader shader, shader_001;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        const vec3 em = shader_001.emission + shader.emission;
        if (s.rnd.x < 0.5f) {
            s.rnd.x = s.rnd.x / 0.5f;
            shader = shader_001;
        } else {
            s.rnd.x = 2.0f * (s.rnd.x - 0.5f);
        }
        shader.emission = em;
        shader.weight *= 2.0f;
    }
};

// ---------------------------------------------------------------------------
// Further converters, textures and BSDFs (device/nodes/nodes.hpp).  Each
// follows the reference's node, simplifications included; field names are
// the codegen's identifiers of the exporter's socket names.

// (device/nodes/nodes.hpp:15-46); flags as 0/1 floats
struct LightPath {
    float isCameraRay = 0.0f, isReflectionRay = 0.0f, isTransmissionRay = 0.0f, isShadowRay = 0.0f;
    float isDiffuseRay = 0.0f, isGlossyRay = 0.0f, isSingularRay = 0.0f, rayLength = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        auto flag = [&](uint32_t f) { return (s.ray_flags & f) ? 1.0f : 0.0f; };
        isCameraRay = flag(RayCamera);
        isReflectionRay = flag(RayReflection);
        isTransmissionRay = flag(RayTransmission);
        isShadowRay = flag(RayShadow);
        isDiffuseRay = flag(RayDiffuse);
        isGlossyRay = flag(RayGlossy);
        isSingularRay = flag(RaySingular);
        rayLength = s.distance;
        if (isSingularRay != 0.0f) isGlossyRay = 1.0f;
    }
};

struct kVectorMath {
    enum Operation {
        OPERATION_ADD, OPERATION_SUBTRACT, OPERATION_MULTIPLY, OPERATION_MULTIPLY_ADD, OPERATION_NORMALIZE,
        OPERATION_SCALE, OPERATION_MINIMUM, OPERATION_LENGTH, OPERATION_DOT_PRODUCT
    };
};

// (device/nodes/nodes.hpp:151-199); the dot product lands in `vector`
template <int Operation>
struct VectorMath {
    float scale = 0.0f, value = 0.0f;
    vec3 vector, vector_001, vector_002;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        switch (Operation) {
        case kVectorMath::OPERATION_ADD: vector = vector + vector_001; break;
        case kVectorMath::OPERATION_SUBTRACT: vector = vector - vector_001; break;
        case kVectorMath::OPERATION_MULTIPLY: vector = vector * vector_001; break;
        case kVectorMath::OPERATION_MULTIPLY_ADD: vector = vector * vector_001 + vector_002; break;
        case kVectorMath::OPERATION_NORMALIZE: vector = normalize(vector); break;
        case kVectorMath::OPERATION_SCALE: vector = vector * scale; break;
        case kVectorMath::OPERATION_MINIMUM: vector = vmin(vector, vector_001); break;
        case kVectorMath::OPERATION_LENGTH: value = length(vector); break;
        case kVectorMath::OPERATION_DOT_PRODUCT: vector = vec3(dot(vector, vector_001)); break;
        }
    }
};

// floor to int, saturated to the int range so that the GPU and the host agree
RM_HD int floor_to_int(float x) {
    const float f = floor_(x);
    if (!(f >= -2147483648.0f)) return (int)0x80000000;
    if (f >= 2147483648.0f) return 0x7fffffff;
    return (int)f;
}

// ---------------------------------------------------------------------------
// Image textures and the Nishita sky (device/nodes/nodes.hpp:443-564,
// 646-870): lookups in the scene's texture slots.

// sRGB transfer function to linear (IEC 61966-2-1)
RM_HD float srgb_to_linear(float c) {
    if (c < 0.04045f) return c < 0.0f ? 0.0f : c * (1.0f / 12.92f);
    return pow_((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}

struct kTexImage {
    enum Interpolation { INTERPOLATION_LINEAR };
    enum Projection { PROJECTION_FLAT, PROJECTION_BOX, PROJECTION_EQUIRECTANGULAR, PROJECTION_MIRROR_BALL };
    enum Extension { EXTENSION_REPEAT };
    enum Alpha { ALPHA_STRAIGHT };
    enum ColorSpace {
        COLOR_SPACE_LINEAR, COLOR_SPACE_SRGB, COLOR_SPACE_NON_COLOR, COLOR_SPACE_RAW, COLOR_SPACE_XYZ,
        COLOR_SPACE_FILMIC_LOG
    };
};

// Image and environment textures: flat projection flips v, equirectangular
// maps a direction to (longitude, polar angle); box and mirror-ball
// projections are not implemented by the reference (its coordinate is left
// unset) and sample (0, 0) here.  The texel format (grey or RGBA) is the
// slot's, read at run time.
template <int Slot, int Interpolation, int Projection, int Extension, int Alpha, int ColorSpace>
struct TexImage {
    vec3 vector;
    vec4 color;
    float alpha = 0.0f;
    RM_HD void compute(const SceneView& ctx, ShadingContext&) {
        vec2 projected(0.0f, 0.0f);
        if (Projection == kTexImage::PROJECTION_FLAT) projected = vec2(vector.x, 1.0f - vector.y);
        if (Projection == kTexImage::PROJECTION_EQUIRECTANGULAR) projected = warp::equirect_sphere_to_square(vector);
        color = texture_sample(ctx.textures[Slot], projected);
        alpha = color.w;
        if (ColorSpace == kTexImage::COLOR_SPACE_SRGB)
            color = vec4(srgb_to_linear(color.x), srgb_to_linear(color.y), srgb_to_linear(color.z), color.w);
        if (ColorSpace == kTexImage::COLOR_SPACE_XYZ) color = vec4(xyz_to_rgb(color.xyz()), color.w);
    }
};

// Nishita sky radiance from the precomputed texture (longitude x sqrt of
// elevation, CIE XYZ) and the sun's two edge colours: data = bottom XYZ (3),
// top XYZ (3), sun elevation, sun rotation, angular diameter (< 0: no disc),
// sun intensity (Blender's Cycles sky kernel, node_sky_texture.osl)
RM_HD vec3 geographical_to_direction(float lat, float lon) {
    const float cl = cos_(lat);
    return vec3(cl * cos_(lon), cl * sin_(lon), sin_(lat));
}
RM_HD vec3 sky_radiance_nishita(vec3 dir, const float* data, const Texture& tex) {
    const float sun_elevation = data[6], sun_rotation = data[7], angular_diameter = data[8], sun_intensity = data[9];
    const bool sun_disc = angular_diameter > 0.0f;
    const float polar = acos_(dir.z), azimuth = atan2_(dir.x, dir.y);
    vec3 xyz(0.0f);
    if (dir.z >= 0.0f) {
        const vec3 sun_dir = geographical_to_direction(sun_elevation, sun_rotation + 0.5f * kPi);
        // the angle between the two unit vectors, accurate near 0
        const float sun_angle = 2.0f * atan2_(length(dir - sun_dir), length(dir + sun_dir));
        const float half_angular = angular_diameter / 2.0f;
        const float dir_elevation = 0.5f * kPi - polar;
        if (sun_angle < half_angular && sun_disc) {
            const vec3 bottom(data[0], data[1], data[2]), top(data[3], data[4], data[5]);
            if (sun_elevation - half_angular > 0.0f) {
                if (sun_elevation + half_angular > 0.0f) {
                    const float y = (dir_elevation - sun_elevation) / angular_diameter + 0.5f;
                    xyz = lerp(bottom, top, y) * sun_intensity;
                }
            } else if (sun_elevation + half_angular > 0.0f) {
                const float y = dir_elevation / (sun_elevation + half_angular);
                xyz = lerp(bottom, top, y) * sun_intensity;
            }
            // limb darkening with coefficient 0.6
            const float f = sun_angle / half_angular;
            xyz *= 1.0f - 0.6f * (1.0f - sqrt_(1.0f - f * f));
        } else {
            float x = (azimuth + kPi + sun_rotation) / (2.0f * kPi);
            const float y = sqrt_(dir_elevation / (0.5f * kPi));  // more texels toward the horizon
            if (x > 1.0f) x = x - 1.0f;
            xyz = texture_sample(tex, vec2(x, y)).xyz();
        }
    } else if (dir.z >= -0.4f) {
        // the ground fades to black below the horizon
        const float fade = pow_(1.0f + dir.z * 2.5f, 3.0f);
        float x = (azimuth + kPi + sun_rotation) / (2.0f * kPi);
        if (x > 1.0f) x = x - 1.0f;
        xyz = texture_sample(tex, vec2(x, 1e-3f)).xyz() * fade;
    }
    return xyz_to_rgb(xyz);
}

// (device/nodes/nodes.hpp:857-870); the direction is the outgoing one in the
// sky texture's frame (y and z flipped)
template <int Slot>
struct TexNishita {
    float scale = 0.0f;
    vec3 vector;
    vec4 color;
    float data[10] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    RM_HD void compute(const SceneView& ctx, ShadingContext& s) {
        const vec3 d = s.wo * vec3(1.0f, -1.0f, -1.0f);
        color = vec4(scale * sky_radiance_nishita(d, data, ctx.textures[Slot]), 1.0f);
    }
};

// ---------------------------------------------------------------------------
// Nodes the reference maps to constant stubs (device/nodes/nodes.hpp:48-150,
// 407-441, 566-644, 1051-1110, 1736-1744): they compile and produce the
// reference's fixed outputs, so scenes that use them render as the
// reference renders them.

struct ObjectInfo {
    float random = 0.0f;
    vec3 location;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

struct AmbientOcclusion {
    vec4 color;
    float distance = 0.0f;
    vec3 normal;
    float ao = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) { ao = 1.0f; }
};

// volume closures are not rendered (nodes.hpp:71-119)
struct VolumeScatter {
    float anisotropy = 0.0f, density = 0.0f, weight = 0.0f;
    vec4 color;
    UberShader volume;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};
struct VolumeAbsorption {
    float density = 0.0f, weight = 0.0f;
    vec4 color;
    UberShader volume;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};
struct PrincipledVolume {
    vec4 absorptionColor, blackbodyTint, color, emissionColor;
    float anisotropy = 0.0f, blackbodyIntensity = 0.0f, density = 0.0f, emissionStrength = 0.0f;
    float temperature = 0.0f, weight = 0.0f;
    UberShader volume;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

struct ParticleInfo {
    float random = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

struct LightFalloff {
    float strength = 0.0f, smooth = 0.0f, quadratic = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) { quadratic = 1.0f; }
};

struct VertexColor {
    vec3 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

struct UVMapCoordinate {
    vec3 uv;
    RM_HD void compute(const SceneView&, ShadingContext& s) { uv = s.uv; }
};

// the curve mapping is not applied: colour and factor pass through
struct ColorCurves {
    vec4 color;
    float fac = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {}
};

// the "Normal" node's dot output is fixed at 1
struct NormalProduct {
    vec3 normal;
    float dot = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) { dot = 1.0f; }
};

struct kTexWave {
    enum Type { TYPE_BANDS };
    enum Direction { DIRECTION_DIAGONAL, DIRECTION_X };
    enum Profile { PROFILE_SIN };
};
template <int Type, int Direction, int Profile>
struct TexWave {
    float detail = 0.0f, detailRoughness = 0.0f, detailScale = 0.0f, distortion = 0.0f, phaseOffset = 0.0f;
    float scale = 0.0f;
    vec3 vector;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) { color = vec4(1.0f); }
};

struct TexIES {
    vec3 vector;
    float strength = 0.0f, fac = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) { fac = strength; }
};

struct TexMagic {
    float distortion = 0.0f, scale = 0.0f;
    vec3 vector;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) { color = vec4(1.0f); }
};

struct TexMusgrave {
    float detail = 0.0f, dimension = 0.0f, gain = 0.0f, lacunarity = 0.0f, offset = 0.0f, scale = 0.0f, w = 0.0f;
    vec3 vector;
    float fac = 0.0f;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        fac = 1.0f;
        color = vec4(1.0f);
    }
};

struct TexBrick {
    vec4 color1, color2, mortar;
    float bias = 0.0f, brickWidth = 0.0f, mortarSize = 0.0f, mortarSmooth = 0.0f, rowHeight = 0.0f, scale = 0.0f;
    vec3 vector;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) { color = vec4(1.0f); }
};

struct Displacement {
    float height = 0.0f, midlevel = 0.0f, scale = 0.0f;
    vec3 normal, displacement;
    RM_HD void compute(const SceneView&, ShadingContext&) { displacement = vec3(0.0f); }
};

// "Attribute" returns the generated coordinates and white (nodes.hpp:1736-1744)
struct Attribute {
    vec3 vector;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        vector = s.generated;
        color = vec4(1.0f);
    }
};

// Tangent-space (or world-space) normal map (nodes.hpp:1060-1097): the
// colour maps to [-1, 1]^3, is blended towards +z by strength and taken
// through the frame (tu, tv, normal).
struct kNormalMap { enum Space { SPACE_TANGENT, SPACE_WORLD }; };
template <int Space>
struct NormalMap {
    vec4 color;
    float strength = 0.0f;
    vec3 normal;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        const float k = max_(strength, 0.0f);
        normal = normalize(2.0f * color.xyz() - 1.0f);
        normal = k * normal + (1.0f - k) * vec3(0.0f, 0.0f, 1.0f);
        normal = normalize(normal);
        if (Space == kNormalMap::SPACE_TANGENT) {
            mat3 onb;
            onb.c0 = s.tu;
            onb.c1 = s.tv;
            onb.c2 = s.normal;
            normal = onb * normal;
        }
    }
};

// Cycles' piecewise fit of the blackbody colour over 800-12000 K, normalised
// to unit luminance (nodes.hpp:1421-1495; Cycles' svm math_util.h): a/t +
// b t + c for red and green, a cubic in t for blue.
struct Blackbody {
    float temperature = 0.0f;
    vec4 color;
    RM_HD static vec3 rgb(float t) {
        if (t >= 12000.0f) return vec3(0.8262954810464208f, 0.9945080501520986f, 1.566307710274283f);
        if (t < 800.0f) return vec3(5.413294490189271f, -0.20319390035873933f, -0.0822535242887164f);
        const int i = (t >= 6365.0f) ? 6 : (t >= 3315.0f) ? 5 : (t >= 1902.0f) ? 4 : (t >= 1449.0f) ? 3
                    : (t >= 1167.0f) ? 2 : (t >= 965.0f) ? 1 : 0;
        // row i of the fit: red and green (a, b, c), blue (a, b, c, d); the
        // rows are picked by value so that no table is indexed at run time
        auto P = [i](float c0, float c1, float c2, float c3, float c4, float c5, float c6) {
            return i == 0 ? c0 : i == 1 ? c1 : i == 2 ? c2 : i == 3 ? c3 : i == 4 ? c4 : i == 5 ? c5 : c6;
        };
        const float ra = P(1.61919106e+03f, 2.48845471e+03f, 3.34143193e+03f, 4.09461742e+03f, 4.67028036e+03f,
                           4.59509185e+03f, 3.78717450e+03f);
        const float rb = P(-2.05010916e-03f, -1.11330907e-03f, -4.86551192e-04f, -1.27446582e-04f, 2.91258199e-05f,
                           2.87495649e-05f, 9.35907826e-06f);
        const float rc = P(5.02995757e+00f, 3.22621544e+00f, 1.76486769e+00f, 7.25731635e-01f, 1.26703442e-01f,
                           1.50345020e-01f, 3.99075871e-01f);
        const float ga = P(-4.88999748e+02f, -7.55994277e+02f, -1.02363977e+03f, -1.26571316e+03f, -1.42529332e+03f,
                           -1.17554822e+03f, -5.00799571e+02f);
        const float gb = P(6.04330754e-04f, 3.16730098e-04f, 1.20223470e-04f, 4.87340896e-06f, -4.01150431e-05f,
                           -2.16378048e-05f, -4.59832026e-06f);
        const float gc = P(-7.55807526e-02f, 4.78306139e-01f, 9.36662319e-01f, 1.27054498e+00f, 1.43972784e+00f,
                           1.30408023e+00f, 1.09098763e+00f);
        const float ba = P(5.96945309e-11f, 2.40430366e-11f, -1.40949732e-11f, -3.61460868e-11f, -1.97075738e-11f,
                           -1.61997957e-13f, 6.72650283e-13f);
        const float bb = P(-4.85742887e-08f, 5.55021075e-08f, 1.89878968e-07f, 2.84822009e-07f, 1.75359352e-07f,
                           -1.64216008e-08f, -2.73078809e-08f);
        const float bc = P(-9.70622247e-05f, -1.98503712e-04f, -3.56632824e-04f, -4.93211319e-04f, -2.50542825e-04f,
                           3.86216271e-04f, 4.24098264e-04f);
        const float bd = P(-4.07936148e-03f, 2.89312858e-02f, 9.10767778e-02f, 1.56723440e-01f, -2.22783266e-02f,
                           -7.38077418e-01f, -7.52335691e-01f);
        const float inv = 1.0f / t;
        return vec3((ra * inv + rb * t) + rc, (ga * inv + gb * t) + gc, ((ba * t + bb) * t + bc) * t + bd);
    }
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const vec3 b = rgb(temperature);
        color = vec4(b / luminance(b), 1.0f);
    }
};

// (device/nodes/nodes.hpp:247-264)
struct TexChecker {
    float scale = 0.0f;
    vec4 color1, color2;
    vec3 vector;
    float fac = 0.0f;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const vec3 p = (vector * scale + 0.000001f) * 0.999999f;
        const bool which = ((floor_to_int(p.x) ^ floor_to_int(p.y) ^ floor_to_int(p.z)) & 1) != 0;
        color = which ? color1 : color2;
        fac = which ? 1.0f : 0.0f;
    }
};

// (device/nodes/nodes.hpp:266-284)
struct SeparateVector {
    vec3 vector;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        x = vector.x;
        y = vector.y;
        z = vector.z;
    }
};
struct CombineVector {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    vec3 vector;
    RM_HD void compute(const SceneView&, ShadingContext&) { vector = vec3(x, y, z); }
};

// (device/nodes/nodes.hpp:304-344): float data, linear interpolation
template <bool Clamp>
struct MapRange {
    float fromMin = 0.0f, fromMax = 0.0f, toMin = 0.0f, toMax = 0.0f, steps = 0.0f;
    vec3 steps_float3, from_min_float3, from_max_float3, to_min_float3, to_max_float3;
    float value = 0.0f;
    vec3 vector;
    float result = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        float v = (value - fromMin) / (fromMax - fromMin);
        if (Clamp) v = saturate(v);
        result = lerpf(toMin, toMax, v);
    }
};

// XYZ Euler rotation, the matrix of node_mapping.osl (device/utils/math.hpp:43-67)
RM_HD vec3 euler_rotate(vec3 e, vec3 v) {
    float cx, cy, cz;
    const float sx = sincos_(e.x, cx);
    const float sy = sincos_(e.y, cy);
    const float sz = sincos_(e.z, cz);
    const vec3 c0(cy * cz, cy * sz, -sy);
    const vec3 c1(sy * sx * cz - cx * sz, sy * sx * sz + cx * cz, cy * sx);
    const vec3 c2(sy * cx * cz + sx * sz, sy * cx * sz - sx * cz, cy * cx);
    return (c0 * v.x + c1 * v.y) + c2 * v.z;
}

// (device/nodes/nodes.hpp:346-364): the point mapping (the exporter writes
// the node's type, "MAPPING", not its vector type)
struct Mapping {
    vec3 scale, rotation, location, vector;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        vector = euler_rotate(rotation, vector * scale) + location;
    }
};

struct kSeparateColor { enum Mode { MODE_RGB, MODE_HSV }; };
struct kCombineColor { enum Mode { MODE_RGB, MODE_HSV }; };

// (device/nodes/nodes.hpp:1195-1259); combining HSV passes the values
// through as the reference does
template <int Mode>
struct SeparateColor {
    vec4 color;
    float red = 0.0f, green = 0.0f, blue = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const vec3 c = Mode == kSeparateColor::MODE_HSV ? rgb2hsv(color.xyz()) : color.xyz();
        red = c.x;
        green = c.y;
        blue = c.z;
    }
};
template <int Mode>
struct CombineColor {
    float red = 0.0f, green = 0.0f, blue = 0.0f;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) { color = vec4(red, green, blue, 1.0f); }
};

// (device/nodes/nodes.hpp:1261-1277)
struct HueSaturation {
    vec4 color;
    float fac = 0.0f, hue = 0.0f, saturation = 0.0f, value = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        vec3 hsv = rgb2hsv(color.xyz());
        hsv.x = fmod_(hsv.x + hue + 0.5f, 1.0f);
        hsv.y = saturate(hsv.y * saturation);
        hsv.z *= value;
        const vec3 result = vmax(hsv2rgb(hsv), vec3(0.0f));
        color = vec4(lerp(color.xyz(), result, fac), color.w);
    }
};

// (device/nodes/nodes.hpp:1279-1290); alpha goes through the same map
struct BrightnessContrast {
    float bright = 0.0f, contrast = 0.0f;
    vec4 color;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const float a = 1.0f + contrast;
        const float b = bright - contrast / 2.0f;
        color = vmax4(a * color + b, vec4(0.0f));
    }
};

// (device/nodes/nodes.hpp:1292-1302)
struct Gamma {
    vec4 color;
    float gamma = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        if (gamma == 0.0f) {
            color.x = color.y = color.z = 1.0f;
        } else {
            if (color.x > 0.0f) color.x = pow_(color.x, gamma);
            if (color.y > 0.0f) color.y = pow_(color.y, gamma);
            if (color.z > 0.0f) color.z = pow_(color.z, gamma);
        }
    }
};

// (device/nodes/nodes.hpp:1387-1394)
struct ColorInvert {
    vec4 color;
    float fac = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const vec3 c = color.xyz();
        const vec3 r = c - fac * (2.0f * c - 1.0f);
        color = vec4(r, color.w);
    }
};

// (device/nodes/nodes.hpp:1726-1734)
struct RGBToBW {
    vec4 color;
    float val = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) { val = luminance(color.xyz()); }
};

// (device/nodes/nodes.hpp:1112-1126)
struct Fresnel {
    float ior = 0.0f;
    vec3 normal;
    float fac = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        const float cos_i = dot(s.wo, normal);
        float eta = max_(ior, 1e-5f);
        if (cos_i < 0.0f) eta = 1.0f / eta;
        fac = fresnel_dielectric_cos(cos_i, eta);
    }
};

// (device/nodes/nodes.hpp:1689-1714)
struct LayerWeight {
    float blend = 0.0f;
    vec3 normal;
    float fresnel = 0.0f, facing = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        const float cos_i = dot(s.wo, normal);
        float eta = max_(1.0f - blend, 1e-5f);
        eta = cos_i < 0.0f ? eta : 1.0f / eta;
        fresnel = fresnel_dielectric_cos(cos_i, eta);
        facing = abs_(cos_i);
        if (blend != 0.5f) {
            float b = min_(max_(blend, 0.0f), 1.0f - 1e-5f);
            b = b < 0.5f ? 2.0f * b : 0.5f / (1.0f - b);
            facing = pow_(facing, b);
        }
        facing = 1.0f - facing;
    }
};

struct kMix { enum FactorMode { FACTOR_MODE_UNIFORM, FACTOR_MODE_NON_UNIFORM }; };

// (device/nodes/nodes.hpp:1933-1974): every data type is computed; colours
// always blend linearly
template <bool ClampFactor, bool ClampResult, int FactorMode>
struct Mix {
    vec4 a_color, b_color, result_color;
    float a_float = 0.0f, b_float = 0.0f, result_float = 0.0f;
    vec3 a_vector, b_vector, result_vector;
    float factor_float = 0.0f;
    vec3 factor_vector;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        if (ClampFactor) {
            factor_float = saturate(factor_float);
            factor_vector = saturate(factor_vector);
        }
        if (FactorMode == kMix::FACTOR_MODE_UNIFORM) result_vector = lerp(a_vector, b_vector, factor_float);
        else result_vector = a_vector + (b_vector - a_vector) * factor_vector;
        result_float = lerpf(a_float, b_float, factor_float);
        result_color = lerp(a_color, b_color, factor_float);
        if (ClampResult) result_color = saturate(result_color);
    }
};

struct kTexGradient {
    enum Type { TYPE_LINEAR, TYPE_SPHERICAL, TYPE_EASING, TYPE_QUADRATIC, TYPE_QUADRATIC_SPHERE };
};

// (device/nodes/nodes.hpp:366-405): the easing and quadratic types use the
// linear ramp, as in the reference
template <int Type>
struct TexGradient {
    vec3 vector;
    vec4 color;
    float fac = 0.0f;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const float v = Type == kTexGradient::TYPE_SPHERICAL ? saturate(length(vector)) : saturate(vector.x);
        color = vec4(v);
        fac = color.x;
    }
};

// Glossy and anisotropic: one GGX specular lobe (device/nodes/nodes.hpp:1534-1567, 1746-1770)
struct BsdfGlossy {
    vec4 color;
    vec3 normal;
    float roughness = 0.0f, weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const float alpha = square(max_(roughness, 1e-4f));
        bsdf.specular.alpha_x = alpha;
        bsdf.specular.alpha_y = alpha;
        bsdf.specular.cspec0 = color.xyz();
        bsdf.specular.ior = 1.45f;
        bsdf.specular.weight = 1.0f;
        bsdf.lobe_prob[1] = 1.0f;
        bsdf.normal = normal;
    }
};
struct BsdfAnisotropic {
    float anisotropy = 0.0f;
    vec4 color;
    vec3 normal;
    float rotation = 0.0f;
    vec3 tangent;
    float roughness = 0.0f, weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView& v, ShadingContext& s) {
        BsdfGlossy g;
        g.color = color;
        g.normal = normal;
        g.roughness = roughness;
        g.compute(v, s);
        bsdf = g.bsdf;
    }
};

// Glass and refraction: the transmission lobe (device/nodes/nodes.hpp:1497-1532, 1772-1795)
struct BsdfGlass {
    vec4 color;
    float ior = 0.0f;
    vec3 normal;
    float roughness = 0.0f, weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const float alpha = square(max_(roughness, 1e-4f));
        bsdf.transmission.reflection_alpha = alpha;
        bsdf.transmission.transmission_alpha = alpha;
        bsdf.transmission.base_color = color.xyz();
        bsdf.transmission.cspec0 = color.xyz();
        bsdf.transmission.ior = ior;
        bsdf.transmission.weight = 1.0f;
        bsdf.lobe_prob[2] = 1.0f;
        bsdf.normal = normal;
    }
};
struct BsdfRefraction {
    vec4 color;
    float ior = 0.0f;
    vec3 normal;
    float roughness = 0.0f, weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        const float alpha = square(max_(roughness, 1e-4f));
        bsdf.transmission.reflection_alpha = alpha;
        bsdf.transmission.transmission_alpha = alpha;
        bsdf.transmission.base_color = color.xyz();
        bsdf.transmission.cspec0 = vec3(0.0f);
        bsdf.transmission.ior = ior;
        bsdf.transmission.weight = 1.0f;
        bsdf.transmission.only_refract = true;
        bsdf.lobe_prob[2] = 1.0f;
        bsdf.normal = normal;
    }
};

// (device/nodes/nodes.hpp:1797-1807)
struct BsdfTransparent {
    vec4 color;
    float weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        bsdf.alpha = 0.0f;
        bsdf.alpha_weight = color.xyz();
    }
};

// Hair is shaded as a diffuse lobe of its colour about the tangent
// (nodes.hpp:1849-1869)
struct BsdfHair {
    vec4 color;
    float offset = 0.0f, roughnessu = 0.0f, roughnessv = 0.0f, weight = 0.0f;
    vec3 tangent;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        bsdf.diffuse.diffuse_weight = color.xyz();
        bsdf.diffuse.sheen_weight = vec3(0.0f);
        bsdf.diffuse.roughness = 0.0f;
        bsdf.normal = tangent;
        bsdf.lobe_prob[0] = 1.0f;
    }
};

// Translucent and velvet: the diffuse lobe (device/nodes/nodes.hpp:1829-1847, 1871-1893)
struct BsdfTranslucent {
    vec4 color;
    vec3 normal;
    float weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        bsdf.diffuse.diffuse_weight = color.xyz();
        bsdf.diffuse.sheen_weight = vec3(0.0f);
        bsdf.diffuse.roughness = 1.0f;
        bsdf.diffuse.translucent = true;
        bsdf.normal = normal;
        bsdf.lobe_prob[0] = 1.0f;
    }
};
struct BsdfVelvet {
    vec4 color;
    vec3 normal;
    float sigma = 0.0f, weight = 0.0f;
    UberShader bsdf;
    RM_HD void compute(const SceneView&, ShadingContext&) {
        bsdf.diffuse.diffuse_weight = color.xyz();
        bsdf.diffuse.sheen_weight = vec3(0.0f);
        bsdf.diffuse.roughness = sigma;
        bsdf.diffuse.translucent = false;
        bsdf.normal = normal;
        bsdf.lobe_prob[0] = 1.0f;
    }
};

// Outputs (device/nodes/nodes.hpp:1976-2005)
struct OutputMaterial {
    vec3 displacement;
    float thickness = 0.0f;
    UberShader surface, volume;
    RM_HD void compute(const SceneView&, ShadingContext& s) { s.material = surface; }
};
struct OutputWorld {
    float thickness = 0.0f;
    UberShader surface;
    RM_HD void compute(const SceneView&, ShadingContext& s) { s.material = surface; }
};
struct OutputLight {
    UberShader surface;
    RM_HD void compute(const SceneView&, ShadingContext& s) {
        s.material.alpha = 0.0f;
        s.material.emission = surface.emission;
    }
};

}  // namespace rm
