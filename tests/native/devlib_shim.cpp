// Host build of the shared device library for CPU property tests: exposes
// RNG, elementary functions, warps, BSDF lobes, lights, the triangle test and
// the lens camera through plain C entry points.  Test infrastructure only.
#include "bvh2l.h"
#include "camera.h"
#include "isect.h"
#include "material_prelude.h"

// material hooks (unused by the functions exported here)
namespace rm {
inline void shade_surface(uint32_t, const SceneView&, ShadingContext&) {}
inline void shade_light(uint32_t, const SceneView&, ShadingContext&) {}
}  // namespace rm

using namespace rm;

extern "C" {

void shim_tea32(const uint32_t* a, const uint32_t* b, uint32_t* out, int n) {
    for (int i = 0; i < n; i++) out[i] = tea32(a[i], b[i]);
}
void shim_tea_float(const uint32_t* a, const uint32_t* b, float* out, int n) {
    for (int i = 0; i < n; i++) out[i] = tea_float(a[i], b[i]);
}
// Prng draw sequence: seed = TEA(a, b), then `count` floats
void shim_prng_sequence(uint32_t a, uint32_t b, float* out, int count) {
    Prng p(a, b);
    for (int i = 0; i < count; i++) out[i] = p.sample();
}

// 0 sin 1 cos 2 atan2(x, y) 3 acos 4 log 5 exp 6 pow(x, y) 7 tan 8 asin 9 atan
void shim_elem(int fn, const float* x, const float* y, float* out, int n) {
    for (int i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = sin_(x[i]); break;
        case 1: out[i] = cos_(x[i]); break;
        case 2: out[i] = atan2_(x[i], y[i]); break;
        case 3: out[i] = acos_(x[i]); break;
        case 4: out[i] = log_(x[i]); break;
        case 5: out[i] = exp_(x[i]); break;
        case 6: out[i] = pow_(x[i], y[i]); break;
        case 7: out[i] = tan_(x[i]); break;
        case 8: out[i] = asin_(x[i]); break;
        default: out[i] = atan_(x[i]); break;
        }
    }
}

// 0 square->sphere, 1 sphere->square, 2 cosine hemisphere, 3 disk, 4 triangle barycentric
void shim_warp(int fn, const float* in, float* out, int n) {
    for (int i = 0; i < n; i++) {
        switch (fn) {
        case 0: {
            const vec3 v = warp::uniform_square_to_sphere(vec2(in[2 * i], in[2 * i + 1]));
            out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
            break;
        }
        case 1: {
            const vec2 v = warp::uniform_sphere_to_square(vec3(in[3 * i], in[3 * i + 1], in[3 * i + 2]));
            out[2 * i] = v.x; out[2 * i + 1] = v.y;
            break;
        }
        case 2: {
            const vec3 v = warp::uniform_square_to_cosine_hemisphere(vec2(in[2 * i], in[2 * i + 1]));
            out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
            break;
        }
        case 3: {
            const vec2 v = warp::uniform_square_to_disk(vec2(in[2 * i], in[2 * i + 1]));
            out[2 * i] = v.x; out[2 * i + 1] = v.y;
            break;
        }
        default: {
            const vec2 v = warp::uniform_square_to_triangle_barycentric(vec2(in[2 * i], in[2 * i + 1]));
            out[2 * i] = v.x; out[2 * i + 1] = v.y;
            break;
        }
        }
    }
}

// Principled material from 14 parameters: base rgb, metallic, roughness,
// specular, specular tint, sheen, sheen tint, clearcoat, clearcoat roughness,
// transmission, ior, anisotropic.  lobe >= 0 keeps only that lobe (prob 1).
static UberShader make_material(const float* p, int lobe) {
    BsdfPrincipled<0, 0> n{};
    n.baseColor = vec4(p[0], p[1], p[2], 1.0f);
    n.metallic = p[3];
    n.roughness = p[4];
    n.specular = p[5];
    n.specularTint = p[6];
    n.sheen = p[7];
    n.sheenTint = p[8];
    n.clearcoat = p[9];
    n.clearcoatRoughness = p[10];
    n.transmission = p[11];
    n.ior = p[12];
    n.anisotropic = p[13];
    n.alpha = 1.0f;
    n.emissionStrength = 1.0f;
    SceneView ctx{};
    ShadingContext sc;
    n.compute(ctx, sc);
    UberShader m = n.bsdf;
    if (lobe >= 0) {
        for (int i = 0; i < 4; i++) m.lobe_prob[i] = i == lobe ? 1.0f : 0.0f;
        // evaluate() of the specular lobe omits its weight; make it 1 so that
        // single-lobe value/pdf comparisons see the same factor on both paths
        if (lobe == 1) m.specular.weight = 1.0f;
        if (lobe == 2) m.transmission.weight = 1.0f;
        if (lobe == 3) m.clearcoat.weight = 1.0f;
    }
    return m;
}

// local shading frame (normal +z): value (3) and pdf of wo -> wi
void shim_bsdf_eval(const float* params, int lobe, const float* wo, const float* wi, float* value, float* pdf, int n) {
    const UberShader m = make_material(params, lobe);
    for (int i = 0; i < n; i++) {
        float pd;
        const vec3 v = m.evaluate_local(vec3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]),
                                        vec3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), pd);
        value[3 * i] = v.x; value[3 * i + 1] = v.y; value[3 * i + 2] = v.z;
        pdf[i] = pd;
    }
}

// world = local frame (normal +z)
void shim_bsdf_sample(const float* params, int lobe, const float* wo, const float* rnd, float* wi, float* weight,
                      float* pdf, uint32_t* flags, int n) {
    const UberShader m = make_material(params, lobe);
    const vec3 nz(0.0f, 0.0f, 1.0f);
    for (int i = 0; i < n; i++) {
        const BsdfSample s = m.sample(vec3(rnd[3 * i], rnd[3 * i + 1], rnd[3 * i + 2]),
                                      vec3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), nz, nz, 0u);
        wi[3 * i] = s.wi.x; wi[3 * i + 1] = s.wi.y; wi[3 * i + 2] = s.wi.z;
        weight[3 * i] = s.weight.x; weight[3 * i + 1] = s.weight.y; weight[3 * i + 2] = s.weight.z;
        pdf[i] = s.pdf;
        flags[i] = s.flags;
    }
}

// triangle test: rays n x 8, triangles n x 9 -> hit flag, t, b0, b1
void shim_intersect(const float* rays, const float* tris, int n, uint32_t* hit, float* t, float* b0, float* b1) {
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        const float* v = tris + 9 * i;
        const vec3 o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
        const RayShear sh = make_shear(d);
        float tt = -1.0f, u = 0.0f, w = 0.0f;
        hit[i] = intersect_triangle(o, sh, vec3(v[0], v[1], v[2]), vec3(v[3], v[4], v[5]), vec3(v[6], v[7], v[8]), r[3],
                                    r[7], tt, u, w) ? 1u : 0u;
        t[i] = tt;
        b0[i] = u;
        b1[i] = w;
    }
}

// world light: sample (u, v) -> direction + pdf; pdf lookup of directions
void shim_world_sample(const float* pdfs, const float* mip, int res, const float* uv, float* dir, float* pdf, int n) {
    WorldLight w{0, res, pdfs, mip};
    for (int i = 0; i < n; i++) {
        float p;
        const vec3 d = world_sample(w, vec2(uv[2 * i], uv[2 * i + 1]), p);
        dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
        pdf[i] = p;
    }
}
void shim_world_pdf(const float* pdfs, const float* mip, int res, const float* dir, float* pdf, int n) {
    WorldLight w{0, res, pdfs, mip};
    for (int i = 0; i < n; i++) pdf[i] = world_pdf(w, vec3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
}

// noise: 3-D Perlin (signed) and fBm
void shim_noise(const float* p, int n, float octaves, float roughness, float* perlin, float* fbm) {
    for (int i = 0; i < n; i++) {
        const vec3 q(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        perlin[i] = perlin_3d(q.x, q.y, q.z);
        fbm[i] = fractal_noise(q, octaves, roughness);
    }
}
uint32_t shim_hash_uint3(uint32_t a, uint32_t b, uint32_t c) { return hash_uint3(a, b, c); }

void shim_hsv(const float* rgb, float* out, int n) {
    for (int i = 0; i < n; i++) {
        const vec3 h = rgb2hsv(vec3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]));
        const vec3 r = hsv2rgb(h);
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}

// ---------------------------------------------------------------------------
// Pins of the shared library against independent restatements (tests/test_pins_cpu.py)

// One lobe from explicit parameters.  kind 1 Specular: ax, ay, cspec0 (3),
// ior, weight.  kind 2 Transmission: reflection alpha, transmission alpha,
// base (3), cspec0 (3), ior, weight, only_refract.  kind 3 Clearcoat: alpha,
// weight.
struct AnyLobe {
    int kind;
    Specular sp;
    Transmission tr;
    Clearcoat cc;
    explicit AnyLobe(int k, const float* p) : kind(k) {
        if (k == 1) {
            sp.alpha_x = p[0]; sp.alpha_y = p[1]; sp.cspec0 = vec3(p[2], p[3], p[4]); sp.ior = p[5]; sp.weight = p[6];
        } else if (k == 2) {
            tr.reflection_alpha = p[0]; tr.transmission_alpha = p[1];
            tr.base_color = vec3(p[2], p[3], p[4]); tr.cspec0 = vec3(p[5], p[6], p[7]);
            tr.ior = p[8]; tr.weight = p[9]; tr.only_refract = p[10] != 0.0f;
        } else {
            cc.alpha = p[0]; cc.weight = p[1];
        }
    }
    vec3 evaluate(vec3 wo, vec3 wi, float& pdf) const {
        return kind == 1 ? sp.evaluate(wo, wi, pdf) : (kind == 2 ? tr.evaluate(wo, wi, pdf) : cc.evaluate(wo, wi, pdf));
    }
    BsdfSample sample(vec2 r, vec3 wo) const {
        return kind == 1 ? sp.sample(r, wo) : (kind == 2 ? tr.sample(r, wo) : cc.sample(r, wo));
    }
};

void shim_lobe_eval(int kind, const float* params, const float* wo, const float* wi, float* value, float* pdf, int n) {
    const AnyLobe L(kind, params);
    for (int i = 0; i < n; i++) {
        float pd = 0.0f;
        const vec3 v = L.evaluate(vec3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), vec3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), pd);
        value[3 * i] = v.x; value[3 * i + 1] = v.y; value[3 * i + 2] = v.z;
        pdf[i] = pd;
    }
}

void shim_lobe_sample(int kind, const float* params, const float* wo, const float* rnd, float* wi, float* weight,
                      float* pdf, int n) {
    const AnyLobe L(kind, params);
    for (int i = 0; i < n; i++) {
        const BsdfSample s = L.sample(vec2(rnd[2 * i], rnd[2 * i + 1]), vec3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]));
        wi[3 * i] = s.wi.x; wi[3 * i + 1] = s.wi.y; wi[3 * i + 2] = s.wi.z;
        weight[3 * i] = s.weight.x; weight[3 * i + 1] = s.weight.y; weight[3 * i + 2] = s.weight.z;
        pdf[i] = s.pdf;
    }
}

// BsdfPrincipled::compute.  in: base (3), metallic, roughness, specular,
// specular tint, sheen, sheen tint, clearcoat, clearcoat roughness,
// transmission, ior, anisotropic, transmission roughness, alpha, emission (3),
// emission strength (20).  out (34): lobe probabilities (4), diffuse weight
// (3), sheen weight (3), diffuse roughness, specular ax, ay, cspec0 (3), ior,
// weight, transmission reflection alpha, transmission alpha, base (3),
// cspec0 (3), ior, weight, clearcoat alpha, weight, alpha, emission (3).
void shim_principled(const float* in, float* out) {
    BsdfPrincipled<0, 0> n{};
    n.baseColor = vec4(in[0], in[1], in[2], 1.0f);
    n.metallic = in[3]; n.roughness = in[4]; n.specular = in[5]; n.specularTint = in[6];
    n.sheen = in[7]; n.sheenTint = in[8]; n.clearcoat = in[9]; n.clearcoatRoughness = in[10];
    n.transmission = in[11]; n.ior = in[12]; n.anisotropic = in[13]; n.transmissionRoughness = in[14];
    n.alpha = in[15]; n.emission = vec4(in[16], in[17], in[18], 1.0f); n.emissionStrength = in[19];
    SceneView ctx{};
    ShadingContext sc;
    n.compute(ctx, sc);
    const UberShader& b = n.bsdf;
    float* o = out;
    for (int i = 0; i < 4; i++) *o++ = b.lobe_prob[i];
    auto put3 = [&](vec3 v) { *o++ = v.x; *o++ = v.y; *o++ = v.z; };
    put3(b.diffuse.diffuse_weight); put3(b.diffuse.sheen_weight); *o++ = b.diffuse.roughness;
    *o++ = b.specular.alpha_x; *o++ = b.specular.alpha_y; put3(b.specular.cspec0); *o++ = b.specular.ior;
    *o++ = b.specular.weight;
    *o++ = b.transmission.reflection_alpha; *o++ = b.transmission.transmission_alpha;
    put3(b.transmission.base_color); put3(b.transmission.cspec0); *o++ = b.transmission.ior;
    *o++ = b.transmission.weight;
    *o++ = b.clearcoat.alpha; *o++ = b.clearcoat.weight;
    *o++ = b.alpha; put3(b.emission);
}

static mat4 mat4_of(const float* m) {  // column-major 4x4
    mat4 r;
    r.c0 = vec4(m[0], m[1], m[2], m[3]); r.c1 = vec4(m[4], m[5], m[6], m[7]);
    r.c2 = vec4(m[8], m[9], m[10], m[11]); r.c3 = vec4(m[12], m[13], m[14], m[15]);
    return r;
}
static mat3 mat3_of(const float* m) {  // column-major 3x3
    mat3 r;
    r.c0 = vec3(m[0], m[1], m[2]); r.c1 = vec3(m[3], m[4], m[5]); r.c2 = vec3(m[6], m[7], m[8]);
    return r;
}

// ShadingContext::build for face `prim` of a mesh instance.  m: point
// transform (column-major 4x4), nm: normal transform (column-major 3x3),
// bounds: min (3), size (3).  out (24): uv (3), position (3), generated (3),
// object (3), normal (3), true normal (3), tu (3), tv (3); shader index.
uint32_t shim_shading_context(const float* verts, const float* normals, const float* uvs, const uint32_t* faces,
                              const uint16_t* mats, const float* m, const float* nm, const float* bounds,
                              uint32_t prim, float b0, float b1, float* out) {
    SceneView ctx{};
    ctx.vertices = verts; ctx.normals = normals; ctx.texcoords = uvs; ctx.indices = faces; ctx.materials = mats;
    PerInstance inst{};
    inst.point_transform = mat4_of(m);
    inst.normal_transform = mat3_of(nm);
    inst.bounds_min = vec3(bounds[0], bounds[1], bounds[2]);
    inst.bounds_size = vec3(bounds[3], bounds[4], bounds[5]);
    ShadingContext sc;
    uint32_t shader = 0;
    sc.build(ctx, inst, prim, vec2(b0, b1), 1.0f, shader);
    const vec3 v[8] = {sc.uv, sc.position, sc.generated, sc.object, sc.normal, sc.true_normal, sc.tu, sc.tv};
    for (int i = 0; i < 8; i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
    return shader;
}

// ShapeLight::sample / pdf from point `origin` over an emissive mesh with the
// face CDF `cdf` (nf entries) and total area `area`.  Per sample i the RNG is
// Prng(seed_a[i], seed_b[i]).  out: direction (3), distance, pdf, weight (x),
// pdf recomputed by shape_pdf_single from the sample's shading context.
void shim_shape_light(const float* verts, const float* normals, const float* uvs, const uint32_t* faces, int nf,
                      const float* cdf, float area, const float* m, const float* nm, const float* origin,
                      const uint32_t* seed_a, const uint32_t* seed_b, int n, float* out) {
    SceneView ctx{};
    ctx.vertices = verts; ctx.normals = normals; ctx.texcoords = uvs; ctx.indices = faces;
    static const uint16_t zero_mats[1 << 16] = {0};
    ctx.materials = zero_mats;
    PerInstance inst{};
    inst.point_transform = mat4_of(m);
    inst.normal_transform = mat3_of(nm);
    inst.light_face_offset = 0;
    inst.light_face_count = (uint32_t)nf;
    ctx.instances = &inst;
    ctx.lights.light_faces = cdf;
    ShapeLight L{0u, area};
    for (int i = 0; i < n; i++) {
        ShadingContext sc;
        sc.position = vec3(origin[0], origin[1], origin[2]);
        Prng prng(seed_a[i], seed_b[i]);
        const LightSample s = sample_shape(ctx, L, sc, prng);
        float* o = out + 7 * i;
        o[0] = s.direction.x; o[1] = s.direction.y; o[2] = s.direction.z;
        o[3] = s.distance; o[4] = s.pdf; o[5] = s.weight.x;
        o[6] = shape_pdf_single(L, sc);
    }
}

// Perlin noise of dimension 1-4 (signed, unscaled) and the noise texture's
// fBm colour for dimension 4 (p: 4 per point)
void shim_perlin(int dim, const float* p, float* out, int n) {
    for (int i = 0; i < n; i++) {
        const float* q = p + 4 * i;
        switch (dim) {
        case 1: out[i] = perlin_1d(q[0]); break;
        case 2: out[i] = perlin_2d(q[0], q[1]); break;
        case 3: out[i] = perlin_3d(q[0], q[1], q[2]); break;
        default: out[i] = perlin_4d(q[0], q[1], q[2], q[3]); break;
        }
    }
}
uint32_t shim_hash_uint(uint32_t a) { return hash_uint(a); }
uint32_t shim_hash_uint2(uint32_t a, uint32_t b) { return hash_uint2(a, b); }
uint32_t shim_hash_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return hash_uint4(a, b, c, d); }
float shim_fractal4(const float* p, float octaves, float roughness) {
    return fractal_noise(vec4(p[0], p[1], p[2], p[3]), octaves, roughness);
}

// ColorRamp with 4 linear stops: pos (4), colours (4 x 4), fac (n) -> colour (n x 4)
void shim_ramp4(const float* pos, const float* col, const float* fac, float* out, int n) {
    for (int i = 0; i < n; i++) {
        ColorRamp<4> r{};
        for (int k = 0; k < 4; k++) {
            r.elements[k].position = pos[k];
            r.elements[k].color = vec4(col[4 * k], col[4 * k + 1], col[4 * k + 2], col[4 * k + 3]);
        }
        r.fac = fac[i];
        SceneView ctx{};
        ShadingContext sc;
        r.compute(ctx, sc);
        out[4 * i] = r.color.x; out[4 * i + 1] = r.color.y; out[4 * i + 2] = r.color.z; out[4 * i + 3] = r.color.w;
    }
}

}  // extern "C"

template <int Op, bool Clamp>
static float math_node(float a, float b, float c) {
    Math<Op, Clamp> m{};
    m.value = a; m.value_001 = b; m.value_002 = c;
    SceneView ctx{};
    ShadingContext sc;
    m.compute(ctx, sc);
    return m.value;
}
template <int Op>
static float math_any(int clamp, float a, float b, float c) {
    return clamp ? math_node<Op, true>(a, b, c) : math_node<Op, false>(a, b, c);
}
extern "C" {

// Math node, operation `op` in kMath order
void shim_math(int op, int clamp, const float* a, const float* b, const float* c, float* out, int n) {
    for (int i = 0; i < n; i++) {
        float r = 0.0f;
        switch (op) {
        case 0: r = math_any<0>(clamp, a[i], b[i], c[i]); break;
        case 1: r = math_any<1>(clamp, a[i], b[i], c[i]); break;
        case 2: r = math_any<2>(clamp, a[i], b[i], c[i]); break;
        case 3: r = math_any<3>(clamp, a[i], b[i], c[i]); break;
        case 4: r = math_any<4>(clamp, a[i], b[i], c[i]); break;
        case 5: r = math_any<5>(clamp, a[i], b[i], c[i]); break;
        case 6: r = math_any<6>(clamp, a[i], b[i], c[i]); break;
        case 7: r = math_any<7>(clamp, a[i], b[i], c[i]); break;
        case 8: r = math_any<8>(clamp, a[i], b[i], c[i]); break;
        case 9: r = math_any<9>(clamp, a[i], b[i], c[i]); break;
        case 10: r = math_any<10>(clamp, a[i], b[i], c[i]); break;
        default: r = math_any<11>(clamp, a[i], b[i], c[i]); break;
        }
        out[i] = r;
    }
}

}  // extern "C"

template <int Blend, bool Clamp>
static vec4 mix_node(vec4 c1, vec4 c2, float fac) {
    ColorMix<Blend, Clamp> m{};
    m.color1 = c1; m.color2 = c2; m.fac = fac;
    SceneView ctx{};
    ShadingContext sc;
    m.compute(ctx, sc);
    return m.color;
}
template <int Blend>
static vec4 mix_any(int clamp, vec4 c1, vec4 c2, float fac) {
    return clamp ? mix_node<Blend, true>(c1, c2, fac) : mix_node<Blend, false>(c1, c2, fac);
}
extern "C" {

// MixRGB node, blend type in kColorMix order: c1, c2 (n x 4), fac (n) -> n x 4
void shim_mix(int blend, int clamp, const float* c1, const float* c2, const float* fac, float* out, int n) {
    for (int i = 0; i < n; i++) {
        const vec4 a(c1[4 * i], c1[4 * i + 1], c1[4 * i + 2], c1[4 * i + 3]);
        const vec4 b(c2[4 * i], c2[4 * i + 1], c2[4 * i + 2], c2[4 * i + 3]);
        vec4 r;
        switch (blend) {
        case 0: r = mix_any<0>(clamp, a, b, fac[i]); break;
        case 1: r = mix_any<1>(clamp, a, b, fac[i]); break;
        case 2: r = mix_any<2>(clamp, a, b, fac[i]); break;
        case 3: r = mix_any<3>(clamp, a, b, fac[i]); break;
        case 4: r = mix_any<4>(clamp, a, b, fac[i]); break;
        case 5: r = mix_any<5>(clamp, a, b, fac[i]); break;
        case 6: r = mix_any<6>(clamp, a, b, fac[i]); break;
        case 7: r = mix_any<7>(clamp, a, b, fac[i]); break;
        case 8: r = mix_any<8>(clamp, a, b, fac[i]); break;
        default: r = mix_any<9>(clamp, a, b, fac[i]); break;
        }
        out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
    }
}

// MixShader of two shaders told apart by their weights (1 and 2) with
// emissions e0, e1 (3 each): out (n x 5) = chosen weight, rescaled rnd.x,
// emission (3)
void shim_mix_shader(float fac, const float* e0, const float* e1, const float* rnd, float* out, int n) {
    for (int i = 0; i < n; i++) {
        MixShader m{};
        m.fac = fac;
        m.shader.weight = 1.0f;
        m.shader.emission = vec3(e0[0], e0[1], e0[2]);
        m.shader_001.weight = 2.0f;
        m.shader_001.emission = vec3(e1[0], e1[1], e1[2]);
        SceneView ctx{};
        ShadingContext sc;
        sc.rnd = vec3(rnd[i], 0.5f, 0.5f);
        m.compute(ctx, sc);
        float* o = out + 5 * i;
        o[0] = m.shader.weight; o[1] = sc.rnd.x;
        o[2] = m.shader.emission.x; o[3] = m.shader.emission.y; o[4] = m.shader.emission.z;
    }
}

// bilinear repeat lookups of a w x h texture with `channels` floats per texel
void shim_texture_sample(const float* pixels, int w, int h, int channels, const float* uv, float* out, int n) {
    const Texture t{pixels, w, h, channels};
    for (int i = 0; i < n; i++) {
        const vec4 c = texture_sample(t, vec2(uv[2 * i], uv[2 * i + 1]));
        out[4 * i] = c.x; out[4 * i + 1] = c.y; out[4 * i + 2] = c.z; out[4 * i + 3] = c.w;
    }
}

// Nishita sky radiance (linear sRGB) toward directions d from the sky
// texture and the node's ten data values
void shim_sky_nishita(const float* pixels, int w, int h, const float* data, const float* dirs, float* out, int n) {
    const Texture t{pixels, w, h, 4};
    for (int i = 0; i < n; i++) {
        const vec3 c = sky_radiance_nishita(vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), data, t);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

// ---------------------------------------------------------------------------
// Lens camera.  A lens is passed as `ns` rows of (radius, thickness,
// aperture, n_d, V_d) and settings[] = {stop_index, relative_stop, blades,
// sensor_scale, camera_scale, focus, spectral}.

static Lens make_lens(const float* surfaces, int ns, const float* settings) {
    Lens L{};
    L.surfaces = reinterpret_cast<const LensSurface*>(surfaces);
    L.num_surfaces = ns;
    L.stop_index = (int)settings[0];
    L.relative_stop = settings[1];
    L.aperture_blades = (int)settings[2];
    L.sensor_scale = settings[3];
    L.camera_scale = settings[4];
    L.focus = settings[5];
    L.spectral = (int)settings[6];
    return L;
}

// the stop's opening at positions relative to its aperture radius
void shim_lens_aperture(const float* pos, int n, float relative_stop, int blades, uint32_t* inside) {
    Lens L{};
    L.relative_stop = relative_stop;
    L.aperture_blades = blades;
    for (int i = 0; i < n; i++) inside[i] = lens_aperture(vec2(pos[2 * i], pos[2 * i + 1]), L);
}

void shim_lens_refract(const float* d, const float* nrm, const float* eta, int n, float* out, uint32_t* ok) {
    for (int i = 0; i < n; i++) {
        vec3 v(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        ok[i] = lens_refract(v, vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), eta[i]);
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}

// n_d and V_d of one surface at wavelengths l (micrometres)
void shim_lens_ior(float ior_d, float abbe, const float* l, float* out, int n) {
    const LensSurface s{0.0f, 0.0f, 1.0f, ior_d, abbe};
    for (int i = 0; i < n; i++) out[i] = lens_ior(s, l[i]);
}

// sensor rays (origin, direction, wavelength) traced toward the object side
void shim_lens_trace(const float* surfaces, int ns, const float* settings, const float* o, const float* d,
                     const float* lambda, float* o_out, float* d_out, uint32_t* ok, int n) {
    const Lens L = make_lens(surfaces, ns, settings);
    for (int i = 0; i < n; i++) {
        vec3 a(o[3 * i], o[3 * i + 1], o[3 * i + 2]), b(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        ok[i] = lens_trace(a, b, lambda[i], L);
        o_out[3 * i] = a.x; o_out[3 * i + 1] = a.y; o_out[3 * i + 2] = a.z;
        d_out[3 * i] = b.x; d_out[3 * i + 1] = b.y; d_out[3 * i + 2] = b.z;
    }
}

// camera rays of pixel (x, y) of a W x H image with the identity camera
// transform, sample s of RNG stream (seed, index + i)
void shim_lens_camera_ray(const float* surfaces, int ns, const float* settings, uint32_t W, uint32_t H, uint32_t x,
                          uint32_t y, uint32_t seed, float* o, float* d, float* w, uint32_t* ok, int n) {
    const Lens L = make_lens(surfaces, ns, settings);
    Camera cam{};
    cam.transform.c0 = vec4(1.0f, 0.0f, 0.0f, 0.0f);
    cam.transform.c1 = vec4(0.0f, 1.0f, 0.0f, 0.0f);
    cam.transform.c2 = vec4(0.0f, 0.0f, 1.0f, 0.0f);
    cam.transform.c3 = vec4(0.0f, 0.0f, 0.0f, 1.0f);
    for (int i = 0; i < n; i++) {
        Prng prng(seed, (uint32_t)i);
        vec3 a, b, c;
        ok[i] = lens_camera_ray(cam, L, x, y, W, H, prng, a, b, c);
        o[3 * i] = a.x; o[3 * i + 1] = a.y; o[3 * i + 2] = a.z;
        d[3 * i] = b.x; d[3 * i + 1] = b.y; d[3 * i + 2] = b.z;
        w[3 * i] = c.x; w[3 * i + 1] = c.y; w[3 * i + 2] = c.z;
    }
}

// CIE 1931 colour matching functions at wavelengths (nm), their linear-sRGB
// image, and the per-channel normalisation of the spectral weight
void shim_spectral(const float* nm, int n, float* xyz, float* rgb, float* norm) {
    for (int i = 0; i < n; i++) {
        const vec3 v = wavelength_to_xyz(nm[i]);
        const vec3 c = xyz_to_rgb(v);
        xyz[3 * i] = v.x; xyz[3 * i + 1] = v.y; xyz[3 * i + 2] = v.z;
        rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
    }
    norm[0] = kCieNormRgbR;
    norm[1] = kCieNormRgbG;
    norm[2] = kCieNormRgbB;
}

void shim_xyz_to_rgb(const float* xyz, float* rgb, int n) {
    for (int i = 0; i < n; i++) {
        const vec3 c = xyz_to_rgb(vec3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
        rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
    }
}

// Two-level hits by exhaustion: every instance (vertex offset, face offset,
// face count, 16 floats of column-major object-to-world transform), the ray
// mapped by its world-to-object rows, every triangle of its shape.
void shim_two_level_brute(const float* vertices, const uint32_t* faces, const uint32_t* inst_ranges,
                          const float* transforms, int ni, const float* rays, int n, int any, float* t,
                          uint32_t* inst, uint32_t* prim, float* bary) {
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        const vec3 wo(r[0], r[1], r[2]), wd(r[4], r[5], r[6]);
        float bt = r[7];
        uint32_t bi = 0xffffffffu, bp = 0xffffffffu;
        float bu = 0.0f, bv = 0.0f, bd = 1.0f;
        bool found = false;
        for (int k = 0; k < ni && !(any && found); k++) {
            const float* m = transforms + 16 * k;
            mat4 M;
            M.c0 = vec4(m[0], m[1], m[2], m[3]);
            M.c1 = vec4(m[4], m[5], m[6], m[7]);
            M.c2 = vec4(m[8], m[9], m[10], m[11]);
            M.c3 = vec4(m[12], m[13], m[14], m[15]);
            TlasInstance ti;
            tlas_inverse(M, ti.m);
            const vec3 o = tlas_point(ti, wo), d = tlas_vector(ti, wd);
            const RayShear sh = make_shear(d);
            const uint32_t vo = inst_ranges[3 * k], fo = inst_ranges[3 * k + 1], fc = inst_ranges[3 * k + 2];
            for (uint32_t f = 0; f < fc; f++) {
                const uint32_t* ix = faces + 3 * (size_t)(fo + f);
                const float* p0 = vertices + 3 * (size_t)(vo + ix[0]);
                const float* p1 = vertices + 3 * (size_t)(vo + ix[1]);
                const float* p2 = vertices + 3 * (size_t)(vo + ix[2]);
                float tt, u, v, det;
                if (intersect_triangle_uvd(o, sh, vec3(p0[0], p0[1], p0[2]), vec3(p1[0], p1[1], p1[2]),
                                           vec3(p2[0], p2[1], p2[2]), r[3], r[7], tt, u, v, det) &&
                    hit_before(tt, (uint32_t)k, f, bt, bi, bp)) {
                    bt = tt;
                    bi = (uint32_t)k;
                    bp = f;
                    bu = u;
                    bv = v;
                    bd = det;
                    found = true;
                    if (any) break;
                }
            }
        }
        t[i] = found ? bt : -1.0f;
        inst[i] = bi;
        prim[i] = bp;
        bary[2 * i] = bu / bd;
        bary[2 * i + 1] = bv / bd;
    }
}

}  // extern "C"
