// Host build of the shared device library for CPU property tests: exposes
// RNG, elementary functions, warps, BSDF lobes, lights and the triangle test
// through plain C entry points.  Test infrastructure only.
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

}  // extern "C"
